"""The C-ABI library loads and exports every symbol include/iss_hip.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import iss_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "iss_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(iss_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "iss_ctx_create" in syms and "iss_sample" in syms
    assert len(syms) >= 14


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(iss_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(iss_amd._SIGNATURES)


def test_hadron_layout_matches_iss_hadron():
    assert iss_amd.HADRON_DTYPE.itemsize == 40
    assert list(iss_amd.HADRON_DTYPE.names) == ["pid", "mass", "E", "px", "py", "pz", "t", "x", "y", "z"]


# ctypes mirrors of the header's structs (Python name -> C typedef)
STRUCTS = {
    "_Species": "iss_species_desc", "_Deltaf": "iss_deltaf_desc", "_Eos": "iss_eos_desc",
    "_YieldOpts": "iss_yield_opts", "_SampleOpts": "iss_sample_opts", "Stats": "iss_stats",
    "_TuneOpts": "iss_tune_opts", "TuneReport": "iss_tune_report", "_TrainOpts": "iss_train_opts",
    "TrainReport": "iss_train_report", "_SpectraOpts": "iss_spectra_opts", "_BatchOpts": "iss_batch_opts",
    "_DecayTable": "iss_decay_table",
}


def test_struct_layouts_match_the_header(tmp_path):
    """Size and every field offset of the Python bindings' structs against
    the C compiler's layout of include/iss_hip.h (a probe built with gcc)."""
    import subprocess
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "iss_hip.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        st = getattr(iss_amd, py)
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for f in st._fields_:
            lines.append(f'printf("{py} {f[0]} %zu\\n", offsetof({c}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for py in STRUCTS:
        st = getattr(iss_amd, py)
        assert got[(py, "size")] == ctypes.sizeof(st), py
        for f in st._fields_:
            assert got[(py, f[0])] == getattr(st, f[0]).offset, (py, f[0])
