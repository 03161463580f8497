"""In-tree build of the gfx950 library, the C++ drop-in layer and the oracle.

Outputs (git-ignored, shipped with the tree):
  iss_amd/libiss_hip.so     HIP kernels + C-ABI (include/iss_hip.h)
  iss_amd/libiSS_amd.so     C++ drop-in `iSS` class over the C-ABI
  iss_amd/iSS.e             command-line driver (the reference's iSS.e)
  oracle/liboracle.so       CPU oracle (test infrastructure)
  oracle/_ref/libissref.so  reference sources, when the reference tree exists
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "iss_amd")
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are kept out of every HIP build
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
REFERENCE_SRC = "/root/reference/src"


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if r.returncode != 0:
        print(out)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def build_hip():
    src = os.path.join(CSRC, "iss_hip.hip")
    out = os.path.join(PKG, "libiss_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out, src])
    return out


def build_host():
    """C++ drop-in layer (iSS, ParameterReader, readers) linked to the HIP lib."""
    hdir = os.path.join(CSRC, "host")
    if not os.path.exists(os.path.join(hdir, "main.cpp")):
        return None
    srcs = sorted(os.path.join(hdir, f) for f in os.listdir(hdir)
                  if f.endswith(".cpp") and f != "main.cpp")
    lib = os.path.join(PKG, "libiSS_amd.so")
    common = ["-O2", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + hdir]
    _run(["g++", *common, "-shared", "-o", lib, *srcs, "-L" + PKG, "-liss_hip",
          "-Wl,-rpath,$ORIGIN"])
    exe = os.path.join(PKG, "iSS.e")
    _run(["g++", *common, "-o", exe, os.path.join(hdir, "main.cpp"), "-L" + PKG, "-liSS_amd",
          "-liss_hip", "-Wl,-rpath,$ORIGIN"])
    return lib


def build_oracle():
    d = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", d, "liboracle.so"])
    if os.path.isdir(REFERENCE_SRC):
        _run(["make", "-s", "-C", os.path.join(d, "ref"), "REF=" + REFERENCE_SRC])


def build_all():
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host, "oracle": build_oracle}[w]()
