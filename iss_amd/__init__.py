"""iss_amd: a gfx950-native Cooper-Frye sampler with iSS's FSSW behaviour.

This module is a ctypes binding of the C-ABI declared in include/iss_hip.h
(the library ``libiss_hip.so`` built next to this file).  It exists for the
tests and the benchmark; the drop-in for iSS users is the C++ ``iSS`` class in
``iss_amd/csrc/host``.  There is no CPU fallback: every call goes to the HIP
library and fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libiss_hip.so")
TABLE_DIR = os.path.join(_HERE, "tables")

HADRON_DTYPE = np.dtype(
    [("pid", "<i4"), ("mass", "<f4"), ("E", "<f4"), ("px", "<f4"), ("py", "<f4"),
     ("pz", "<f4"), ("t", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4")])

# field order of a rest-frame cell record (FO_surf_LRF, src/data_struct.h:66-76)
LRF_FIELDS = ["tau", "xpt", "ypt", "eta", "da0", "da1", "da2", "da3", "u0", "u1", "u2", "u3",
              "Edec", "Tdec", "Pdec", "Bn", "muB", "muS", "muQ", "bulkPi",
              "pixx", "pixy", "pixz", "piyy", "piyz", "qx", "qy", "qz"]
# field order of a lab-frame cell record (FO_surf, src/data_struct.h:53-63)
LAB_FIELDS = ["tau", "xpt", "ypt", "eta", "da0", "da1", "da2", "da3", "u0", "u1", "u2", "u3",
              "Edec", "Tdec", "Pdec", "Bn", "muB", "muS", "muQ",
              "pi00", "pi01", "pi02", "pi03", "pi11", "pi12", "pi13", "pi22", "pi23", "pi33",
              "bulkPi", "qmu0", "qmu1", "qmu2", "qmu3"]


class ISSError(RuntimeError):
    pass


class _Species(C.Structure):
    _fields_ = [("n", C.c_int32), ("pid", C.c_void_p), ("mass", C.c_void_p),
                ("degeneracy", C.c_void_p), ("baryon", C.c_void_p), ("strange", C.c_void_p),
                ("charge", C.c_void_p), ("sign", C.c_void_p)]


class _Deltaf(C.Structure):
    _fields_ = [("include_shear", C.c_int32), ("include_bulk", C.c_int32),
                ("bulk_kind", C.c_int32), ("include_diffusion", C.c_int32),
                ("k11_nT", C.c_int32), ("k11_nmu", C.c_int32),
                ("k11_T0", C.c_double), ("k11_mu0", C.c_double),
                ("k11_dT", C.c_double), ("k11_dmu", C.c_double),
                ("k11_c0", C.c_void_p), ("k11_c1", C.c_void_p), ("k11_c2", C.c_void_p),
                ("neos_ncol", C.c_int32), ("neos", C.c_void_p),
                ("fermion_cdf2_quirk", C.c_int32), ("qmu_table", C.c_void_p)]


class _Eos(C.Structure):
    _fields_ = [("iEOS", C.c_int32), ("n_rows", C.c_int64), ("rows", C.c_void_p)]


class _YieldOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("hydro_mode", C.c_int32),
                ("y_LB", C.c_double), ("y_RB", C.c_double)]


class _SampleOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("event0", C.c_int64), ("n_events", C.c_int32),
                ("proposal", C.c_int32), ("sampling_model", C.c_int32),
                ("hydro_mode", C.c_int32), ("y_LB", C.c_double), ("y_RB", C.c_double),
                ("max_trials", C.c_int32), ("precision", C.c_int32),
                ("local_charge_conservation", C.c_int32), ("sampling_para1", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("n_hadrons", C.c_int64), ("n_trials", C.c_int64), ("n_inner", C.c_int64),
                ("n_failed", C.c_int64), ("n_range_errors", C.c_int64),
                ("ms_yield", C.c_double), ("ms_sample", C.c_double),
                ("ms_sample_kernel", C.c_double), ("ms_yield_kernel", C.c_double),
                ("n_fallback", C.c_int64), ("n_violations", C.c_int64),
                ("ms_setup_kernel", C.c_double), ("ms_trials_kernel", C.c_double),
                ("ms_decay", C.c_double), ("n_decay_generations", C.c_int64),
                ("n_pairs", C.c_int64), ("n_routed", C.c_int64),
                ("ms_learned_trials_kernel", C.c_double), ("ms_routed_setup_kernel", C.c_double),
                ("ms_routed_trials_kernel", C.c_double), ("n_routed_trials", C.c_int64),
                ("n_routed_inner", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _TrainOpts(C.Structure):
    _fields_ = [("batch", C.c_int64), ("steps", C.c_int32), ("lr", C.c_float), ("seed", C.c_uint64)]


class TrainReport(C.Structure):
    _fields_ = [("n_samples", C.c_int64), ("steps", C.c_int32), ("loss_first", C.c_double),
                ("loss_last", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _TuneOpts(C.Structure):
    _fields_ = [("proposal", C.c_int32), ("n_events", C.c_int32), ("seed", C.c_uint64), ("repeats", C.c_int32)]


class TuneReport(C.Structure):
    _fields_ = [("refill_min", C.c_int32), ("blocks_per_cu", C.c_int32), ("ms_best", C.c_double),
                ("ms_default", C.c_double), ("n_tried", C.c_int32), ("learned_min_hadrons", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _DecayTable(C.Structure):
    _fields_ = [("n", C.c_int32), ("monval", C.c_void_p), ("mass", C.c_void_p), ("width", C.c_void_p),
                ("stable", C.c_void_p), ("first_channel", C.c_void_p), ("n_channels", C.c_void_p),
                ("n_channels_total", C.c_int32), ("npart", C.c_void_p), ("branching_ratio", C.c_void_p),
                ("daughters", C.c_void_p)]


class _BatchOpts(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("hydro_mode", C.c_int32), ("y_LB", C.c_double), ("y_RB", C.c_double),
                ("events_per_round", C.c_int32), ("max_trials", C.c_int32), ("precision", C.c_int32),
                ("use_graph", C.c_int32), ("proposal", C.c_int32)]


class _SpectraOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("mode", C.c_int32), ("all_cells", C.c_int32),
                ("use_pos_dN_only", C.c_int32), ("restrict_deltaf", C.c_int32), ("deltaf_max_ratio", C.c_double)]


# network layout (parameters, FP32, row-major [out][in])
N_IN, N_HID, N_OUT = 16, 64, 8
N_BANDS = 9
BAND_DIAG = 47


# every symbol include/iss_hip.h declares, with its ctypes signature
_SIGNATURES = {
    "iss_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "iss_ctx_destroy": (C.c_int, [C.c_void_p]),
    "iss_last_error": (C.c_char_p, [C.c_void_p]),
    "iss_set_species": (C.c_int, [C.c_void_p, C.POINTER(_Species)]),
    "iss_set_deltaf": (C.c_int, [C.c_void_p, C.POINTER(_Deltaf)]),
    "iss_set_surface_lab": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(_Eos)]),
    "iss_set_surface_lrf": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "iss_num_cells": (C.c_int64, [C.c_void_p]),
    "iss_get_surface_lrf": (C.c_int, [C.c_void_p, C.c_void_p]),
    "iss_compute_yields": (C.c_int, [C.c_void_p, C.POINTER(_YieldOpts), C.c_void_p, C.c_void_p]),
    "iss_compute_yields_noalloc": (C.c_int, [C.c_void_p, C.POINTER(_YieldOpts), C.c_void_p, C.c_void_p,
                                             C.c_int64]),
    "iss_sample": (C.c_int, [C.c_void_p, C.POINTER(_SampleOpts), C.POINTER(C.c_int64)]),
    "iss_get_event_counts": (C.c_int, [C.c_void_p, C.c_void_p]),
    "iss_get_hadrons": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "iss_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "iss_autotune": (C.c_int, [C.c_void_p, C.POINTER(_TuneOpts), C.POINTER(TuneReport)]),
    "iss_mlp_num_params": (C.c_int, []),
    "iss_mlp_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "iss_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "iss_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "iss_set_learned_routing": (C.c_int, [C.c_void_p, C.c_int32]),
    "iss_get_learned_routing": (C.c_int, [C.c_void_p]),
    "iss_mlp_collect": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "iss_mlp_train": (C.c_int, [C.c_void_p, C.POINTER(_TrainOpts), C.POINTER(TrainReport)]),
    "iss_mlp_eval": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p]),
    "iss_band_diag": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
    "iss_set_decay_table": (C.c_int, [C.c_void_p, C.POINTER(_DecayTable)]),
    "iss_decay": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int64)]),
    "iss_set_hadrons": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "iss_set_surfaces_lrf": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "iss_num_surfaces": (C.c_int, [C.c_void_p]),
    "iss_sample_batch": (C.c_int, [C.c_void_p, C.POINTER(_BatchOpts), C.c_void_p, C.POINTER(C.c_int64)]),
    "iss_batch_rounds": (C.c_int, [C.c_void_p]),
    "iss_batch_events_per_round": (C.c_int, [C.c_void_p]),
    "iss_batch_ms": (C.c_double, [C.c_void_p]),
    "iss_get_batch_layout": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "iss_autotune_batch": (C.c_int, [C.c_void_p, C.POINTER(_BatchOpts), C.c_void_p, C.POINTER(TuneReport)]),
    "iss_spectra": (C.c_int, [C.c_void_p, C.POINTER(_SpectraOpts), C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                              C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "iss_spectra_ms": (C.c_double, [C.c_void_p]),
    "iss_flows": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                            C.c_double, C.c_int32, C.c_void_p, C.c_void_p]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libiss_hip.so; raise if it is not built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ISSError(f"HIP library {path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """One device context of the sampler (wraps iss_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.iss_ctx_create(device, C.byref(h))
        if rc != 0:
            raise ISSError(f"iss_ctx_create failed with status {rc}")
        self.h = h
        self.n_species = 0
        self._keep = []

    def close(self):
        if self.h:
            self.lib.iss_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.iss_last_error(self.h)
            raise ISSError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    # ---------------------------------------------------------------- tables
    def set_species(self, sp: dict):
        arrs = {k: np.ascontiguousarray(sp[k], dtype=np.float64 if k == "mass" else np.int32)
                for k in ("pid", "mass", "g", "B", "S", "Q", "sign")}
        d = _Species(len(arrs["pid"]), _ptr(arrs["pid"]), _ptr(arrs["mass"]), _ptr(arrs["g"]),
                     _ptr(arrs["B"]), _ptr(arrs["S"]), _ptr(arrs["Q"]), _ptr(arrs["sign"]))
        self._check(self.lib.iss_set_species(self.h, C.byref(d)), "iss_set_species")
        self.n_species = len(arrs["pid"])
        self.species = arrs

    def set_deltaf(self, shear=0, bulk=0, bulk_kind=1, k11=None, neos=None,
                   fermion_cdf2_quirk=1, diffusion=0, qmu_table=None):
        d = _Deltaf()
        d.include_shear, d.include_bulk, d.bulk_kind = int(shear), int(bulk), int(bulk_kind)
        d.include_diffusion = int(diffusion)
        d.fermion_cdf2_quirk = int(fermion_cdf2_quirk)
        keep = []
        if k11 is not None:
            d.k11_nT, d.k11_nmu = k11["nT"], k11["nmu"]
            d.k11_T0, d.k11_mu0, d.k11_dT, d.k11_dmu = k11["T0"], k11["mu0"], k11["dT"], k11["dmu"]
            for nm in ("c0", "c1", "c2"):
                a = np.ascontiguousarray(k11[nm], dtype=np.float64)
                keep.append(a)
                setattr(d, "k11_" + nm, _ptr(a))
        if neos is not None:
            a = np.ascontiguousarray(neos, dtype=np.float64)
            keep.append(a)
            d.neos_ncol = a.shape[1]
            d.neos = _ptr(a)
        if qmu_table is not None:
            a = np.ascontiguousarray(qmu_table, dtype=np.float64).reshape(150, 100)
            keep.append(a)
            d.qmu_table = _ptr(a)
        self._check(self.lib.iss_set_deltaf(self.h, C.byref(d)), "iss_set_deltaf")

    # -------------------------------------------------------------- surfaces
    def set_surface_lrf(self, cells28: np.ndarray):
        a = np.ascontiguousarray(cells28, dtype=np.float32).reshape(-1, 28)
        self._check(self.lib.iss_set_surface_lrf(self.h, a.shape[0], _ptr(a)), "iss_set_surface_lrf")

    def set_surface_lab(self, cells34: np.ndarray, eos: dict | None = None):
        a = np.ascontiguousarray(cells34, dtype=np.float32).reshape(-1, 34)
        e = None
        if eos is not None:
            rows = np.ascontiguousarray(eos["rows"], dtype=np.float64)
            self._keep = [rows]
            e = C.byref(_Eos(int(eos["iEOS"]), rows.shape[0], _ptr(rows)))
        self._check(self.lib.iss_set_surface_lab(self.h, a.shape[0], _ptr(a), e), "iss_set_surface_lab")

    def num_cells(self) -> int:
        return int(self.lib.iss_num_cells(self.h))

    def get_surface_lrf(self) -> np.ndarray:
        out = np.zeros((self.num_cells(), 28), dtype=np.float32)
        self._check(self.lib.iss_get_surface_lrf(self.h, _ptr(out)), "iss_get_surface_lrf")
        return out

    # ---------------------------------------------------------------- yields
    def compute_yields(self, precision=32, per_cell=False, hydro_mode=2, y_LB=-5.0, y_RB=5.0):
        tot = np.zeros(self.n_species, dtype=np.float64)
        pc = None
        if per_cell:
            pc = np.zeros((self.num_cells(), self.n_species),
                          dtype=np.float64 if precision == 64 else np.float32)
        o = _YieldOpts(precision, hydro_mode, y_LB, y_RB)
        self._check(self.lib.iss_compute_yields(self.h, C.byref(o), _ptr(tot), _ptr(pc)),
                    "iss_compute_yields")
        return (tot, pc) if per_cell else tot

    def compute_yields_noalloc(self, species_totals_dev: int, per_cell_dev: int, capacity: int, precision=32,
                               hydro_mode=2, y_LB=-5.0, y_RB=5.0) -> int:
        """iss_compute_yields_noalloc into caller DEVICE buffers (addresses as
        integers): species totals [n_species] FP64, per-cell rows
        [n_cells][row_stride(precision)].  Returns the status code."""
        o = _YieldOpts(precision, hydro_mode, y_LB, y_RB)
        return self.lib.iss_compute_yields_noalloc(self.h, C.byref(o), C.c_void_p(species_totals_dev),
                                                   C.c_void_p(per_cell_dev), capacity)

    def row_stride(self, precision=32) -> int:
        """Row length of the allocation-free yields: the species count rounded
        up to 16 (FP32) or 8 (FP64)."""
        q = 16 if precision == 32 else 8
        return (self.n_species + q - 1) // q * q

    # -------------------------------------------------------------- sampling
    def sample(self, n_events, seed=1, event0=0, hydro_mode=2, y_LB=-5.0, y_RB=5.0,
               max_trials=5000, precision=32, proposal=0, sampling_model=30,
               local_charge_conservation=0, sampling_para1=0.16) -> int:
        """One sampling pass of n_events events (FSSW's per-event loop,
        src/FSSW.cpp:884-1070).  sampling_model is dN_dy_sampling_model: 30
        Poisson per cell, 1 / 10 / 20 per (event, species) from the species
        yield (src/FSSW.cpp:250-309), with sampling_para1 the negative
        binomial parameter of 10 and 20."""
        o = _SampleOpts(seed, event0, n_events, proposal, sampling_model, hydro_mode, y_LB, y_RB,
                        max_trials, precision, int(local_charge_conservation), float(sampling_para1))
        n = C.c_int64()
        self._check(self.lib.iss_sample(self.h, C.byref(o), C.byref(n)), "iss_sample")
        self.n_events = n_events
        return int(n.value)

    def event_counts(self) -> np.ndarray:
        out = np.zeros(self.n_events, dtype=np.int64)
        self._check(self.lib.iss_get_event_counts(self.h, _ptr(out)), "iss_get_event_counts")
        return out

    def hadrons(self) -> np.ndarray:
        n = self.stats().n_hadrons
        out = np.zeros(n, dtype=HADRON_DTYPE)
        self._check(self.lib.iss_get_hadrons(self.h, _ptr(out), n), "iss_get_hadrons")
        return out

    def stats(self) -> Stats:
        s = Stats()
        self._check(self.lib.iss_get_stats(self.h, C.byref(s)), "iss_get_stats")
        return s

    def autotune(self, proposal=0, n_events=200, seed=99, repeats=2) -> dict:
        o = _TuneOpts(proposal, n_events, seed, repeats)
        r = TuneReport()
        self._check(self.lib.iss_autotune(self.h, C.byref(o), C.byref(r)), "iss_autotune")
        return r.as_dict()

    # ------------------------------------------------------ learned proposal
    def num_params(self) -> int:
        return int(self.lib.iss_mlp_num_params())

    def mlp_init(self, seed: int = 1):
        self._check(self.lib.iss_mlp_init(self.h, seed), "iss_mlp_init")

    def mlp_set_params(self, params: np.ndarray):
        a = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._check(self.lib.iss_mlp_set_params(self.h, _ptr(a), a.size), "iss_mlp_set_params")

    def mlp_get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        self._check(self.lib.iss_mlp_get_params(self.h, _ptr(out), out.size), "iss_mlp_get_params")
        return out

    def set_learned_routing(self, min_hadrons: int = 4):
        """(cell, species) pairs with fewer hadrons than min_hadrons (1..64)
        in a learned pass take the plain sampler; 65 routes every pair (the
        pass runs as the plain one)."""
        self._check(self.lib.iss_set_learned_routing(self.h, min_hadrons), "iss_set_learned_routing")

    def learned_routing(self) -> int:
        return int(self.lib.iss_get_learned_routing(self.h))

    def mlp_collect(self, max_samples: int = 10_000_000) -> int:
        n = C.c_int64()
        self._check(self.lib.iss_mlp_collect(self.h, max_samples, C.byref(n)), "iss_mlp_collect")
        return int(n.value)

    def mlp_train(self, batch=8192, steps=500, lr=2e-3, seed=7) -> dict:
        o = _TrainOpts(batch, steps, lr, seed)
        r = TrainReport()
        self._check(self.lib.iss_mlp_train(self.h, C.byref(o), C.byref(r)), "iss_mlp_train")
        return r.as_dict()

    def mlp_eval(self, cell, species):
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        species = np.ascontiguousarray(species, dtype=np.int32)
        n = cell.size
        loss = np.zeros(n, dtype=np.float32)
        dz = np.zeros((n, N_OUT), dtype=np.float32)
        x = np.zeros((n, N_IN), dtype=np.float32)
        grad = np.zeros(self.num_params(), dtype=np.float32)
        self._check(self.lib.iss_mlp_eval(self.h, n, _ptr(cell), _ptr(species), _ptr(loss), _ptr(dz),
                                          _ptr(x), _ptr(grad)), "iss_mlp_eval")
        return {"loss": loss, "dz": dz, "x": x, "grad": grad}

    def band_diag(self, cell, species, rho):
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        species = np.ascontiguousarray(species, dtype=np.int32)
        rho = np.ascontiguousarray(rho, dtype=np.float32).reshape(cell.size, N_OUT)
        tab = np.zeros((cell.size, BAND_DIAG), dtype=np.float32)
        par = np.zeros((cell.size, 24), dtype=np.float32)
        self._check(self.lib.iss_band_diag(self.h, cell.size, _ptr(cell), _ptr(species), _ptr(rho),
                                           _ptr(tab), _ptr(par)), "iss_band_diag")
        return tab, par

    # ------------------------------------------------------ resonance decays
    def set_decay_table(self, t: dict):
        """t: tables.decay_table(...) (entries with their channels)"""
        arrs = {"monval": np.ascontiguousarray(t["monval"], dtype=np.int32),
                "mass": np.ascontiguousarray(t["mass"], dtype=np.float64),
                "width": np.ascontiguousarray(t["width"], dtype=np.float64),
                "stable": np.ascontiguousarray(t["stable"], dtype=np.int32),
                "first_channel": np.ascontiguousarray(t["first_channel"], dtype=np.int32),
                "n_channels": np.ascontiguousarray(t["n_channels"], dtype=np.int32),
                "npart": np.ascontiguousarray(t["npart"], dtype=np.int32),
                "branching_ratio": np.ascontiguousarray(t["branching_ratio"], dtype=np.float64),
                "daughters": np.ascontiguousarray(t["daughters"], dtype=np.int32).reshape(-1)}
        d = _DecayTable(arrs["monval"].size, _ptr(arrs["monval"]), _ptr(arrs["mass"]), _ptr(arrs["width"]),
                        _ptr(arrs["stable"]), _ptr(arrs["first_channel"]), _ptr(arrs["n_channels"]),
                        arrs["npart"].size, _ptr(arrs["npart"]), _ptr(arrs["branching_ratio"]),
                        _ptr(arrs["daughters"]))
        self._check(self.lib.iss_set_decay_table(self.h, C.byref(d)), "iss_set_decay_table")

    def set_hadrons(self, counts, hadrons):
        c = np.ascontiguousarray(counts, dtype=np.int64)
        h = np.ascontiguousarray(hadrons, dtype=HADRON_DTYPE)
        self._check(self.lib.iss_set_hadrons(self.h, c.size, _ptr(c), _ptr(h)), "iss_set_hadrons")
        self.n_events = c.size

    def decay(self, seed=1):
        """replace the last sampled hadron list by its stable descendants"""
        n = C.c_int64()
        self._check(self.lib.iss_decay(self.h, seed, C.byref(n)), "iss_decay")
        return n.value

    # ------------------------------------------------------ batched surfaces
    def set_surfaces_lrf(self, surfaces):
        """surfaces: list of [n_k, 28] rest-frame cell arrays, held as one record"""
        n = np.array([np.asarray(c).reshape(-1, 28).shape[0] for c in surfaces], dtype=np.int64)
        a = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.float32).reshape(-1, 28) for c in surfaces]))
        self._check(self.lib.iss_set_surfaces_lrf(self.h, n.size, _ptr(n), _ptr(a)), "iss_set_surfaces_lrf")
        self.n_surfaces = n.size

    def _batch_opts(self, seed, hydro_mode, y_LB, y_RB, events_per_round, max_trials, precision, use_graph,
                    proposal=0):
        return _BatchOpts(seed, hydro_mode, y_LB, y_RB, events_per_round, max_trials, precision, use_graph,
                          proposal)

    def sample_batch(self, n_events, seed=1, hydro_mode=2, y_LB=-5.0, y_RB=5.0, events_per_round=0,
                     max_trials=5000, precision=32, use_graph=1, proposal=0):
        ev = np.ascontiguousarray(n_events, dtype=np.int32)
        o = self._batch_opts(seed, hydro_mode, y_LB, y_RB, events_per_round, max_trials, precision, use_graph,
                             proposal)
        n = C.c_int64()
        self._check(self.lib.iss_sample_batch(self.h, C.byref(o), _ptr(ev), C.byref(n)), "iss_sample_batch")
        self._batch_events = ev.copy()
        return n.value

    def batch_ms(self):
        return self.lib.iss_batch_ms(self.h)

    def batch_layout(self):
        """(segments [rounds, S, 2], event_sizes [rounds, S, R]) of the last batch"""
        r = self.lib.iss_batch_rounds(self.h)
        R = self.lib.iss_batch_events_per_round(self.h)
        S = self._batch_events.size
        seg = np.zeros((max(r, 0), S, 2), dtype=np.int64)
        ev = np.zeros((max(r, 0), S, max(R, 1)), dtype=np.int64)
        if r > 0:
            self._check(self.lib.iss_get_batch_layout(self.h, _ptr(seg), _ptr(ev)), "iss_get_batch_layout")
        return seg, ev

    def batch_surface(self, k, hadrons=None, layout=None):
        """hadrons of surface k in event order and its event sizes"""
        h = self.hadrons() if hadrons is None else hadrons
        seg, ev = self.batch_layout() if layout is None else layout
        parts = [h[o:o + c] for o, c in seg[:, k]]
        sizes = ev[:, k, :].ravel()[: int(self._batch_events[k])]
        return (np.concatenate(parts) if parts else h[:0]), sizes

    def autotune_batch(self, n_events, seed=1, hydro_mode=2, precision=32, proposal=0):
        ev = np.ascontiguousarray(n_events, dtype=np.int32)
        o = self._batch_opts(seed, hydro_mode, -5.0, 5.0, 0, 5000, precision, 1, proposal)
        r = TuneReport()
        self._check(self.lib.iss_autotune_batch(self.h, C.byref(o), _ptr(ev), C.byref(r)), "iss_autotune_batch")
        return r.as_dict()

    # ------------------------------------------------------- smooth spectra
    def spectra(self, species, pT, phi, eta, eta_weight=None, *, mode=0, precision=64, all_cells=1,
                use_pos_dN_only=0, restrict_deltaf=1, deltaf_max_ratio=1.0):
        """dN/(pT dpT dphi dy) [GeV^-2] of the given species (indices into the
        species table) on pT x phi; mode 0 integrates each cell over the eta_s
        table, mode 1 evaluates at the rapidities ``eta``.  Returns
        [n_species, n_pT, n_phi] (mode 0) or [n_species, n_y, n_pT, n_phi]."""
        species = np.ascontiguousarray(np.atleast_1d(species), dtype=np.int32)
        pT = np.ascontiguousarray(pT, dtype=np.float64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        eta = np.ascontiguousarray(np.atleast_1d(eta), dtype=np.float64)
        w = None if eta_weight is None else np.ascontiguousarray(eta_weight, dtype=np.float64)
        o = _SpectraOpts(precision, mode, all_cells, use_pos_dN_only, restrict_deltaf, deltaf_max_ratio)
        shape = (species.size, pT.size, phi.size) if mode == 0 else (species.size, eta.size, pT.size, phi.size)
        out = np.zeros(shape, dtype=np.float64)
        self._check(self.lib.iss_spectra(self.h, C.byref(o), _ptr(species), species.size, _ptr(pT), pT.size,
                                         _ptr(phi), phi.size, _ptr(eta), _ptr(w), eta.size, _ptr(out)),
                    "iss_spectra")
        return out

    def spectra_ms(self):
        return self.lib.iss_spectra_ms(self.h)


def flows(dN, pT, pT_weight, phi, phi_weight, mass, to_order=9):
    """calculate_flows of one spectrum dN[n_pT, n_phi]: (vn_diff [n_pT, 3 + 3 to_order],
    vn_inte [to_order + 1, 6]) in the reference's column order."""
    lib = load_library()
    dN = np.ascontiguousarray(dN, dtype=np.float64)
    pT = np.ascontiguousarray(pT, dtype=np.float64)
    pw = np.ascontiguousarray(pT_weight, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    fw = np.ascontiguousarray(phi_weight, dtype=np.float64)
    if dN.shape != (pT.size, phi.size):
        raise ISSError("flows: dN must be [n_pT, n_phi]")
    diff = np.zeros((pT.size, 3 + 3 * to_order))
    inte = np.zeros((to_order + 1, 6))
    rc = lib.iss_flows(_ptr(dN), _ptr(pT), _ptr(pw), pT.size, _ptr(phi), _ptr(fw), phi.size, float(mass),
                       to_order, _ptr(diff), _ptr(inte))
    if rc != 0:
        raise ISSError(f"iss_flows: status {rc}")
    return diff, inte
