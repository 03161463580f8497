"""The allocation-free yield entry point, iss_compute_yields_noalloc, against
iss_compute_yields through the C-ABI.

Its outputs are caller device buffers (allocated here with hipMalloc from the
HIP runtime the library itself loaded) and its scratch is a static device
array.  Per-cell rows and FP64 species totals must equal the allocating entry
point's bit for bit: on two calls in a row, the second on a smaller surface,
in FP32 and FP64, on a surface whose heavy species leave the staged Bessel
rows (the exact-K rerun, decided on the device) and on one that does not.
The host-side refusals (capacity, alignment, state, diffusion) are checked
with no kernel launched.
"""
import ctypes as C

import numpy as np
import pytest

import iss_amd
from iss_amd import synth, tables

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_STATE, ERR_UNSUPPORTED, ERR_RANGE = -1, -3, -4, -6


def _hip():
    """The HIP runtime libiss_hip.so was linked against (same instance)."""
    iss_amd.load_library()
    with open("/proc/self/maps") as f:
        path = next(l.split()[-1] for l in f if "libamdhip64" in l)
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


class DevBuf:
    """A device buffer of n elements of dtype, filled with 0xff bytes so that
    every entry the entry point claims to write is seen to be written."""

    def __init__(self, hip, n, dtype):
        self.hip, self.n, self.dtype = hip, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), self.nbytes) == 0
        self.p = p.value
        assert hip.hipMemset(C.c_void_p(self.p), 0xFF, self.nbytes) == 0

    def get(self):
        out = np.empty(self.n, dtype=self.dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.p), self.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(C.c_void_p(self.p))


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


def _ctx(sp, diffusion=0):
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1), diffusion=diffusion,
                   qmu_table=tables.read_qmu_table() if diffusion else None)
    return ctx


def _surface(sp, n, seed, cold):
    cells, _ = synth.viscous_lrf(n, sp, seed=seed, muB_max=0.3)
    if cold:
        # T = 0.03 GeV: 20 m / T - 10 > 1024 for m > 1.55 GeV, beyond the
        # staged rows, so the exact-K pass runs
        cells[::97, 13] = 0.03
    return cells


def _compare(ctx, hip, prec, tot_d, pc_d, cap):
    print(f"noalloc: precision {prec}, cells {ctx.num_cells()}, capacity {cap}", flush=True)
    tot_ref, pc_ref = ctx.compute_yields(precision=prec, per_cell=True)
    rc = ctx.compute_yields_noalloc(tot_d.p, pc_d.p, cap, precision=prec)
    assert rc == 0, ctx.lib.iss_last_error(ctx.h)
    nc, ns, st = ctx.num_cells(), ctx.n_species, ctx.row_stride(prec)
    rows = pc_d.get()[: nc * st].reshape(nc, st)
    assert np.array_equal(rows[:, :ns].view(np.uint8), np.ascontiguousarray(pc_ref).view(np.uint8))
    assert np.all(rows[:, ns:] == 0)
    tot = tot_d.get()
    assert np.array_equal(tot.view(np.uint64), tot_ref.view(np.uint64))
    assert np.all(np.isfinite(tot)) and tot.sum() > 0


@pytest.mark.parametrize("cold", [False, True], ids=["staged", "exact_rerun"])
@pytest.mark.parametrize("prec", [32, 64])
def test_noalloc_equals_compute_yields(species, prec, cold):
    hip = _hip()
    ctx = _ctx(species)
    ns = len(species["mass"])
    st = ctx.row_stride(prec)
    dt = np.float32 if prec == 32 else np.float64
    n1, n2 = 30000, 7000
    tot_d = DevBuf(hip, ns, np.float64)
    pc_d = DevBuf(hip, n1 * st, dt)
    try:
        ctx.set_surface_lrf(_surface(species, n1, 5, cold))
        _compare(ctx, hip, prec, tot_d, pc_d, n1 * st)
        # a second call, smaller surface, the same buffers (the scratch keeps
        # the first call's contents)
        ctx.set_surface_lrf(_surface(species, n2, 6, cold))
        _compare(ctx, hip, prec, tot_d, pc_d, n1 * st)
    finally:
        tot_d.free()
        pc_d.free()
        ctx.close()


def test_noalloc_refusals(species):
    hip = _hip()
    ns = len(species["mass"])
    st = (ns + 15) // 16 * 16
    tot_d = DevBuf(hip, ns, np.float64)
    pc_d = DevBuf(hip, 2000 * st, np.float32)
    ctx = _ctx(species)
    ctxd = _ctx(species, diffusion=1)
    try:
        ctx.set_surface_lrf(_surface(species, 2000, 7, False))
        # no Bessel tables yet on this context
        assert ctx.compute_yields_noalloc(tot_d.p, pc_d.p, 2000 * st) == ERR_STATE
        ctx.compute_yields(precision=32)
        assert ctx.compute_yields_noalloc(tot_d.p, pc_d.p, 2000 * st - 1) == ERR_RANGE
        assert ctx.compute_yields_noalloc(tot_d.p, pc_d.p + 4, 2000 * st) == ERR_ARG
        assert ctx.compute_yields_noalloc(tot_d.p, 0, 2000 * st) == ERR_ARG
        assert ctx.compute_yields_noalloc(tot_d.p, pc_d.p, 2000 * st, precision=16) == ERR_ARG
        ctxd.set_surface_lrf(_surface(species, 2000, 7, False))
        ctxd.compute_yields(precision=32)
        assert ctxd.compute_yields_noalloc(tot_d.p, pc_d.p, 2000 * st) == ERR_UNSUPPORTED
        # nothing was written by the refused calls
        assert np.all(pc_d.get().view(np.uint32) == 0xFFFFFFFF)
        # and a valid call still works after them
        assert ctx.compute_yields_noalloc(tot_d.p, pc_d.p, 2000 * st) == 0
    finally:
        tot_d.free()
        pc_d.free()
        ctx.close()
        ctxd.close()
