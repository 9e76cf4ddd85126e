"""The allocation-free kinetic sum (kineticSumClusterNoAllocGFX950) against the entry point it
varies (kineticSumClusterGFX950): the same value, bit for bit, on two calls in a row, the second
on a smaller system; the output goes to device memory the caller owns.

The caller's buffer comes from the HIP runtime the library itself is linked against (torch
bundles a second HIP runtime, which cannot start once this one is up in the process)."""
import ctypes as C

import numpy as np
import pytest

from conftest import product

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")  # the runtime libmdgfx950.so is linked against
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _noalloc_sum(L, hip, sim, d_out):
    nan = np.full(1, np.nan)
    assert hip.hipMemcpy(d_out, nan.ctypes.data, 8, H2D) == 0
    rc = L.kineticSumClusterNoAllocGFX950(sim.param_p, sim.atom_p, d_out)
    assert rc == 0, product().status()
    L.mdgfx950Synchronize()
    got = np.zeros(1)
    assert hip.hipMemcpy(got.ctypes.data, d_out, 8, D2H) == 0
    return float(got[0])


def test_noalloc_kinetic_sum_equals_kinetic_sum():
    P = product()
    L = P.lib()
    hip = _hip()
    got = []
    d_out = None
    try:
        for shape, steps in (((8, 8, 8), 30), ((5, 6, 7), 10)):
            sim = P.CPSim(nx=shape[0], ny=shape[1], nz=shape[2], cluster_m=4, cluster_n=4,
                          precision=0, ntimes=steps)
            try:
                if d_out is None:
                    d_out = C.c_void_p()
                    assert hip.hipMalloc(C.byref(d_out), 8) == 0
                sim.run(steps)
                ref = sim.kinetic_sum()
                val = _noalloc_sum(L, hip, sim, d_out)
                print("kinetic sum %dx%dx%d: %.17g (entry point) %.17g (no-alloc)"
                      % (shape + (ref, val)), flush=True)
                assert val == ref
                assert np.isfinite(ref) and ref > 0
                got.append(val)
            finally:
                sim.close()
    finally:
        if d_out is not None and d_out.value:
            hip.hipFree(d_out)
    assert got[1] < got[0]  # the second system is the smaller one


def test_noalloc_kinetic_sum_rejects_null_output():
    P = product()
    L = P.lib()
    sim = P.CPSim(nx=4, ny=4, nz=4, cluster_m=4, cluster_n=4, precision=0, ntimes=0)
    try:
        rc = L.kineticSumClusterNoAllocGFX950(sim.param_p, sim.atom_p, None)
        assert rc == 4  # MDG_ERR_ARG
        code, msg = P.status()
        assert code == 4 and "no output pointer" in msg
    finally:
        L.mdgfx950ClearStatus()
        sim.close()
