"""pecs_solve_dev_noalloc through the C-ABI against pecs_solve_dev.

The no-allocation variant runs the same kernels in the same order on a static
device scratch array instead of the context's buffers, so its solutions must
equal pecs_solve_dev's bit for bit, for every system, on two contexts in a
row (the second one smaller: the array keeps the first call's contents).  It
must leave the device's free memory unchanged, and reject batched contexts
and systems whose scratch exceeds the limit with an error.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _lib():
    import pecs_amd
    L = pecs_amd.lib()
    for name in ("pecs_solve_dev", "pecs_solve_dev_noalloc"):
        getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        getattr(L, name).restype = ctypes.c_int
    return L


class _Hip:
    """Device buffers through the HIP runtime the library itself loaded
    (no second runtime in the process)."""

    def __init__(self):
        import pecs_amd
        pecs_amd.lib()
        self.h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipMemGetInfo"):
            getattr(self.h, name).restype = ctypes.c_int
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def alloc(self, n):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(8 * max(n, 1))) == 0
        self.bufs.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = self.alloc(a.size)
        assert self.h.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0
        return d

    def get(self, d, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free_bytes(self):
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        assert self.h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
        return free.value

    def release(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def test_noalloc_matches_solve_dev(seed):
    import pecs_amd
    L = _lib()
    hip = _Hip()
    rng = np.random.default_rng(seed)
    for r, k in ((5, 2), (3, 1)):  # the second context is smaller
        p = pecs_amd.workload_params(r, k, 5e-4)
        p.update(phi_bi=0.4, phi_app=0.1)
        _say(f"[gpu] no-alloc solve vs pecs_solve_dev: refinement {r}, k={k}")
        g = pecs_amd.SolarCellGPU(p)
        for which, n in ((pecs_amd.ELECTRONS, g.n_semi), (pecs_amd.HOLES, g.n_semi),
                         (pecs_amd.REDUCTANTS, g.n_elec), (pecs_amd.OXIDANTS, g.n_elec),
                         (pecs_amd.POISSON, g.n_poisson)):
            bh = rng.standard_normal(n)
            b = hip.put(bh)
            x_ref = hip.put(np.zeros(n))
            x_new = hip.put(np.full(n, np.nan))
            hip.sync()
            assert L.pecs_solve_dev(g.ctx, which, b, x_ref) == 0
            hip.sync()
            free0 = hip.free_bytes()
            rc = L.pecs_solve_dev_noalloc(g.ctx, which, b, x_new)
            assert rc == 0, L.pecs_last_error(g.ctx)
            hip.sync()
            assert hip.free_bytes() == free0
            xr, xn = hip.get(x_ref, n), hip.get(x_new, n)
            assert np.array_equal(xn, xr), (r, which, float(np.abs(xn - xr).max()))
            # and the host-pointer solve (pecs_solve) agrees with both
            assert np.array_equal(g.solve(which, bh), xr)
        g.close()
    hip.release()


def test_noalloc_rejects_batches_and_oversize():
    import pecs_amd
    L = _lib()
    hip = _Hip()
    p = pecs_amd.workload_params(3, 2, 5e-4)
    _say("[gpu] no-alloc solve: batched context and size limit rejected")
    gb = pecs_amd.SolarCellGPU([p, dict(p, phi_app=0.2)])
    b = hip.put(np.zeros(gb.n_semi))
    x = hip.put(np.zeros(gb.n_semi))
    hip.sync()
    assert L.pecs_solve_dev_noalloc(gb.ctx, pecs_amd.ELECTRONS, b, x) != 0
    assert b"single-instance" in L.pecs_last_error(gb.ctx)
    gb.close()
    g = pecs_amd.SolarCellGPU(p)
    os.environ["PECS_NOALLOC_LIMIT_BYTES"] = "4096"
    try:
        assert L.pecs_solve_dev_noalloc(g.ctx, pecs_amd.ELECTRONS, b, x) != 0
        assert b"above the limit" in L.pecs_last_error(g.ctx)
    finally:
        del os.environ["PECS_NOALLOC_LIMIT_BYTES"]
    assert L.pecs_solve_dev_noalloc(g.ctx, 7, b, x) != 0
    hip.sync()
    g.close()
    hip.release()
