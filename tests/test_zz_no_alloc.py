"""The allocation-free entry cfd_bicg_ex_noalloc (device arrays, work vectors
in the library's static device scratch) against cfd_bicg_ex (host arrays,
allocated work vectors) on the same mesh-Laplace systems: the same kernels in
the same order, so x, the iteration count, r.z and the converged flag agree
bit for bit.  Two calls in a row, the second on a smaller system (the scratch
keeps the first call's contents), then the refusals: a system beyond the
scratch's rows, and an out-of-range column index."""
import ctypes as C

import numpy as np
import pytest

from cfd_amd._lib import CfdError, check, lib
from cfd_amd.api import MeshContext, biCG_ex
from cfd_amd.mesh import ellipse_case

pytestmark = pytest.mark.gpu


def _system(nr, nt, seed):
    cs = ellipse_case(nr, nt)
    m = cs.mesh
    ctx = MeshContext(m.inpoel, m.npoin)
    g = ctx.deriv(m.x, m.y)
    A, dg = ctx.laplace(m.x, m.y, g["dNx"], g["dNy"])
    rp, idx = ctx.lap_structure()
    fidx = np.ascontiguousarray(cs.lists["ilaux"], np.int32)
    xfix = np.random.default_rng(seed).standard_normal(fidx.size) * 1e-3
    return dict(A=A, idx=idx, rp=rp, dg=dg, fidx=fidx, xfix=xfix, n=m.npoin)


class _Dev:
    """Device arrays through the HIP runtime libcfd_amd.so itself links (the
    test allocates; the entry under test does not)."""
    H2D, D2H = 1, 2

    def __init__(self):
        lib()                                        # loads libcfd_amd.so and its runtime
        self.hip = C.CDLL("libamdhip64.so.7")        # the same, already loaded, instance
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                                  self.H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes),
                                  self.D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(sy, x0, tol, maxit, rel, idx=None):
    d = _Dev()
    x0 = np.ascontiguousarray(x0, np.float64)
    p = [d.put(np.asarray(sy["A"], np.float64)),
         d.put(np.asarray(sy["idx"] if idx is None else idx, np.int32)),
         d.put(np.asarray(sy["rp"], np.int32)), d.put(np.asarray(sy["dg"], np.float64)),
         d.put(x0), d.put(np.zeros(sy["n"])), d.put(np.asarray(sy["xfix"], np.float64)),
         d.put(np.asarray(sy["fidx"], np.int32))]
    it, rz, cv = C.c_int(), C.c_double(), C.c_int()
    try:
        check(lib().cfd_bicg_ex_noalloc(sy["n"], p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7],
                                        int(sy["fidx"].size), C.c_double(tol), int(maxit),
                                        int(rel), None, C.byref(it), C.byref(rz), C.byref(cv)))
        x = d.get(p[4], x0)
    finally:
        d.free()
    return x, it.value, rz.value, bool(cv.value)


def _allocating(sy, x0, tol, maxit, rel):
    x = np.array(x0, np.float64)
    it, rz, cv = biCG_ex(sy["A"], sy["idx"], sy["rp"], sy["dg"], x, np.zeros(sy["n"]), sy["xfix"],
                         sy["fidx"], tol=tol, maxit=maxit, rel=rel)
    return x, it, rz, cv


def test_noalloc_pcg_matches_the_allocating_entry():
    big, small = _system(64, 128, 1), _system(24, 48, 2)
    assert small["n"] < big["n"]
    for sy, tol, rel in ((big, 1e-8, 1), (small, 1e-6, 1), (big, 1e-10, 0)):
        x0 = np.random.default_rng(3).standard_normal(sy["n"]) * 1e-4
        print(f"noalloc PCG on {sy['n']} rows, rel {rel}, tol {tol:g}", flush=True)
        xa, ita, rza, cva = _noalloc(sy, x0, tol, 5000, rel)
        xb, itb, rzb, cvb = _allocating(sy, x0, tol, 5000, rel)
        assert ita > 5 and cva
        assert (ita, rza, cva) == (itb, rzb, cvb)
        assert np.array_equal(xa.view(np.int64), xb.view(np.int64))


def test_noalloc_pcg_refusals():
    sy = _system(24, 48, 4)
    x0 = np.zeros(sy["n"])
    bad = sy["idx"].copy()
    bad[7] = sy["n"] + 3
    with pytest.raises(CfdError, match="out of range"):
        _noalloc(sy, x0, 1e-8, 100, 1, idx=bad)
    # refused on its size before any work (small real arrays, never read)
    d = _Dev()
    p = [d.put(np.zeros(8)) for _ in range(6)]
    rc = lib().cfd_bicg_ex_noalloc((1 << 19) + 1, p[0], p[1], p[2], p[3], p[4], p[5], None, None,
                                   0, C.c_double(1e-8), 100, 1, None, None, None, None)
    d.free()
    assert rc != 0 and b"exceeds" in lib().cfd_last_error()
    # and the entry still works afterwards
    xa, ita, _, cva = _noalloc(sy, x0, 1e-8, 5000, 1)
    xb, itb, _, _ = _allocating(sy, x0, 1e-8, 5000, 1)
    assert cva and ita == itb and np.array_equal(xa, xb)
