"""Parity at the benchmark's sizes: the cylinder O-grid 1024 x 2048 (2,097,152
nodes, 4,190,208 triangles; the bench workload) and the pitching ellipse
512 x 1024 (524,288 nodes, 1,046,528 triangles; the ALE comparison).

- the neighbour and Laplace tables bit for bit against the oracle's;
- one step of each FP64 session path from the oracle's state, within 1e-12 per
  conservative component (the same bar as the small teacher-forced tests);
- the FP32-element path's one-step residual within its binary32 bound.

The oracle runs its element loops on up to 16 host threads here (a few
seconds per step at these sizes; the summation order of its atomics then
varies at round-off, far below the 1e-12 bar) and is put back to one thread
afterwards."""
import os

import numpy as np
import pytest

from cfd_amd.mesh import cylinder_case, ellipse_case
from oracle import oracle as orc
from oracle.oracle import OracleSolver
from tests.helpers import relnorm, relnorm_cols

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def threads():
    orc.set_threads(min(16, os.cpu_count() or 1))
    yield
    orc.set_threads(1)


def _advanced(cs, nsteps):
    o = OracleSolver(cs)
    for _ in range(nsteps):
        o.step()
    return o


@pytest.fixture(scope="module")
def big_cyl(threads):
    cs = cylinder_case(1024, 2048, mach=0.8, re=1000.0)
    return cs, _advanced(cs, 2)


@pytest.fixture(scope="module")
def big_ell(threads):
    cs = ellipse_case()
    return cs, _advanced(cs, 2)


def test_full_size_tables(big_cyl):
    from cfd_amd.api import MeshContext
    cs, o = big_cyl
    ctx = MeshContext(cs.mesh.inpoel, cs.mesh.npoin)
    e1, e2 = ctx.getEsup()
    p1, p2 = ctx.getPsup()
    r, i = ctx.lap_structure()
    t = o.tab
    assert e1.size == 3 * cs.nelem and i.size == 2048 * (7 * 1024 - 4)   # Ntheta (7 Nr - 4)
    for name, a, b in [("esup1", e1, t["esup1"]), ("esup2", e2, t["esup2"]),
                       ("psup1", p1, t["psup1"]), ("psup2", p2, t["psup2"]),
                       ("lap_rowptr", r, t["lap_rowptr"]), ("lap_idx", i, t["lap_idx"])]:
        assert a.shape == b.shape and np.array_equal(a, b), name


def _one_step(cs, o0, precision):
    from cfd_amd.solver import Session
    o = OracleSolver(cs)
    o.set_state(o0.state())
    o.disn2 = o0.disn2
    S = Session(cs, precision=precision)
    S.set_state(o.state())
    S.set_scalars(DISN2=o.disn2)
    o.step()
    S.step(1)
    ref, got = o.state(), S.state()
    S.close()
    return ref, got


@pytest.mark.parametrize("precision", [0, 2])
def test_full_size_cylinder_step(big_cyl, precision):
    cs, o0 = big_cyl
    ref, got = _one_step(cs, o0, precision)
    rel = relnorm_cols(got["U"], ref["U"])
    print(f"precision {precision}: U rel err per component {rel}")
    assert got["DTMIN"] == ref["DTMIN"]
    assert np.all(rel < TOL), rel
    rhs = relnorm_cols(got["RHS"], ref["RHS"])
    print(f"precision {precision}: RHS rel err per component {rhs}")
    assert np.all(rhs < TOL), rhs
    for f in ("T", "VEL_X", "VEL_Y", "P"):
        assert relnorm(got[f], ref[f]) < TOL, f
    # the state moved: the comparison is not of two copies of the input
    assert relnorm(ref["U"], o0.state()["U"]) > 1e-9


@pytest.mark.parametrize("precision", [0, 2])
def test_full_size_pitching_ellipse_step(big_ell, precision):
    cs, o0 = big_ell
    ref, got = _one_step(cs, o0, precision)
    rel = relnorm_cols(got["U"], ref["U"])
    print(f"precision {precision}: U rel err per component {rel}")
    assert got["DTMIN"] == ref["DTMIN"]
    assert np.all(rel < TOL), rel
    rhs = relnorm_cols(got["RHS"], ref["RHS"])
    print(f"precision {precision}: RHS rel err per component {rhs}")
    assert np.all(rhs < TOL), rhs
    for f in ("T", "VEL_X", "VEL_Y", "P", "X", "Y", "M"):
        assert relnorm(got[f], ref[f]) < TOL, f
    # both mesh solves stop under the reference's absolute rule
    assert relnorm(got["XPOS"], ref["XPOS"]) < 1e-6 or np.abs(ref["XPOS"]).max() == 0


def test_full_size_relative_pcg(big_ell):
    """The mesh-Laplace system of the ALE comparison (524,288 rows) solved by
    the multi-kernel PCG under the relative rule (||r||_z <= tol ||r0||_z of
    the guess that is zero off the fixed nodes), as the bench's ALE runs
    solve it: for tol 1e-6, 1e-8 and 1e-10, from a zero guess and from the
    solution of a slightly smaller rotation (a previous step's), the rule is
    met by the residual recomputed with scipy and the Dirichlet values are
    kept; the warm start needs fewer iterations."""
    import math
    import scipy.sparse as sp
    from cfd_amd.api import biCG_ex
    cs, o = big_ell
    s, t = o.s, o.tab
    n = cs.npoin
    m = cs.mesh
    body = m.node(0, np.arange(m.nt)).astype(np.int32)
    outer = m.node(m.nr - 1, np.arange(m.nt)).astype(np.int32)
    fidx = np.concatenate([body, outer]).astype(np.int32)
    px, py = s["X"][body - 1] - cs.xref, s["Y"][body - 1] - cs.yref
    rows = np.repeat(np.arange(n), np.diff(t["lap_rowptr"]))
    K = sp.csr_matrix((s["LAP_SPARSE"], (rows, t["lap_idx"] - 1)), shape=(n, n))
    fixed = np.zeros(n, bool)
    fixed[fidx - 1] = True
    args = (s["LAP_SPARSE"], t["lap_idx"], t["lap_rowptr"], s["LAP_DIAG"])

    def rot(deg):
        a = math.radians(deg)
        return np.concatenate([math.cos(a) * px + math.sin(a) * py - px, np.zeros(outer.size)])

    def rz_of(x, xfix):
        v = x.copy()
        v[fidx - 1] = xfix
        r = -(K @ v)
        r[fixed] = 0.0
        return float(r @ (r / s["LAP_DIAG"]))

    xfix, xprev = rot(0.05), rot(0.049)
    rz0 = rz_of(np.zeros(n), xfix)
    x_old = np.zeros(n)
    biCG_ex(*args, x_old, np.zeros(n), xprev, fidx, tol=1e-11, maxit=20000, rel=1)
    for tol in (1e-6, 1e-8, 1e-10):
        its = []
        for x in (np.zeros(n), x_old.copy()):
            it, rz, cv = biCG_ex(*args, x, np.zeros(n), xfix, fidx, tol=tol, maxit=20000, rel=1)
            rzn = rz_of(x, xfix)
            print(f"tol {tol:g}: {it} iterations, recurrence r.z {rz:.3e}, recomputed "
                  f"{rzn:.3e}, bound {tol * tol * rz0:.3e}")
            assert cv
            assert abs(rz) <= tol * tol * rz0 * 1.0001
            assert rzn <= 1.5 * tol * tol * rz0, (tol, rzn, tol * tol * rz0)
            assert np.array_equal(x[fidx - 1], xfix)
            its.append(it)
        assert its[1] < its[0], its


def test_full_size_fp32_residual(big_cyl):
    cs, o0 = big_cyl
    ref, got = _one_step(cs, o0, 1)
    rel_rhs = relnorm_cols(got["RHS"], ref["RHS"])
    print("FP32 element: RHS rel err per component", rel_rhs)
    assert np.all(rel_rhs < 2e-4), rel_rhs
    # two steps from the free stream the y-momentum is still small (its norm
    # is set by the flow near the body), so its relative error is the largest
    # (measured 2.9e-7; the other components about 1e-10)
    assert np.all(relnorm_cols(got["U"], ref["U"]) < 1e-6)
    assert got["DTMIN"] == ref["DTMIN"]


@pytest.mark.parametrize("precision", [0, 1])
def test_full_size_renumbered_session_bitwise(big_cyl, precision):
    """The bench's configuration (fused path, patch-order numbering inside the
    session) at full size: two steps from the oracle's state, bit for bit the
    session in the caller's numbering."""
    from cfd_amd.solver import Session
    cs, o0 = big_cyl
    st = o0.state()
    out = []
    for rn in (0, 1):
        S = Session(cs, precision=precision, renumber=rn)
        S.set_state(st)
        S.set_scalars(DISN2=o0.disn2)
        S.step(2)
        out.append({k: S.get(k) for k in ("U", "T", "RHS", "DTL", "SHOC")} | S.scalars())
        S.close()
    a, b = out
    for k in ("U", "T", "RHS", "DTL", "SHOC"):
        assert np.array_equal(a[k], b[k]), k
    assert a["DTMIN"] == b["DTMIN"] and a["FX1"] == b["FX1"]


def test_full_size_batch_sweep():
    """The 256-case sweep (10.8M nodes, 21.5M triangles) as the bench runs it:
    every case finite after 20 steps with its own positive time step, and two
    of its cases bit for bit against their single-case sessions."""
    from cfd_amd.batch import Batch
    from cfd_amd.mesh import sweep_cases
    from cfd_amd.solver import Session
    cases = sweep_cases(256)
    B = Batch(cases, precision=1)
    B.step(20)
    B.sync()
    dts = [B.scalars(k)["DTMIN"] for k in range(len(cases))]
    assert all(np.isfinite(d) and d > 0 for d in dts)
    for k in range(len(cases)):
        assert np.isfinite(B.get(k, "U")).all(), k
    for k in (0, 137):
        S = Session(cases[k], precision=1)
        S.step(20)
        assert np.array_equal(B.get(k, "U"), S.get("U")), k
        assert B.scalars(k)["DTMIN"] == S.scalars()["DTMIN"], k
        S.close()
    B.close()
