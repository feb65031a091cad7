"""GPU parity: the gfx950 library (through its C ABI) against the oracle on
identical inputs.

Bars: neighbour / CSR tables bit-exact; gather-based sums (MASAS, laplace)
bit-exact; FP64 element residuals and stage updates within 1e-12 of the
oracle, norm-relative per conservative component (pow() on the device and in
glibc may differ in the last bit; every other operation is IEEE-rounded in the
same order).  States for the residual tests are developed states produced by
the oracle itself.
"""
import numpy as np
import pytest

from cfd_amd.mesh import cylinder_case, ellipse_case
from oracle.oracle import OracleSolver
from tests.helpers import bitdiff, relnorm, relnorm_cols

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def cyl():
    cs = cylinder_case(64, 128, mach=0.8, re=1000.0)
    o = OracleSolver(cs)
    for _ in range(30):
        o.step()
    return cs, o


@pytest.fixture(scope="module")
def ell():
    cs = ellipse_case(32, 64)
    o = OracleSolver(cs)
    for _ in range(10):
        o.step()
    return cs, o


@pytest.fixture(scope="module")
def slip():
    # body edges as slip walls: NORMALES + NORMALVEL on every stage
    from cfd_amd.mesh import acoustic_fsafe
    cs = cylinder_case(40, 80, mach=0.5, re=500.0, slip_body=True, fsafe=acoustic_fsafe(0.5))
    o = OracleSolver(cs)
    for _ in range(15):
        o.step()
    return cs, o


@pytest.fixture(scope="module")
def c1flow():
    # the CPU-runnable configuration's flow: M 0.3, Re 100, FSAFE 0.5
    cs = cylinder_case(64, 128, mach=0.3, re=100.0)
    o = OracleSolver(cs)
    for _ in range(40):
        o.step()
    return cs, o


@pytest.fixture(scope="module")
def itloc():
    # local time stepping (ITLOCAL /= 0): DTL blends DTMIN with each element's DT
    from cfd_amd.mesh import acoustic_fsafe
    cs = cylinder_case(40, 80, mach=0.5, re=500.0, itlocal=25, fsafe=acoustic_fsafe(0.5))
    o = OracleSolver(cs)
    for _ in range(8):
        o.step()
    assert np.abs(o.s["DTL"] - o.dtmin).max() > 0          # the blend is active
    return cs, o


@pytest.fixture(scope="module")
def pitch1():
    # the parameterised pitching law alpha(t) = amp sin(omega t) (motion = 1)
    cs = ellipse_case(32, 64, motion=1, motion_amp=0.2, motion_omega=60.0)
    o = OracleSolver(cs)
    for _ in range(10):
        o.step()
    assert o.disn2 != 0.0
    return cs, o


@pytest.fixture(scope="module")
def mctx(cyl):
    from cfd_amd.api import MeshContext
    cs, _ = cyl
    return MeshContext(cs.mesh.inpoel, cs.mesh.npoin)


def test_tables_bit_exact(cyl, mctx):
    cs, o = cyl
    e1, e2 = mctx.getEsup()
    p1, p2 = mctx.getPsup()
    r, i = mctx.lap_structure()
    t = o.tab
    for name, a, b in [("esup1", e1, t["esup1"]), ("esup2", e2, t["esup2"]), ("psup1", p1, t["psup1"]),
                       ("psup2", p2, t["psup2"]), ("lap_rowptr", r, t["lap_rowptr"]),
                       ("lap_idx", i, t["lap_idx"])]:
        assert a.shape == b.shape, name
        assert np.array_equal(a, b), name


def test_tables_ellipse_and_ragged():
    """A second mesh family plus an unstructured, renumbered mesh."""
    from cfd_amd.api import MeshContext
    from oracle.oracle import tables
    rng = np.random.default_rng(7)
    cs = ellipse_case(24, 40)
    m = cs.mesh
    # renumber nodes and shuffle elements: tables must still match the
    # reference algorithm on the permuted input
    perm = rng.permutation(m.npoin)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.npoin)
    inp = (inv[m.inpoel - 1] + 1).astype(np.int32)[rng.permutation(m.nelem)]
    for ip in (m.inpoel, inp):
        ctx = MeshContext(ip, m.npoin)
        ref = tables(np.ascontiguousarray(ip), m.npoin)
        e1, e2 = ctx.getEsup()
        p1, p2 = ctx.getPsup()
        r, i = ctx.lap_structure()
        assert np.array_equal(e1, ref["esup1"]) and np.array_equal(e2, ref["esup2"])
        assert np.array_equal(p1, ref["psup1"]) and np.array_equal(p2, ref["psup2"])
        assert np.array_equal(r, ref["lap_rowptr"]) and np.array_equal(i, ref["lap_idx"])


def test_mesh_rejects_bad_ids():
    from cfd_amd._lib import CfdError
    from cfd_amd.api import MeshContext
    with pytest.raises(CfdError):
        MeshContext(np.array([[1, 2, 5]], np.int32), 4)
    with pytest.raises(CfdError):
        MeshContext(np.array([[1, 2, 2]], np.int32), 4)


def test_deriv_masas_laplace(ell):
    from cfd_amd.api import MeshContext
    cs, o = ell
    s = o.s
    ctx = MeshContext(cs.mesh.inpoel, cs.mesh.npoin)
    g = ctx.deriv(s["X"], s["Y"])
    for k, ref in [("area", s["AREA"]), ("dNx", s["DNX"]), ("dNy", s["DNY"]), ("HH", s["HH"]),
                   ("HHX", s["HHX"]), ("HHY", s["HHY"])]:
        assert bitdiff(g[k], ref) == 0, k
    assert g["hmin"] == o.hmin
    M = ctx.MASAS(s["AREA"])
    assert bitdiff(M, s["M"]) == 0
    ls, ld = ctx.laplace(s["X"], s["Y"], s["DNX"], s["DNY"])
    assert bitdiff(ls, s["LAP_SPARSE"]) == 0
    assert bitdiff(ld, s["LAP_DIAG"]) == 0


def test_deltat(cyl, mctx):
    from cfd_amd.api import consts_of
    cs, o = cyl
    s = o.s
    DT, dtmin = mctx.deltat(consts_of(cs), s["AREA"], s["T"], s["VEL_X"], s["VEL_Y"], s["W_X"], s["W_Y"])
    dt_ref = s["DT"].copy()
    dtmin_ref = o.deltat()
    assert abs(dtmin - dtmin_ref) <= 1e-14 * dtmin_ref
    assert relnorm(DT, s["DT"]) < 1e-14
    s["DT"][:] = dt_ref


def test_estab(cyl, mctx):
    from cfd_amd.api import consts_of
    cs, o = cyl
    s = o.s
    dtmin = o.dtmin
    out = mctx.ESTAB(consts_of(cs), s["U"], s["T"], s["GAMM"], s["VEL_X"], s["VEL_Y"], s["W_X"],
                     s["W_Y"], s["DNX"], s["DNY"], dtmin)
    o.estab(dtmin)
    for a, k in zip(out, ("SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3")):
        assert relnorm(a, s[k]) < TOL, k


@pytest.mark.parametrize("with_theta", [False, True])
def test_calcrhs(cyl, mctx, with_theta):
    from cfd_amd.api import consts_of
    cs, o = cyl
    s = o.s
    rng = np.random.default_rng(3)
    theta = rng.standard_normal((cs.npoin, 4)) * 1e-3 if with_theta else None
    rhs0 = rng.standard_normal((cs.npoin, 4)) * 1e-6      # accumulate semantics
    ref = o.calcrhs(rhs=rhs0.copy(), theta=theta)
    got = mctx.calcRHS(consts_of(cs), rhs0.copy(), s["U"], theta, s["T"], s["DNX"], s["DNY"],
                       s["AREA"], s["SHOC"], s["DTL"], s["T_SUGN1"], s["T_SUGN2"], s["T_SUGN3"])
    assert np.all(relnorm_cols(got, ref) < TOL), relnorm_cols(got, ref)


def test_fuente(cyl, mctx):
    cs, o = cyl
    s = o.s
    rng = np.random.default_rng(4)
    wx = rng.standard_normal(cs.npoin)
    wy = rng.standard_normal(cs.npoin)
    save = s["W_X"].copy(), s["W_Y"].copy()
    s["W_X"][:], s["W_Y"][:] = wx, wy
    ref = o.fuente()
    s["W_X"][:], s["W_Y"][:] = save
    got = mctx.FUENTE(np.zeros((cs.npoin, 4)), s["U"], wx, wy, s["DNX"], s["DNY"], s["AREA"], s["DTL"])
    assert np.all(relnorm_cols(got, ref) < TOL)


def test_force_visc_entry(cyl, mctx):
    """cfd_force_visc (FORCE_VISC, ns2DComp.ALE.f90:819-893) against the
    oracle on a developed state (the device sums the edge terms by a fixed
    tree, the oracle in edge order: equal within round-off of the terms'
    magnitudes); the set repeated 101 times (more than 64 partial sums to
    fold) gives 101 times the force."""
    cs, o = cyl
    s = o.s
    iset = np.asarray(cs.lists["iset"], np.int32)
    got = mctx.force_visc(iset, s["X"], s["Y"], s["P"], s["T"], s["VEL_X"], s["VEL_Y"], s["DNX"],
                          s["DNY"], cs.t_inf)
    ref = o.force_visc()
    sc = force_visc_scale(cs, s["X"], s["Y"], s["P"], s)
    print("FORCE_VISC got", got, "ref", ref, "scale", sc)
    for a, b in zip(got, ref):
        assert abs(a - b) <= 1e-14 * sc
    rep = np.tile(iset.reshape(-1, 3), (101, 1))
    assert rep.shape[0] > 64 * 64
    got_r = mctx.force_visc(rep, s["X"], s["Y"], s["P"], s["T"], s["VEL_X"], s["VEL_Y"],
                            s["DNX"], s["DNY"], cs.t_inf)
    for a, b in zip(got_r, ref):
        assert abs(a - 101 * b) <= 101 * 1e-14 * sc
    # an empty set gives zero, a bad element id is refused
    assert mctx.force_visc(np.zeros((0, 3), np.int32), s["X"], s["Y"], s["P"], s["T"],
                           s["VEL_X"], s["VEL_Y"], s["DNX"], s["DNY"], cs.t_inf) == (0.0, 0.0)
    from cfd_amd._lib import CfdError
    bad = iset[:2].copy()
    bad[1, 2] = cs.nelem + 1
    with pytest.raises(CfdError):
        mctx.force_visc(bad, s["X"], s["Y"], s["P"], s["T"], s["VEL_X"], s["VEL_Y"], s["DNX"],
                        s["DNY"], cs.t_inf)


def test_normalvel_entry(slip):
    """cfd_normalvel (NORMALVEL, subrutinas.f90:66-85) with the normals of
    cfd_normales, against the oracle's, bit for bit; refused before
    cfd_normales on a fresh mesh context."""
    from cfd_amd._lib import CfdError
    from cfd_amd.api import MeshContext
    cs, o = slip
    s = o.s
    ctx = MeshContext(cs.mesh.inpoel, cs.mesh.npoin)
    rng = np.random.default_rng(8)
    vx, vy, wx, wy = (rng.standard_normal(cs.npoin) for _ in range(4))
    with pytest.raises(CfdError):
        ctx.normalvel(vx, vy, wx, wy)
    ctx.normales(cs.lists, s["X"], s["Y"])
    gx, gy = ctx.normalvel(vx, vy, wx, wy)
    o.normales()
    rx, ry = o.normalvel(vx, vy, wx, wy)
    assert bitdiff(gx, rx) == 0 and bitdiff(gy, ry) == 0
    assert bitdiff(gx, vx) > 0


def test_rk(cyl, mctx):
    from cfd_amd.api import consts_of
    cs, o = cyl
    st = o.state()
    mctx.normales(cs.lists, st["X"], st["Y"])
    got = mctx.RK(consts_of(cs), 4, o.dtmin, st["DTL"], dict(st))
    o.set_state(st)
    o.rk(o.dtmin)
    ref = o.state()
    o.set_state(st)
    assert np.all(relnorm_cols(got["U1"], ref["U1"]) < TOL)
    for k in ("T", "VEL_X", "VEL_Y", "RHO", "E", "P", "RMACH", "SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3"):
        assert relnorm(got[k], ref[k]) < TOL, k
    assert np.all(relnorm_cols(got["RHS"], ref["RHS"]) < TOL), relnorm_cols(got["RHS"], ref["RHS"])


def test_bicg(ell):
    from cfd_amd.api import biCG
    cs, o = ell
    s, t = o.s, o.tab
    rng = np.random.default_rng(5)
    L = cs.lists
    fidx = L["ilaux"]
    xfix = rng.standard_normal(fidx.size) * 1e-3
    x0 = np.zeros(cs.npoin)
    xr = x0.copy()
    it_ref, rz_ref = o.bicg(xr, xfix, fidx)
    xg = x0.copy()
    it, rz = biCG(s["LAP_SPARSE"], t["lap_idx"], t["lap_rowptr"], s["LAP_DIAG"], xg,
                  np.zeros(cs.npoin), xfix, fidx)
    assert it_ref > 5
    assert abs(it - it_ref) <= 2        # same iteration count up to reduction order
    assert abs(rz) <= 1e-10             # the reference's stopping test holds
    # both solutions are as close to the exact solution of the Dirichlet
    # problem as the absolute 1e-10 test allows (the sums differ in order only)
    x_exact = _direct_solve(s["LAP_SPARSE"], t["lap_idx"], t["lap_rowptr"], fidx, xfix, cs.npoin)
    err_ref = relnorm(xr, x_exact)
    assert relnorm(xg, x_exact) <= 3.0 * err_ref + 1e-12
    assert relnorm(xg, xr) < 1e-4


def _residual_tests(A, idx, rowptr, diag, fidx, xfix, x, b):
    """(r.r, r.z) of x and of the guess that is zero off the fixed nodes,
    with the solver's conventions (fixed rows carry no residual)."""
    import scipy.sparse as sp
    n = x.size
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    K = sp.csr_matrix((A, (rows, idx - 1)), shape=(n, n))
    fixed = np.zeros(n, bool)
    fixed[fidx - 1] = True
    xd = np.zeros(n)
    xd[fidx - 1] = xfix
    out = []
    for v in (x, xd):
        r = b - K @ v
        r[fixed] = 0.0
        out.append((float(r @ r), float(r @ (r / diag))))
    return out


def test_bicg_relative_mode_and_warm_starts(ell):
    """Relative stopping rule: the test is set by the guess that is zero off
    the fixed nodes, so a better warm start stops sooner at the same
    tolerance; an exact guess stops at once; the result meets the rule and
    approaches the direct solution."""
    from cfd_amd.api import biCG_ex
    cs, o = ell
    s, t = o.s, o.tab
    rng = np.random.default_rng(11)
    fidx = cs.lists["ilaux"]
    xfix = rng.standard_normal(fidx.size) * 1e-3
    n = cs.npoin
    A, idx, rp, dg = s["LAP_SPARSE"], t["lap_idx"], t["lap_rowptr"], s["LAP_DIAG"]
    x_exact = _direct_solve(A, idx, rp, fidx, xfix, n)
    tol = 1e-6
    x0 = np.zeros(n)
    it0, rz0, cv0 = biCG_ex(A, idx, rp, dg, x0, np.zeros(n), xfix, fidx, tol=tol, rel=1)
    (rr, rz), (rr_d, rz_d) = _residual_tests(A, idx, rp, dg, fidx, xfix, x0, np.zeros(n))
    assert cv0 and it0 > 3
    assert abs(rz) <= tol * tol * rz_d * 1.0001
    assert relnorm(x0, x_exact) < 1e-3
    # a warm start 1e-3 away from the solution: fewer iterations, same rule
    xw = x_exact + 1e-3 * np.abs(xfix).max() * rng.standard_normal(n)
    itw, _, cvw = biCG_ex(A, idx, rp, dg, xw, np.zeros(n), xfix, fidx, tol=tol, rel=1)
    (rr, rz), (rr_d2, rz_d2) = _residual_tests(A, idx, rp, dg, fidx, xfix, xw, np.zeros(n))
    assert cvw and itw < it0
    assert rz_d2 == pytest.approx(rz_d, rel=1e-12)
    assert abs(rz) <= tol * tol * rz_d * 1.0001
    # the exact solution passes the first residual test
    xe = x_exact.copy()
    ite, _, cve = biCG_ex(A, idx, rp, dg, xe, np.zeros(n), xfix, fidx, tol=1e-4, rel=1)
    assert cve and ite == 0


@pytest.mark.skip(reason="written for a persistent cooperative PCG grid (several rows per "
                         "thread); that kernel went silent for 180 s on this 563k-row system "
                         "with 2048 blocks, no cause was shown, and it was switched off. The "
                         "multi-kernel PCG runs one row per thread.")
def test_bicg_many_rows_per_thread():
    """A system with more rows than the co-resident grid has threads (each
    thread walks several rows): the absolute and relative rules hold for the
    returned x."""
    from cfd_amd.api import MeshContext, biCG_ex
    from cfd_amd.mesh import ogrid
    mesh = ogrid(1100, 512, "circle")
    mc = MeshContext(mesh.inpoel, mesh.npoin)
    g = mc.deriv(mesh.x, mesh.y)
    A, dg = mc.laplace(mesh.x, mesh.y, g["dNx"], g["dNy"])
    rp, idx = mc.lap_structure()
    n = mesh.npoin
    nt = 512
    fidx = np.concatenate([np.arange(1, nt + 1), np.arange(n - nt + 1, n + 1)]).astype(np.int32)
    rng = np.random.default_rng(12)
    xfix = np.concatenate([rng.standard_normal(nt) * 1e-3, np.zeros(nt)])
    _, (rr_d, rz_d) = _residual_tests(A, idx, rp, dg, fidx, xfix, np.zeros(n), np.zeros(n))
    # relative 1e-5 (r.z down 1e-10), then the absolute rule at the same level
    for rel, tol in ((1, 1e-5), (0, 1e-10 * rz_d)):
        x = np.zeros(n)
        it, rz, cv = biCG_ex(A, idx, rp, dg, x, np.zeros(n), xfix, fidx, tol=tol, maxit=20000,
                             rel=rel)
        (rr, rzn), _ = _residual_tests(A, idx, rp, dg, fidx, xfix, x, np.zeros(n))
        assert cv and it > 10, (rel, it)
        bound = tol * tol * rz_d if rel else tol
        assert abs(rz) <= bound, (rel, rz, bound)
        # the recomputed (true) residual agrees with the recurrence's within a few %
        assert abs(rzn) <= 1.1 * bound, (rel, rzn, bound)
        assert np.array_equal(x[fidx - 1], xfix)


def _direct_solve(A, idx, rowptr, fidx, xfix, n):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    K = sp.csr_matrix((A, (rows, idx - 1)), shape=(n, n))
    fixed = np.zeros(n, bool)
    fixed[fidx - 1] = True
    x = np.zeros(n)
    x[fidx - 1] = xfix
    free = ~fixed
    rhs = -K[free][:, fixed] @ x[fixed]
    x[free] = spl.spsolve(K[free][:, free].tocsc(), rhs)
    return x


@pytest.mark.parametrize("mode", [0, 1])
def test_gcl(ell, mode):
    from cfd_amd.api import MeshContext
    cs, o = ell
    s = o.s
    rng = np.random.default_rng(6)
    ctx = MeshContext(cs.mesh.inpoel, cs.mesh.npoin)
    wx, wy, wxo, wyo = (rng.standard_normal(cs.npoin) for _ in range(4))
    area_old = s["AREA"] * (1 + 1e-3 * rng.standard_normal(cs.nelem))
    save = {k: s[k].copy() for k in ("M", "W_X", "W_Y", "W_X_OLD", "W_Y_OLD", "AREA_OLD")}
    s["W_X"][:], s["W_Y"][:], s["W_X_OLD"][:], s["W_Y_OLD"][:], s["AREA_OLD"][:] = wx, wy, wxo, wyo, area_old
    o.gcl(1e-3, mode)
    ref = s["M"].copy()
    for k, v in save.items():
        s[k][:] = v
    from cfd_amd._lib import CfdError
    M = s["M"].copy()
    with pytest.raises(CfdError):
        ctx.gcl_main(M, wx, wy, s["DNX"], s["DNY"], s["AREA"], 1e-3, mode)
    ctx.putW(wxo, wyo)
    ctx.putArea(area_old)
    ctx.gcl_main(M, wx, wy, s["DNX"], s["DNY"], s["AREA"], 1e-3, mode)
    assert bitdiff(M, ref) == 0


def _session_from_oracle(cs, o, precision=0):
    from cfd_amd.solver import Session
    S = Session(cs, precision=precision)
    S.set_state(o.state())
    S.set_scalars(DISN2=o.disn2)
    return S


@pytest.mark.parametrize("precision", [0, 2])
@pytest.mark.parametrize("which", ["cyl", "ell", "slip", "c1flow", "itloc", "pitch1"])
def test_session_teacher_forced(which, precision, request):
    """Each step from the oracle's state; the GPU step must land within 1e-12:
    precision 0 = the fused FP64 path (regrouped element arithmetic, patch
    sums), 2 = the FP64 path in the reference's operation and accumulation
    order."""
    cs, o0 = request.getfixturevalue(which)
    o = OracleSolver(cs)
    o.set_state(o0.state())
    o.disn2 = o0.disn2
    S = _session_from_oracle(cs, o, precision)
    for k in range(3):
        st = o.state()
        S.set_state(st)
        S.set_scalars(DISN2=o.disn2)
        o.step()
        S.step(1)
        ref = o.state()
        got = S.state()
        assert got["DTMIN"] == ref["DTMIN"]
        assert abs(got["TIME"] - ref["TIME"]) <= 1e-15 * ref["TIME"]
        assert np.all(relnorm_cols(got["U"], ref["U"]) < TOL), (k, relnorm_cols(got["U"], ref["U"]))
        # the final stage's residual itself, not only the state it moved
        rhs = relnorm_cols(got["RHS"], ref["RHS"])
        print(f"{which} precision {precision} step {k}: RHS rel err per component {rhs}")
        assert np.all(rhs < TOL), (k, rhs)
        for f in ("T", "VEL_X", "VEL_Y", "P", "X", "Y", "M"):
            assert relnorm(got[f], ref[f]) < TOL, (k, f)
        assert relnorm(got["XPOS"], ref["XPOS"]) < 1e-6 or np.abs(ref["XPOS"]).max() == 0
        assert_forces_match(cs, st, ref, got)


@pytest.mark.parametrize("precision", [0, 1])
def test_mesh_at_rest_skips_the_w_reads(cyl, precision):
    """On a fixed mesh whose W the caller never set (W = 0), the patch-order
    deltat and ESTAB do not read W (v - 0 = v exactly): a step from the
    oracle's state loaded without W lands within the path's bar of the
    oracle (1e-12 for the fused FP64 path, 2e-4 for FP32 elements), and its
    DT, SHOC and T_SUGN1..3 equal bit for bit those of the same step on a
    session that loaded W = 0 and therefore reads it."""
    from cfd_amd.solver import FIELD_OF, Session
    cs, o0 = cyl
    st = o0.state()
    assert np.abs(st["W_X"]).max() == 0 and np.abs(st["W_Y"]).max() == 0
    no_w = [k for k in FIELD_OF if k in st and k not in ("W_X", "W_Y", "XPOS", "YPOS")]
    o = OracleSolver(cs)
    o.set_state(st)
    o.step()
    ref = o.state()
    out = {}
    for tag, fields in (("skip", no_w), ("read", no_w + ["W_X", "W_Y"])):
        S = Session(cs, precision=precision)
        S.set_state(st, fields=fields)
        S.step(1)
        out[tag] = S.state()
        S.close()
    got = out["skip"]
    tol = TOL if precision == 0 else 2e-4
    assert got["DTMIN"] == ref["DTMIN"]
    assert np.all(relnorm_cols(got["U"], ref["U"]) < tol), relnorm_cols(got["U"], ref["U"])
    assert np.all(relnorm_cols(got["RHS"], ref["RHS"]) < tol), relnorm_cols(got["RHS"], ref["RHS"])
    for f in ("DT", "SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3"):
        assert bitdiff(out["skip"][f], out["read"][f]) == 0, f


def assert_forces_match(cs, st, ref, got, tol=TOL):
    """FORCES (meshMove.f90:153-194) on the GPU against the oracle after a
    teacher-forced step: FX, FY and RM within tol of the sum of the
    magnitudes of their edge terms (a closed body's forces cancel, so a
    relative error of FY itself means nothing on a symmetric flow)."""
    iset = np.asarray(cs.lists["iset"])
    n1, n2 = iset[:, 0] - 1, iset[:, 1] - 1
    X, Y, P = st["X"], st["Y"], ref["P"]
    pav = np.abs(P[n1] + P[n2]) / 2.0
    dx, dy = np.abs(X[n1] - X[n2]), np.abs(Y[n2] - Y[n1])
    arm = np.abs(X[n1] - cs.xref) + np.abs(X[n2] - cs.xref) + np.abs(Y[n1] - cs.yref) + \
        np.abs(Y[n2] - cs.yref)
    scale = {"FX1": (pav * dy).sum(), "FY1": (pav * dx).sum(), "RM1": (pav * (dx + dy) * arm).sum()}
    # FORCE_VISC: element-average pressure plus viscous stresses on each edge
    if cs.fmu != 0.0:
        vs = force_visc_scale(cs, X, Y, P, ref)
        scale.update(FVX1=vs, FVY1=vs)
        assert abs(ref["FVX1"]) > 0
    for k, sc in scale.items():
        assert abs(got[k] - ref[k]) <= tol * sc, (k, got[k], ref[k], sc)


def force_visc_scale(cs, X, Y, P, st, iset=None):
    """Sum over the force-set edges of |traction| x edge length, bounded by
    the element-average |p| plus 0.1 |velocity gradient| (0.1 exceeds the
    Sutherland viscosity here): the scale a summation-order difference of
    FORCE_VISC is measured against (a symmetric flow's lift is near zero)."""
    iset = np.asarray(cs.lists["iset"] if iset is None else iset)
    n1, n2, el = iset[:, 0] - 1, iset[:, 1] - 1, iset[:, 2] - 1
    vtx = cs.mesh.inpoel[el] - 1
    pe = np.abs(P[vtx]).mean(axis=1)
    dnx, dny = np.asarray(st["DNX"]).reshape(-1, 3), np.asarray(st["DNY"]).reshape(-1, 3)
    g = sum(np.abs((dnx[el] if k == 0 else dny[el]) * np.asarray(st[v])[vtx]).sum(axis=1)
            for k in (0, 1) for v in ("VEL_X", "VEL_Y"))
    ln = np.hypot(X[n1] - X[n2], Y[n1] - Y[n2])
    return float(((pe + 0.1 * g) * ln).sum())


def test_session_free_run_and_determinism(cyl):
    cs, _ = cyl
    from cfd_amd.solver import Session
    o = OracleSolver(cs)
    runs = []
    for _ in range(2):
        S = Session(cs)
        S.step(20)
        runs.append(S.state())
        S.close()
    for _ in range(20):
        o.step()
    a, b = runs
    for k in ("U", "T", "RHS", "SHOC", "T_SUGN2"):
        assert bitdiff(a[k], b[k]) == 0, k           # run-to-run bitwise
    # Free running, trajectories separate: ESTAB takes a unit vector from the
    # rounding residue of vel2*sum(dN) (subrutinas.f90:380-393), so a last-bit
    # difference of pow() (device vs glibc) changes tau_2/tau_3 by O(1) in some
    # elements one step later.  The parity bar is the teacher-forced step test
    # above; here the drift only has to stay bounded.
    rel = relnorm_cols(a["U"], o.s["U"])
    print("free-run U drift after 20 steps per component:", rel)
    assert np.all(np.isfinite(a["U"]))
    assert np.all(rel < 5e-2), rel


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_session_timing_counts_and_leaves_results(cyl, precision):
    """HIP-event timing of the session's slots: the element slot counts every
    stage launch (4 a step; on the fused paths one event pair spans the
    step's stages), one interval a step for deltat, ESTAB and the whole step,
    positive totals, and the state after timed steps equal bit for bit to
    the untimed run's."""
    cs, _ = cyl
    from cfd_amd.solver import Session
    runs = []
    for timed in (False, True):
        S = Session(cs, precision=precision)
        if timed:
            S.timing(True, slots=["ELEMENT"])
            S.step(3)
            S.sync()
            t, n = S.kernel_time("ELEMENT")
            assert n == 12 and t > 0.0, (t, n)
            S.timing(True)  # every slot
            S.step(2)
            S.sync()
            counts = {"ELEMENT": 8, "DELTAT": 2, "ESTAB": 2, "STEP": 2}
            if precision == 2:
                counts["STAGE"] = 8
            for k, nk in counts.items():
                t, n = S.kernel_time(k)
                assert n == nk and t > 0.0, (k, t, n)
            S.timing(False)
        else:
            S.step(5)
        runs.append(S.state())
        S.close()
    for k in ("U", "T", "RHS", "P"):
        assert bitdiff(runs[0][k], runs[1][k]) == 0, k


def test_estab_bitwise_fraction(cyl, mctx):
    """Diagnostic: share of ESTAB outputs identical to the last bit (only the
    pow() calls can differ)."""
    from cfd_amd.api import consts_of
    cs, o = cyl
    s = o.s
    out = mctx.ESTAB(consts_of(cs), s["U"], s["T"], s["GAMM"], s["VEL_X"], s["VEL_Y"], s["W_X"],
                     s["W_Y"], s["DNX"], s["DNY"], o.dtmin)
    o.estab(o.dtmin)
    fr = {k: 1.0 - bitdiff(a, s[k]) / a.size for a, k in zip(out, ("SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3"))}
    print("ESTAB bitwise-identical fraction:", fr)
    assert fr["SHOC"] == 1.0      # no pow on this path


# ---------------------------------------------------------------------------
# FP32-element / FP64-accumulation path
# ---------------------------------------------------------------------------
def test_fp32_element_step_against_fp64_oracle(cyl):
    """One teacher-forced step: the FP32 element residual (FP64 gather and
    update) against the FP64 oracle; binary32 arithmetic bounds the error."""
    from cfd_amd.solver import Session
    cs, o0 = cyl
    o = OracleSolver(cs)
    o.set_state(o0.state())
    S = Session(cs, precision=1)
    S.set_state(o.state())
    o.step()
    S.step(1)
    ref, got = o.state(), S.state()
    rel_rhs = relnorm_cols(got["RHS"], ref["RHS"])
    print("FP32 element: RHS rel err per component", rel_rhs)
    assert np.all(rel_rhs < 2e-4), rel_rhs
    assert np.all(relnorm_cols(got["U"], ref["U"]) < 1e-7)
    assert got["DTMIN"] == ref["DTMIN"]


def test_fp32_element_moving_mesh_step_against_fp64_oracle(ell):
    """The same on the pitching ellipse: FP32 calcRHS plus the FUENTE mesh
    velocity term, against the FP64 oracle after a teacher-forced step (the
    state carries a nonzero mesh velocity from the oracle's own motion)."""
    from cfd_amd.solver import Session
    cs, o0 = ell
    o = OracleSolver(cs)
    o.set_state(o0.state())
    o.disn2 = o0.disn2
    assert np.abs(o.s["W_X"]).max() > 0
    S = Session(cs, precision=1)
    S.set_state(o.state())
    S.set_scalars(DISN2=o.disn2)
    o.step()
    S.step(1)
    ref, got = o.state(), S.state()
    rel_rhs = relnorm_cols(got["RHS"], ref["RHS"])
    print("FP32 element, moving mesh: RHS rel err per component", rel_rhs)
    assert np.all(rel_rhs < 2e-4), rel_rhs
    assert np.all(relnorm_cols(got["U"], ref["U"]) < 1e-7)
    for k in ("X", "Y"):                          # FP64 mesh solve, same rule
        assert relnorm(got[k], ref[k]) < 1e-10


def test_fp32_free_run_density_and_forces(cyl):
    """Free-running FP32-element vs FP64 paths on the GPU: L2 density error
    and pressure force after 200 steps."""
    from cfd_amd.solver import Session
    cs, _ = cyl
    a = Session(cs, precision=0)
    b = Session(cs, precision=1)
    a.step(200)
    b.step(200)
    ra, rb = a.get("U")[:, 0], b.get("U")[:, 0]
    l2 = relnorm(rb, ra)
    sa, sb = a.scalars(), b.scalars()
    dfx = abs(sb["FX1"] - sa["FX1"]) / abs(sa["FX1"])
    print(f"FP32 vs FP64 after 200 steps: L2(rho) {l2:.3e}, Fx rel {dfx:.3e}")
    assert np.all(np.isfinite(b.get("U")))
    assert l2 < 1e-3 and dfx < 1e-2


def test_learned_warm_start_cuts_mesh_solve_iterations():
    """Moving ellipse, relative PCG rule 1e-6: the learned warm start (the
    solution history extrapolated, plus the MLP's correction on the matrix
    cores, with fallback) needs far fewer mesh-solve iterations than the
    previous-step guess, starts
    from the learned guess on most solves, and the mesh agrees with the
    reference-guess run to the solve tolerance.  The forces of this pitching
    flow are chaotic at round-off level (tests/test_oracle_chaos.py shows it
    on the sequential oracle): the learned run's force difference is compared
    with the difference a one-ulp density perturbation of the plain run makes
    over the same steps."""
    from cfd_amd.solver import Session
    cs = ellipse_case(48, 96)
    nsteps = 30
    out = {}
    for name, ws, ulp in (("plain", 0, False), ("plain_ulp", 0, True), ("learned", 1, False)):
        S = Session(cs, pcg_rel=1, pcg_tol=1e-6, warm_start=ws)
        if ulp:
            U = S.get("U")
            U[5, 0] = np.nextafter(U[5, 0], 2.0)
            S.set("U", U)
            S.set("U1", U)
        f = np.empty((nsteps, 2))
        for k in range(nsteps):
            S.step(1)
            sc = S.scalars()
            f[k] = (sc["FX1"], sc["FY1"])
        out[name] = dict(sc=sc, X=S.get("X"), Y=S.get("Y"), f=f)
        S.close()
        print(f"{name}: PCG iterations {sc['PCG_ITSUM']:.0f}, learned "
              f"{sc['WS_LEARNED']:.0f}/{sc['WS_SOLVES']:.0f}, loss {sc['WS_LOSS']:.3e}")
    p, q = out["plain"]["sc"], out["learned"]["sc"]
    assert q["WS_SOLVES"] == 2 * (nsteps - 2)          # candidates from the third step on
    assert q["WS_LEARNED"] >= 0.8 * q["WS_SOLVES"]
    assert p["PCG_ITSUM"] > 50 and q["PCG_ITSUM"] < 0.5 * p["PCG_ITSUM"]
    assert np.isfinite(q["WS_LOSS"])
    # the rule bounds the residual, not the error, of an ill-conditioned
    # Laplacian: the two runs stop at different points within it
    X0, Y0 = out["plain"]["X"], out["plain"]["Y"]
    disp = np.abs(X0 - cs.mesh.x).max() + np.abs(Y0 - cs.mesh.y).max()
    dx = max(np.abs(out["learned"]["X"] - X0).max(), np.abs(out["learned"]["Y"] - Y0).max())
    print(f"mesh difference {dx:.3e} of displacement {disp:.3e}")
    assert disp > 0 and dx <= 1e-2 * disp
    fp = out["plain"]["f"]
    scale = np.abs(fp).max()
    floor = np.abs(out["plain_ulp"]["f"] - fp).max(axis=1) / scale
    diff = np.abs(out["learned"]["f"] - fp).max(axis=1) / scale
    print("one-ulp floor", ["%.1e" % v for v in floor[4::5]])
    print("learned diff ", ["%.1e" % v for v in diff[4::5]])
    assert floor[nsteps // 3:].max() > 1e-4           # 1e-4 force agreement is below round-off chaos
    assert diff.max() < 0.1                             # the same flow


def test_trained_network_gains_over_the_extrapolation_alone():
    """What the network adds: the same moving-ellipse run (FP32 elements)
    with the learned start (extrapolation + the network's Galerkin-scaled
    correction) and with the extrapolation alone (warm_start 2, no network),
    both under the same relative PCG rule, at 1e-7 and 1e-8, over 120 steps.
    From the fourth step on (the first with the network) the learned start
    needs clearly fewer mesh-solve iterations (measured, on two builds whose
    flows differ at round-off: 0.71-0.80 of the extrapolation's per rule,
    0.72-0.74 pooled), and both meshes agree with each other to the rule's
    accuracy.  The ratio is taken over a long run because over a few dozen
    steps it scatters (0.71-0.97 measured): DTMIN's 1 % hysteresis jumps,
    which follow the flow's round-off, set the extrapolation's error."""
    from cfd_amd.solver import Session
    cs = ellipse_case(48, 96)
    nsteps, n0 = 120, 3
    tot = {1: 0.0, 2: 0.0}
    for tol in (1e-7, 1e-8):
        out = {}
        for ws in (2, 1):
            S = Session(cs, precision=1, pcg_rel=1, pcg_tol=tol, warm_start=ws)
            S.step(n0)
            it0 = S.scalars()["PCG_ITSUM"]
            S.step(nsteps - n0)
            sc = S.scalars()
            out[ws] = dict(its=sc["PCG_ITSUM"] - it0, sc=sc, X=S.get("X"), Y=S.get("Y"))
            tot[ws] += out[ws]["its"]
            S.close()
        e, q = out[2], out[1]
        print(f"tol {tol:g}: iterations after step {n0}: extrapolation {e['its']:.0f}, "
              f"learned {q['its']:.0f} (loss {q['sc']['WS_LOSS']:.3e})")
        assert e["sc"]["WS_SOLVES"] == q["sc"]["WS_SOLVES"] == 2 * (nsteps - 2)
        assert q["its"] <= 0.9 * e["its"]
        assert np.isfinite(q["sc"]["WS_LOSS"]) and q["sc"]["WS_LOSS"] < 1.0
        disp = np.abs(e["X"] - cs.mesh.x).max() + np.abs(e["Y"] - cs.mesh.y).max()
        dx = max(np.abs(q["X"] - e["X"]).max(), np.abs(q["Y"] - e["Y"]).max())
        assert disp > 0 and dx <= 1e-2 * disp
    print(f"pooled ratio {tot[1] / tot[2]:.3f}")
    assert tot[1] <= 0.85 * tot[2]


def test_learned_warm_start_gate_under_the_absolute_rule():
    """Under the reference's absolute PCG rule most solves stop at their first
    test: after a step whose two solves both did, the learned candidate is
    skipped (the reference's guess is used), so fewer solves see a candidate
    than with the relative rule, and the run equals the reference-guess run's
    mesh to the rule's accuracy."""
    from cfd_amd.solver import Session
    cs = ellipse_case(48, 96)
    nsteps = 20
    out = {}
    for ws in (0, 1):
        S = Session(cs, pcg_rel=0, warm_start=ws)
        S.step(nsteps)
        out[ws] = dict(sc=S.scalars(), X=S.get("X"), Y=S.get("Y"))
        S.close()
    p, q = out[0]["sc"], out[1]["sc"]
    print(f"absolute rule: iterations {p['PCG_ITSUM']:.0f} / {q['PCG_ITSUM']:.0f}, candidates on "
          f"{q['WS_SOLVES']:.0f} of {2 * (nsteps - 2)} solves")
    assert q["WS_SOLVES"] < 2 * (nsteps - 2)
    assert np.isfinite(out[1]["X"]).all()
    # |r.z| < 1e-10 is absolute: two starts may stop a few 1e-6 apart on a
    # displacement of 2e-4 (measured 1.0e-2 of it)
    disp = np.abs(out[0]["X"] - cs.mesh.x).max()
    assert np.abs(out[1]["X"] - out[0]["X"]).max() <= 1e-1 * disp


def test_session_rejects_patch_shapes_beyond_lds():
    """A patch shape whose fused-kernel LDS exceeds 64 KB (FP64: 682
    elements need 65 KB of contribution slots alone) is refused when the
    session is created; bad thread counts too."""
    from cfd_amd._lib import CfdError
    from cfd_amd.solver import Session
    cs = cylinder_case(24, 48)
    with pytest.raises(CfdError):
        Session(cs, precision=0, patch_elems=682)
    with pytest.raises(CfdError):
        Session(cs, precision=1, patch_block=96)
    S = Session(cs, precision=1, patch_elems=682, patch_block=512)   # fits in binary32
    S.step(2)
    assert np.isfinite(S.get("U")).all()
    S.close()


def test_session_rejects_bad_modes():
    """Out-of-range precision and warm-start modes are refused at creation."""
    from cfd_amd._lib import CfdError
    from cfd_amd.solver import Session
    cs = ellipse_case(8, 24)
    for kw in ({"precision": 3}, {"precision": -1}, {"warm_start": 3}, {"warm_start": -1}):
        with pytest.raises(CfdError):
            Session(cs, **kw)
    for ws in (1, 2):   # the learned start and its ablation both run
        S = Session(cs, pcg_rel=1, pcg_tol=1e-8, warm_start=ws)
        S.step(4)
        sc = S.scalars()
        assert np.isfinite(S.get("X")).all() and sc["WS_SOLVES"] == 4
        S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_renumbered_session_fixed_mesh_bitwise(cyl, precision):
    """A session that numbers nodes and elements in patch order inside
    (renumber = 1) steps a fixed mesh bit for bit as the session in the
    caller's numbering, with every field read and written in the caller's
    numbering (node, element and Laplace fields; a set-get round trip)."""
    from cfd_amd.solver import Session
    cs, o0 = cyl
    st = o0.state()
    A = Session(cs, precision=precision)
    B = Session(cs, precision=precision, renumber=1)
    for S in (A, B):
        S.set_state(st)
        S.set_scalars(DISN2=o0.disn2)
    for name in ("U", "T", "DNX", "AREA", "LAP_SPARSE", "X"):
        assert np.array_equal(B.get(name), np.asarray(st[name]).reshape(B.get(name).shape)), name
    A.step(6)
    B.step(6)
    for name in ("U", "U1", "RHS", "T", "VEL_X", "P", "RMACH", "DTL", "DT", "SHOC", "T_SUGN2"):
        assert np.array_equal(A.get(name), B.get(name)), name
    sa, sb = A.scalars(), B.scalars()
    for key in ("DTMIN", "TIME", "FX1", "FY1"):
        assert sa[key] == sb[key], key
    A.close()
    B.close()
    # from the free stream: the masses (MASAS) sum in the caller's element order
    A = Session(cs, precision=precision)
    B = Session(cs, precision=precision, renumber=1)
    assert np.array_equal(A.get("M"), B.get("M"))
    A.step(8)
    B.step(8)
    assert np.array_equal(A.get("U"), B.get("U"))
    A.close()
    B.close()


def test_renumbered_session_moving_mesh_teacher_forced(ell):
    """The pitching ellipse with the patch-order numbering inside, each step
    from the oracle's state: within 1e-12 as the caller-numbered session (its
    mesh solves sum their dot products in another row order).  Free runs of
    the two numberings part at round-off and then, through ESTAB's
    rounding-residue direction on the moving mesh, grow apart like any two
    differently rounded runs (1e-4 after 5 steps here)."""
    from cfd_amd.solver import Session
    cs, o0 = ell
    o = OracleSolver(cs)
    o.set_state(o0.state())
    o.disn2 = o0.disn2
    S = Session(cs, precision=0, renumber=1)
    for k in range(3):
        S.set_state(o.state())
        S.set_scalars(DISN2=o.disn2)
        o.step()
        S.step(1)
        ref, got = o.state(), S.state()
        assert got["DTMIN"] == ref["DTMIN"]
        assert np.all(relnorm_cols(got["U"], ref["U"]) < TOL), (k, relnorm_cols(got["U"], ref["U"]))
        for f in ("T", "VEL_X", "VEL_Y", "P", "X", "Y", "M", "LAP_SPARSE"):
            assert relnorm(got[f], ref[f]) < TOL, (k, f)
        assert relnorm(got["XPOS"], ref["XPOS"]) < 1e-6 or np.abs(ref["XPOS"]).max() == 0
    assert S.scalars()["PCG_ITSUM"] > 0
    S.close()


def test_renumber_needs_a_fused_path():
    from cfd_amd._lib import CfdError
    from cfd_amd.solver import Session
    with pytest.raises(CfdError):
        Session(cylinder_case(12, 24), precision=2, renumber=1)
