"""The oracle (C restatement) reproduces the compiled reference bit for bit on
the committed golden vectors (tests/golden/*.npz, written from the reference
by tests/golden/make_golden.py): neighbour tables, start-up geometry, three
full time steps, and each hot-path routine on a developed state."""
import os

import numpy as np
import pytest

from cfd_amd.mesh import cylinder_case, ellipse_case
from oracle import oracle as orc
from tests.helpers import bitdiff

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {
    "cyl_fixed": (cylinder_case, dict(nr=12, nt=24, mach=0.8, re=1000.0), True),
    "ell_moving": (ellipse_case, dict(nr=12, nt=24, mach=0.5, re=5000.0), True),
    "cyl_slip": (cylinder_case, dict(nr=12, nt=24, mach=0.5, re=500.0, slip_body=True), False),
}
STATE_KEYS = ["U", "RHS", "T", "VEL_X", "VEL_Y", "RHO", "E", "P", "X", "Y", "XPOS", "YPOS",
              "W_X", "W_Y", "M", "SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3", "DT", "AREA", "DNX",
              "DNY", "LAP_SPARSE", "LAP_DIAG", "DTL", "GAMM", "U1", "RMACH"]
SCALAR_KEYS = ["DTMIN", "TIME", "DTMIN1", "BANDERA", "ITER", "HMIN", "FX1", "FY1", "RM1"]


def _load(name):
    path = os.path.join(GOLDEN, f"{name}.npz")
    return dict(np.load(path, allow_pickle=False))


def _case(name, g):
    builder, kw, _ = CASES[name]
    cs = builder(**kw)
    assert np.array_equal(cs.mesh.inpoel, g["inpoel"])
    cs.mesh.x[:] = g["x"]
    cs.mesh.y[:] = g["y"]
    return cs


def _state(g, prefix):
    st = {k: g[f"{prefix}/{k}"] for k in STATE_KEYS}
    st.update(dict(zip(SCALAR_KEYS, g[f"{prefix}/scalars"].tolist())))
    return st


def _assert_state(o, g, prefix, keys=STATE_KEYS):
    for k in keys:
        assert bitdiff(o.s[k], g[f"{prefix}/{k}"]) == 0, f"{prefix}:{k}"


@pytest.mark.parametrize("name", list(CASES))
def test_tables_and_startup(name):
    g = _load(name)
    cs = _case(name, g)
    o = orc.OracleSolver(cs)
    for k in ("esup1", "esup2", "psup1", "psup2", "lap_rowptr", "lap_idx"):
        assert np.array_equal(o.tab[k], g[f"tables/{k}"]), k
    _assert_state(o, g, "init", ["U", "T", "AREA", "DNX", "DNY", "M", "LAP_SPARSE", "LAP_DIAG"])


@pytest.mark.parametrize("name", list(CASES))
def test_three_steps_bitwise(name):
    g = _load(name)
    cs = _case(name, g)
    o = orc.OracleSolver(cs)
    for i in range(1, 4):
        o.step()
        _assert_state(o, g, f"step{i}")
        sc = dict(zip(SCALAR_KEYS, g[f"step{i}/scalars"].tolist()))
        assert o.dtmin == sc["DTMIN"] and o.time == sc["TIME"]
        assert (o.fx, o.fy, o.rm) == (sc["FX1"], sc["FY1"], sc["RM1"])


@pytest.mark.parametrize("name", [n for n, c in CASES.items() if c[2]])
def test_units_bitwise(name):
    g = _load(name)
    cs = _case(name, g)
    o = orc.OracleSolver(cs)
    base = _state(g, "unit/base")

    def reset():
        o.set_state(base)

    reset()
    rhs = o.calcrhs(rhs=g["unit/calcrhs/RHS_in"].copy())
    assert bitdiff(rhs, g["unit/calcrhs/RHS"]) == 0
    reset()
    o.s["W_X"][:], o.s["W_Y"][:] = g["unit/fuente/W_X"], g["unit/fuente/W_Y"]
    assert bitdiff(o.fuente(), g["unit/fuente/RHS"]) == 0
    reset()
    dtmin = o.deltat()
    assert dtmin == float(g["unit/deltat/DTMIN"])
    assert bitdiff(o.s["DT"], g["unit/deltat/DT"]) == 0
    reset()
    o.estab(base["DTMIN"])
    for k in ("SHOC", "T_SUGN1", "T_SUGN2", "T_SUGN3"):
        assert bitdiff(o.s[k], g[f"unit/estab/{k}"]) == 0, k
    reset()
    o.rk(base["DTMIN"])
    for k in ("U1", "RHS", "T", "VEL_X", "VEL_Y", "RHO", "E", "P", "RMACH"):
        assert bitdiff(o.s[k], g[f"unit/rk/{k}"]) == 0, k
    reset()
    nn = len(cs.lists["ilaux"])
    x = np.zeros(cs.npoin)
    o.bicg(x, g["unit/bicg/POS_AUX"][:nn], cs.lists["ilaux"])
    assert bitdiff(x, g["unit/bicg/XPOS"]) == 0
    reset()
    for k in ("W_X", "W_Y", "W_X_OLD", "W_Y_OLD", "AREA_OLD"):
        o.s[k][:] = g[f"unit/gcl/{k}"]
    o.gcl(1.0e-3, 0)
    assert bitdiff(o.s["M"], g["unit/gcl/M"]) == 0
