"""Sensitivity of the reference's time step to round-off, on the sequential
oracle (bit-exact with the compiled reference): a one-ulp change of one
node's density, or of one node's coordinate, moves the pitching ellipse's
pressure forces by more than 1e-4 (relative) within 30 steps.  The flow
trajectory is chaotic at round-off level (ESTAB's shock-capturing direction
is built from a velocity "gradient" that is a rounding residue of
sum(vel * dN), subrutinas.f90:378-383), so two runs that round differently --
another summation order,
FP32 elements, another mesh-solve start -- cannot agree on the forces to
1e-4 after tens of steps; the agreement a faster path can show is bounded
below by this floor."""
import numpy as np

from cfd_amd.mesh import ellipse_case
from oracle import oracle as orc


def _forces(cs, nsteps, perturb=None):
    orc.set_threads(1)
    o = orc.OracleSolver(cs)
    if perturb is not None:
        perturb(o.s)
    f = np.empty((nsteps, 2))
    for k in range(nsteps):
        o.step()
        f[k] = (o.fx, o.fy)
    return f


def test_one_ulp_perturbation_moves_forces_beyond_1e4():
    cs = ellipse_case(48, 96)
    n = 30
    base = _forces(cs, n)
    again = _forces(cs, n)
    assert np.array_equal(base, again)           # the sequential oracle is deterministic

    def ulp_rho(s):
        s["U"][5, 0] = np.nextafter(s["U"][5, 0], 2.0)

    def ulp_x(s):
        s["X"][200] = np.nextafter(s["X"][200], 1e9)

    scale = np.abs(base).max()
    for pert in (ulp_rho, ulp_x):
        f = _forces(cs, n, pert)
        d = np.abs(f - base).max(axis=1) / scale
        print(pert.__name__, ["%.1e" % v for v in d[4::5]])
        assert d[0] < 1e-12                      # the first step sees the perturbation at round-off
        assert d[n // 3:].max() > 1e-4           # ... and tens of steps amplify it past 1e-4
        assert d.max() < 0.2                     # the same flow, not a blow-up
