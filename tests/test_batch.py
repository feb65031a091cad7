"""Batches of independent fixed-mesh cases (cfd_batch): every case of a
batch agrees bit for bit with the same case run alone in a Session of the
same precision and patch shape; the recorded-graph replay agrees with plain
launches on every replay; the autotuner keeps both properties.  The sweep
generator's draws (CPU)."""
import numpy as np
import pytest

from cfd_amd.mesh import acoustic_fsafe, cylinder_case, sweep_cases


def _cases():
    # uneven sizes, Mach numbers, angles of attack and Reynolds numbers, each
    # at an acoustic CFL of 0.5 (stable on the oracle for 300 steps)
    return [cylinder_case(n, m, mach=M, re=re, alpha_deg=a, fsafe=acoustic_fsafe(M))
            for n, m, M, re, a in ((12, 40, 0.3, 100.0, 0.0), (20, 64, 0.8, 1000.0, 3.0),
                                   (9, 30, 0.55, 60.0, -4.5), (16, 96, 0.7, 3000.0, 1.0))]


def test_sweep_cases_draws():
    cs = sweep_cases(40)
    assert len(cs) == 40
    for c in cs:
        assert 32 <= c.mesh.nr <= 256 and 64 <= c.mesh.nt <= 512
        assert 0.2 <= c.mach <= 0.85 and -5.0 <= c.alpha_deg <= 5.0 and 50.0 <= c.re <= 5000.0
        assert c.npoin == c.mesh.nr * c.mesh.nt and c.nelem == 2 * (c.mesh.nr - 1) * c.mesh.nt
    again = sweep_cases(3)
    assert [(c.mesh.nr, c.mesh.nt, c.mach) for c in again] == \
           [(c.mesh.nr, c.mesh.nt, c.mach) for c in cs[:3]]           # seeded
    assert len({c.mesh.nr for c in cs}) > 10 and len({round(c.re) for c in cs}) > 10
    # the free stream carries the angle of attack
    c = cs[0]
    assert np.isclose(np.degrees(np.arctan2(c.v_inf, c.u_inf)), c.alpha_deg)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 0])
def test_batch_matches_single_sessions_bitwise(precision):
    from cfd_amd.batch import Batch
    from cfd_amd.solver import Session
    cases = _cases()
    nsteps = 12
    B = Batch(cases, precision=precision)
    B.step(nsteps)
    B.sync()
    for k, cs in enumerate(cases):
        S = Session(cs, precision=precision)
        S.step(nsteps)
        assert np.isfinite(S.get("U")).all()
        for name in ("U", "T", "P", "RMACH", "DTL"):
            a, b = B.get(k, name), S.get(name)
            assert a.shape == b.shape and np.array_equal(a, b), (k, name)
        sb, ss = B.scalars(k), S.scalars()
        for key in ("DTMIN", "TIME", "DTMIN1", "BANDERA", "ITER"):
            assert sb[key] == ss[key], (k, key)
        assert np.array_equal(B.get(k, "SHOC"), S.get("SHOC"))
        S.close()
    # the cases differ (own free stream and step control)
    assert len({B.scalars(k)["DTMIN"] for k in range(len(cases))}) == len(cases)
    B.close()


@pytest.mark.gpu
def test_batch_graph_replays_match_plain_launches():
    """Each replay of the recorded step against the same step launched
    directly, with host reads and synchronisations between the replays."""
    from cfd_amd.batch import Batch
    cases = _cases()
    G = Batch(cases, precision=1)
    P = Batch(cases, precision=1)
    P.use_graph(False)
    for step in range(7):
        G.step(1)
        P.step(1)
        for k in range(len(cases)):
            assert np.array_equal(G.get(k, "U"), P.get(k, "U")), (step, k)
            assert G.scalars(k) == P.scalars(k), (step, k)
    G.close()
    P.close()


@pytest.mark.gpu
def test_batch_autotune_picks_a_fitting_shape_and_keeps_results():
    from cfd_amd.batch import Batch
    from cfd_amd.solver import Session
    cases = _cases()
    B = Batch(cases, precision=1)
    cand = [(512, 256), (256, 128), (384, 192), (682, 512)]
    best, times = B.autotune(cand, steps=3)
    print("autotune", times, "->", best)
    assert best in cand and times[best] is not None
    assert all(t is None or t > 0 for t in times.values())
    info = B.info()
    assert (info["patch_elems"], info["patch_block"]) == best
    B.step(5)
    S = Session(cases[1], precision=1, patch_elems=best[0], patch_block=best[1])
    S.step(5)
    assert np.array_equal(B.get(1, "U"), S.get("U"))
    S.close()
    B.close()


@pytest.mark.gpu
def test_batch_rejects_moving_and_bad_shapes():
    from cfd_amd._lib import CfdError
    from cfd_amd.batch import Batch
    from cfd_amd.mesh import ellipse_case
    with pytest.raises(ValueError):
        Batch([ellipse_case(8, 24)])
    with pytest.raises(CfdError):
        Batch(_cases()[:1], patch_elems=700)
    with pytest.raises(CfdError):
        Batch(_cases()[:1], patch_block=100)
