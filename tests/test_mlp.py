"""The warm-start MLP.

CPU: the oracle (oracle/mlp_oracle.py) is pinned -- its exact-mode gradient
equals central finite differences of its loss for every parameter, its BF16
rounding matches known cases (ties to even, carries into the exponent,
specials), and BF16 mode with no rounding reduces to exact mode.

GPU: the gfx950 matrix-core kernels (forward, back-propagation, weight and
bias gradients) agree with the oracle's BF16 mode, through the C ABI.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _params(rng, scale=0.4):
    p = rng.standard_normal(mo.NPARAMS) * scale
    p[mo.W3 + mo.USED * mo.HID:mo.B3] = 0.0        # unused output rows
    p[mo.B3 + mo.USED:] = 0.0
    return p.astype(np.float32)


def test_bf16_rounding_cases():
    cases = {1.0: 0x3F80, -2.0: 0xC000, 1.0 + 2 ** -8: 0x3F80, 1.0 + 3 * 2 ** -8: 0x3F82,
             1.0 + 2 ** -8 + 2 ** -20: 0x3F81, 255.99: 0x4380, 0.0: 0x0000}
    for v, bits in cases.items():
        assert int(mo.bf16_bits(np.array([v], np.float32))[0]) == bits, v
    specials = np.array([np.inf, -np.inf], np.float32)
    assert list(mo.bf16_bits(specials)) == [0x7F80, 0xFF80]
    assert mo.bf16_value(mo.bf16_bits(np.array([np.nan], np.float32)))[0] != 0.0


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(rng).astype(np.float64)
    x = rng.uniform(-1, 1, (7, mo.IN))
    t = rng.uniform(-1, 1, (7, 2))
    loss, g = mo.loss_and_grad(p, x, t, mode="exact")
    h = 1e-6
    fd = np.zeros_like(g)
    for k in range(mo.NPARAMS):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd[k] = (mo.loss_and_grad(pp, x, t, "exact")[0] - mo.loss_and_grad(pm, x, t, "exact")[0]) / (2 * h)
    used = np.ones(mo.NPARAMS, bool)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())
    assert np.abs(g[used]).max() > 1e-3                                  # not trivially zero
    # the bias gradients are the row sums of dZ: a check independent of the weights' path
    assert np.allclose(g[mo.B3:mo.B3 + 2], 2.0 * (mo.forward(p, x, "exact")[0] - t).sum(0) / 7)


def test_oracle_bf16_mode_bounds_exact_mode():
    rng = np.random.default_rng(4)
    p = _params(rng)
    x = mo.round_bf16(rng.uniform(-1, 1, (64, mo.IN)).astype(np.float32))
    t = rng.uniform(-1, 1, (64, 2))
    y_e, _ = mo.forward(p, x, "exact")
    y_b, _ = mo.forward(p, x, "bf16")
    rel = np.linalg.norm(y_b - y_e) / np.linalg.norm(y_e)
    assert 1e-5 < rel < 3e-2      # BF16 operands: a few parts in a thousand
    _, g_e = mo.loss_and_grad(p, x, t, "exact")
    _, g_b = mo.loss_and_grad(p, x, t, "bf16")
    assert np.linalg.norm(g_b - g_e) / np.linalg.norm(g_e) < 5e-2


@pytest.mark.gpu
@pytest.mark.parametrize("n", [16, 1000, 20000])
def test_gpu_mlp_forward_and_gradient(n):
    from cfd_amd.api import mlp_forward, mlp_gradient
    rng = np.random.default_rng(5 + n)
    p = _params(rng)
    xv = rng.uniform(-1, 1, (n, mo.IN)).astype(np.float32)
    xb = mo.bf16_bits(xv)
    x = mo.bf16_value(xb)
    t = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    print(f"MLP n={n}", flush=True)
    y_gpu = mlp_forward(p, xb)
    y_ref, _ = mo.forward(p, x, "bf16")
    assert np.linalg.norm(y_gpu - y_ref) / np.linalg.norm(y_ref) < 2e-3
    loss_gpu, g_gpu = mlp_gradient(p, xb, t)
    loss_ref, g_ref = mo.loss_and_grad(p, x, t, "bf16")
    assert loss_gpu == pytest.approx(loss_ref, rel=1e-4)
    for name, a, b in (("W1", mo.W1, mo.B1), ("b1", mo.B1, mo.W2), ("W2", mo.W2, mo.B2),
                       ("b2", mo.B2, mo.W3), ("W3", mo.W3, mo.B3), ("b3", mo.B3, mo.NPARAMS)):
        ga, gr = g_gpu[a:b], g_ref[a:b]
        assert np.linalg.norm(ga - gr) <= 5e-3 * np.linalg.norm(gr) + 1e-12, name
