"""CPU checks of the learned proposal's references: the hand-written numpy
backpropagation against finite differences, and the bfloat16 rounding."""
import numpy as np

from sdl_amd import proposal as P


def test_param_layout():
    assert P.N_PARAMS == 24 * 64 + 64 + 2 * (64 * 64 + 64) + 64 + 1 == 9985


def test_numpy_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    flat = P.init_params(seed=5).astype(np.float64)
    X = rng.standard_normal((40, P.N_IN))
    Y = (rng.random(40) < 0.3).astype(np.float64)
    loss, g = P.loss_grad_np(flat, X, Y, pos_weight=3.0)
    idx = rng.choice(P.N_PARAMS, 60, replace=False)
    eps = 1e-6
    for k in idx:
        up, dn = flat.copy(), flat.copy()
        up[k] += eps
        dn[k] -= eps
        num = (P.loss_grad_np(up, X, Y, 3.0)[0] - P.loss_grad_np(dn, X, Y, 3.0)[0]) / (2 * eps)
        assert abs(num - g[k]) <= 1e-6 + 1e-4 * abs(num), (k, num, g[k])


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e38, 1e-40], np.float32)
    r = P.round_bf16(vals)
    # 1 + 2^-8 is exactly halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)
    assert r[3] == -2.5
    assert np.isfinite(r[4])
