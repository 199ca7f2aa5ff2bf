"""Pins the numpy MLP oracle (oracle/mlp_oracle.py) by central finite
differences of its loss, and checks the parameter layout shared with the GPU
kernels.  CPU only."""
import numpy as np

from imcom_amd import kappa_net
from oracle import mlp_oracle as M


def _setup(seed=0, B=37):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, 32)) * 3 + 1
    y = rng.standard_normal(B) * 2 - 5
    p = kappa_net.init_params(X.astype(np.float32), y.astype(np.float32), seed).astype(np.float64)
    # make the output layer non-trivial
    p[M.OFF_W3:M.N_TRAIN] += rng.standard_normal(M.N_TRAIN - M.OFF_W3) * 0.3
    return p, X, y


def test_layout():
    assert M.N_TRAIN == kappa_net.N_TRAIN == 6402
    assert M.N_PARAMS == kappa_net.N_PARAMS == 6402 + 66


def test_gradient_finite_differences():
    p, X, y = _setup()
    g = M.grad(p, X, y)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(M.N_TRAIN, 60, replace=False),
                          np.arange(M.OFF_B3, M.N_TRAIN), [M.OFF_W1, M.OFF_B1, M.OFF_W2]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(pp, X, y) - M.loss(pm, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_gelu_gradient():
    x = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(x))) < 1e-8


def test_bracket_contains_mean():
    p, X, y = _setup(1)
    br = M.bracket(p, X, 3.0)
    out, _ = M.forward(p, X)
    q = M.unpack(p)
    mu = q["ym"] + q["ys"] * out[:, 0]
    assert np.all(br[:, 0] < mu) and np.all(mu < br[:, 1])


def test_bracket_width_floor():
    """The bracket is never narrower than 2 z sigma_floor, however confident
    the network is (a very negative log sigma output)."""
    p, X, y = _setup(2)
    p2 = p.copy()
    p2[M.OFF_B3 + 1] = -50.0  # log sigma bias: sigma -> 0
    br = M.bracket(p2, X, 3.0)
    assert np.all(br[:, 1] - br[:, 0] >= 2 * 3.0 * M.SIGMA_FLOOR * (1 - 1e-9))
