"""CPU oracle of the learned kappa-bracket network -- TEST INFRASTRUCTURE ONLY.

A numpy restatement (float64) of the 32-64-64-2 GELU MLP used by the GPU
library's hand-written kernels (imcom_amd/csrc/mlp.hip), with the same
parameter layout:

  W1[32][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2] | xmean[32] xstd[32] ymean ystd

forward:  x = (X - xmean)/xstd ; h1 = gelu(x W1 + b1) ; h2 = gelu(h1 W2 + b2) ;
          (mu, s) = h2 W3 + b3
loss:     mean over samples of 0.5 ((t - mu)/sigma)^2 + s,  sigma = exp(s),
          t = (y - ymean)/ystd, with s clamped to [-10, 5] (zero gradient
          outside)
gelu:     0.5 x (1 + erf(x / sqrt 2))

Pinned by central finite differences of the loss (tests/test_mlp_oracle.py).
There is no reference counterpart: the reference always bisects the full
kappa range (imcom_bisect.f90:174-267).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NF, NH, NO = 32, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NF * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NO
N_TRAIN = OFF_B3 + NO
OFF_XM = N_TRAIN
OFF_XS = OFF_XM + NF
OFF_YM = OFF_XS + NF
OFF_YS = OFF_YM + 1
N_PARAMS = OFF_YS + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(NF, NH), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(NH, NH), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3].reshape(NH, NO), b3=p[OFF_B3:N_TRAIN],
                xm=p[OFF_XM:OFF_XS], xs=p[OFF_XS:OFF_YM], ym=p[OFF_YM], ys=p[OFF_YS])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def forward(p, X):
    q = unpack(p)
    x = (np.asarray(X, dtype=np.float64) - q["xm"]) / q["xs"]
    z1 = x @ q["W1"] + q["b1"]
    h1 = gelu(z1)
    z2 = h1 @ q["W2"] + q["b2"]
    h2 = gelu(z2)
    y = h2 @ q["W3"] + q["b3"]
    return y, (x, z1, h1, z2, h2)


def loss(p, X, y):
    q = unpack(p)
    out, _ = forward(p, X)
    t = (np.asarray(y, dtype=np.float64) - q["ym"]) / q["ys"]
    s = np.clip(out[:, 1], -10.0, 5.0)
    r = (t - out[:, 0]) * np.exp(-s)
    return float(np.mean(0.5 * r * r + s))


def grad(p, X, y):
    """Gradient of the mean loss with respect to the trainable block
    (N_TRAIN entries, parameter layout)."""
    q = unpack(p)
    out, (x, z1, h1, z2, h2) = forward(p, X)
    B = x.shape[0]
    t = (np.asarray(y, dtype=np.float64) - q["ym"]) / q["ys"]
    y1 = out[:, 1]
    s = np.clip(y1, -10.0, 5.0)
    inv = np.exp(-s)
    r = (t - out[:, 0]) * inv
    g0 = -r * inv / B
    g1 = np.where((y1 > -10.0) & (y1 < 5.0), (1.0 - r * r) / B, 0.0)
    dY = np.stack([g0, g1], axis=1)
    dZ2 = (dY @ q["W3"].T) * gelu_grad(z2)
    dZ1 = (dZ2 @ q["W2"].T) * gelu_grad(z1)
    g = np.zeros(N_TRAIN)
    g[OFF_W1:OFF_B1] = (x.T @ dZ1).ravel()
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    g[OFF_W2:OFF_B2] = (h1.T @ dZ2).ravel()
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    g[OFF_W3:OFF_B3] = (h2.T @ dY).ravel()
    g[OFF_B3:N_TRAIN] = dY.sum(0)
    return g


SIGMA_FLOOR = 0.1  # decades; the bracket's sigma is sqrt(sigma^2 + floor^2)


def bracket(p, X, zsig=3.0):
    q = unpack(p)
    out, _ = forward(p, X)
    mu = q["ym"] + q["ys"] * out[:, 0]
    sn = q["ys"] * np.exp(np.clip(out[:, 1], -10.0, 5.0))
    sg = np.sqrt(sn * sn + SIGMA_FLOOR ** 2)
    return np.stack([mu - zsig * sg, mu + zsig * sg], axis=1)
