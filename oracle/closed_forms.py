"""Closed forms of the lookup results for Gaussian PSFs (test infrastructure:
imported by tests only, never by the product path).

For unit-height Gaussian PSF images of width sigma (input px) sampled at
psfscale, the A table is the autocorrelation psfscale^2/(4 pi sigma^2)
exp(-|D|^2 / (4 sigma^2)) and the B table is -2 x the Gamma-G
cross-correlation -2 psfscale^2/(2 pi v) exp(-|D|^2 / (2 v)), v = sigma^2 +
sigma_Gamma^2 (tests/test_oracle.py::test_tables_gaussian_closed_form).  An
exposure whose PSF is displaced by s shifts its argument: A_ij depends on
x_j - x_i - (s_j - s_i) and B_ia on x_i - X_a - s_i, the signs the overlap
integrals A_ij = sum_y G_i(x_i - y) G_j(x_j - y) and
B_ia = -2 sum_y Gamma(X_a - y) G_i(x_i - y) fix for an image G(x - y0) of a
point source at y0."""
import numpy as np


def A_closed_form(tile, shifts, sigma, psfscale):
    """A (n x n) for exposure e's PSF displaced by shifts[e - 1]."""
    s = np.asarray(shifts, dtype=np.float64)[np.asarray(tile.exp_i) - 1]
    dx = (tile.x_i[None, :] - s[None, :, 0]) - (tile.x_i[:, None] - s[:, None, 0])
    dy = (tile.y_i[None, :] - s[None, :, 1]) - (tile.y_i[:, None] - s[:, None, 1])
    return psfscale ** 2 / (4 * np.pi * sigma ** 2) * np.exp(-(dx ** 2 + dy ** 2) / (4 * sigma ** 2))


def B_closed_form(tile, shifts, sigma, gamma_sigma, psfscale):
    """B (n x m) for exposure e's PSF displaced by shifts[e - 1]."""
    s = np.asarray(shifts, dtype=np.float64)[np.asarray(tile.exp_i) - 1]
    v = sigma ** 2 + gamma_sigma ** 2
    dx = tile.x_i[:, None] - s[:, 0, None] - tile.X_a[None, :]
    dy = tile.y_i[:, None] - s[:, 1, None] - tile.Y_a[None, :]
    return -2 * psfscale ** 2 / (2 * np.pi * v) * np.exp(-(dx ** 2 + dy ** 2) / (2 * v))
