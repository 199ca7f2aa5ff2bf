"""Pins of the lookup-table oracle (oracle/tables.py) on the CPU: the DFT
sums written out with the reference's index ranges on tiny grids, Hermitian
symmetry, the Gaussian closed forms, and agreement with the host set-up."""
import cmath
import math

import numpy as np
import pytest

from imcom_amd import tile as TL
from oracle import oracle as O
from oracle import tables as OT


def _cshift_loops(a, shift, dim):
    """Fortran CSHIFT by explicit 1-based index arithmetic."""
    n1, n2 = a.shape
    out = np.empty_like(a)
    for i in range(1, n1 + 1):
        for j in range(1, n2 + 1):
            if dim == 1:
                out[i - 1, j - 1] = a[(i - 1 + shift) % n1, j - 1]
            else:
                out[i - 1, j - 1] = a[i - 1, (j - 1 + shift) % n2]
    return out


def _brute_table(prod, npad):
    """imcom_build_Alookup's steps for one table, with loops and the DFT sum
    (imcom_matrix.f90:397-414), before the division by n1psf n2psf."""
    n1psf, n2psf = prod.shape
    N1, N2 = n1psf * npad, n2psf * npad
    n1min, n2min = (N1 - n1psf) // 2 + 1, (N2 - n2psf) // 2 + 1
    inner = _cshift_loops(_cshift_loops(prod, n2psf // 2, 2), n1psf // 2, 1)
    u = np.zeros((N1, N2), dtype=complex)
    for i in range(n1psf):
        for j in range(n2psf):
            u[n1min - 1 + i, n2min - 1 + j] = inner[i, j]
    u = _cshift_loops(_cshift_loops(u, N2 // 2, 2), N1 // 2, 1)
    tmp = np.zeros((N1, N2))
    for x1 in range(N1):
        for x2 in range(N2):
            s = 0j
            for k1 in range(N1):
                for k2 in range(N2):
                    if u[k1, k2] != 0:
                        s += u[k1, k2] * cmath.exp(2j * math.pi * (k1 * x1 / N1 + k2 * x2 / N2))
            tmp[x1, x2] = s.real
    return _cshift_loops(_cshift_loops(tmp, N2 // 2, 2), N1 // 2, 1)


@pytest.mark.parametrize("shape", [(6, 6), (5, 7), (8, 4)])
def test_ft_r2c_is_the_dft(shape):
    rng = np.random.default_rng(3)
    F = rng.standard_normal(shape)
    n1, n2 = shape
    ref = np.zeros(shape, dtype=complex)
    for k1 in range(n1):
        for k2 in range(n2):
            ref[k1, k2] = sum(F[x1, x2] * cmath.exp(-2j * math.pi * (k1 * x1 / n1 + k2 * x2 / n2))
                              for x1 in range(n1) for x2 in range(n2))
    got = OT.ft_r2c(F)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    # the completed half is exactly the conjugate of the computed half
    for r in range(n1 // 2 + 1, n1):
        for j in range(n2):
            assert got[r, j] == np.conj(got[n1 - r, (n2 - j) % n2])


@pytest.mark.parametrize("n,npad", [(4, 1), (4, 3), (6, 2)])
def test_tables_match_written_out_sums(n, npad):
    rng = np.random.default_rng(n * 10 + npad)
    G1, G2, Gam = (OT.ft_r2c(rng.standard_normal((n, n))) for _ in range(3))
    A = OT.build_Alookup([G1, G2], npad)
    B = OT.build_Blookup(Gam, [G1, G2], npad)
    pairs = {0: np.conj(G1) * G1, 1: np.conj(G1) * G2, 2: np.conj(G2) * G2}  # ial-1
    for ial, prod in pairs.items():
        ref = _brute_table(prod, npad) / n / n
        assert np.max(np.abs(A[:, :, ial] - ref)) <= 1e-12 * np.max(np.abs(ref))
    for e, G in enumerate([G1, G2]):
        ref = -2.0 * _brute_table(np.conj(Gam) * G, npad) / n / n
        assert np.max(np.abs(B[:, :, e] - ref)) <= 1e-12 * np.max(np.abs(ref))


def _gauss_G(npix, sigma, psfscale):
    c = npix // 2
    s = sigma / psfscale
    i = np.arange(npix) - c
    return np.asfortranarray(np.exp(-0.5 * (i[:, None] ** 2 + i[None, :] ** 2) / s ** 2))


def test_oracle_tables_gaussian_closed_form_and_host_setup():
    ps, sg, sG = 0.25, 1.2, 1.5
    n1pad = 256
    G = TL.embed_psf(_gauss_G(128, sg, ps), n1pad)
    Gam = TL.embed_psf(_gauss_G(128, sG, ps), n1pad)
    Gt, Gamt = OT.ft_r2c(G), OT.ft_r2c(Gam)
    assert np.max(np.abs(Gt - TL.ft_r2c(G))) <= 1e-12 * np.max(np.abs(Gt))
    Al = OT.build_Alookup([Gt], 3)
    Bl = OT.build_Blookup(Gamt, [Gt], 3)
    assert np.max(np.abs(Al - TL.build_Alookup([Gt], 3))) <= 1e-14 * np.max(np.abs(Al))
    assert np.max(np.abs(Bl - TL.build_Blookup(Gamt, [Gt], 3))) <= 1e-14 * np.max(np.abs(Bl))
    for D in [(0.0, 0.0), (0.37, -1.21), (2.5, 3.1)]:
        xa, ya = 3 * D[0] / ps, 3 * D[1] / ps
        r2 = D[0] ** 2 + D[1] ** 2
        a_ref = ps ** 2 / (4 * np.pi * sg ** 2) * np.exp(-r2 / (4 * sg ** 2))
        assert abs(O.interp_lookup(Al[:, :, 0], xa, ya) - a_ref) <= 1e-9 * a_ref
        v = sg ** 2 + sG ** 2
        b_ref = -2 * ps ** 2 / (2 * np.pi * v) * np.exp(-r2 / (2 * v))
        assert abs(O.interp_lookup(Bl[:, :, 0], xa, ya) - b_ref) <= 1e-9 * abs(b_ref)
    C = OT.calc_C(Gamt)
    assert abs(C - ps ** 2 / (4 * np.pi * sG ** 2)) <= 1e-12 * C
    assert C == TL.calc_C(Gamt)


# ---------------------------------------------------------------- rotation
def test_rotate_identity_and_quarter_turn():
    rng = np.random.default_rng(11)
    im = rng.standard_normal((17, 17))
    assert np.array_equal(OT.rotate_image(im, 0.0, 7, 0), im)
    assert np.array_equal(OT.rotate_image(im, 360.0, 7, 0), im)
    # a quarter turn about ((n+1)/2, (n+1)/2): out(i, j) = in(j, n+1-i)
    q = OT.rotate_image(im, 90.0, 7, 0)
    assert np.max(np.abs(q - np.rot90(im))) <= 1e-12 * np.max(np.abs(im))


def test_rotate_reproduces_a_cubic_image():
    """Every output pixel whose rotated position keeps its floor/ceiling in
    the image equals the cubic at that position (the stencil holds >= 4
    nodes), and the others are 0."""
    n1, n2, th = 24, 20, 23.0

    def P(x, y):
        return 0.3 + 0.1 * x - 0.02 * y + 0.003 * x * x + 0.001 * x * y - 0.0004 * y ** 3

    i = np.arange(1, n1 + 1, dtype=float)[:, None]
    j = np.arange(1, n2 + 1, dtype=float)[None, :]
    im = np.asfortranarray(P(i, j) + 0 * j)
    out = OT.rotate_image(im, th, 7, 0)
    c, s = math.cos(math.radians(th)), math.sin(math.radians(th))
    xc, yc = (n1 + 1) / 2, (n2 + 1) / 2
    xs = (i - xc) * c + (j - yc) * s + xc
    ys = -(i - xc) * s + (j - yc) * c + yc
    inside = (np.floor(xs) >= 1) & (np.floor(ys) >= 1) & (np.ceil(xs) <= n1) & (np.ceil(ys) <= n2)
    assert inside.sum() > 200 and (~inside).sum() > 20
    assert np.max(np.abs(out[inside] - P(xs, ys)[inside])) <= 1e-11
    assert np.all(out[~inside] == 0.0)


def test_rotate_conserves_flux_and_matches_host_setup():
    c = 32
    k = np.arange(64) - c
    G = np.asfortranarray(np.exp(-0.5 * (k[:, None] ** 2 / 5.0 ** 2 + k[None, :] ** 2 / 3.0 ** 2)))
    out = OT.rotate_image(G, 37.5, 7, 1)
    assert abs(out.sum() - G.sum()) <= 1e-12 * G.sum()
    host = TL.rotate_image(G, 37.5, 7, 1)
    assert np.max(np.abs(out - host)) <= 1e-12 * np.max(np.abs(out))


# ---------------------------------------------------------- input vectors
def test_assemble_inputs_hand_worked():
    """3 x 2 pixels x 2 exposures, saturation 5: the values 5 (not < 5) and 9
    are masked; x = (i1-1) * scale + dither in 1-based terms."""
    Im = np.zeros((3, 2, 2), order="F")
    Im[:, :, 0] = [[1.0, 5.0], [2.0, 3.0], [9.0, 4.0]]
    Im[:, :, 1] = [[0.5, 0.25], [6.0, -1.0], [4.999, 7.0]]
    sx, sy = [1.0, 2.0], [1.0, 0.5]
    dither = [[0.1, 0.2], [0.3, 0.4]]
    x, y, e, I = OT.assemble_inputs(Im, 5.0, sx, sy, dither)
    # column-major order: (1,1),(2,1),(3,1),(1,2),(2,2),(3,2) per exposure
    assert list(I) == [1.0, 2.0, 3.0, 4.0, 0.5, 4.999, 0.25, -1.0]
    assert list(e) == [1, 1, 1, 1, 2, 2, 2, 2]
    assert np.allclose(x, [0.1, 1.1, 1.1, 2.1, 0.3, 4.3, 0.3, 2.3], rtol=0, atol=1e-15)
    assert np.allclose(y, [0.2, 0.2, 1.2, 1.2, 0.4, 0.4, 0.9, 0.9], rtol=0, atol=1e-15)


def test_assemble_inputs_matches_host_setup():
    rng = np.random.default_rng(5)
    Im = np.asfortranarray(rng.uniform(0, 1, size=(7, 5, 3)))
    sx, sy = np.array([1.0, 0.9, 1.1]), np.array([1.0, 1.05, 0.95])
    dither = rng.uniform(0, 1, size=(3, 2))
    x, y, e, I = OT.assemble_inputs(Im, 0.8, sx, sy, dither)
    subs, I_h, e_h = TL.mask_pixels(Im, 0.8)
    x_h, y_h = TL.build_xy(7, 5, sx, sy, dither, subs)
    assert np.array_equal(I, I_h) and np.array_equal(e, e_h)
    assert np.array_equal(x, x_h) and np.array_equal(y, y_h)
