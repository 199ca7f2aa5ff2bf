"""Device lookup-table set-up (ft_r2c, build_Alookup/Blookup, calc_C) against
the oracle (oracle/tables.py), and a tile solved from device-built tables
against the same tile from host tables."""
import numpy as np
import pytest

from imcom_amd import synth
from imcom_amd import tile as TL
from oracle import tables as OT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from imcom_amd.gpu import ImcomGPU
    with ImcomGPU(0) as ctx:
        yield ctx


def _gauss(npix, sigma, psfscale, shift=(0.0, 0.0)):
    c = npix // 2
    s = sigma / psfscale
    i = np.arange(npix) - c - shift[0] / psfscale
    j = np.arange(npix) - c - shift[1] / psfscale
    return np.asfortranarray(np.exp(-0.5 * (i[:, None] ** 2 + j[None, :] ** 2) / s ** 2))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("n", [16, 64, 256])
def test_ft_r2c_matches_oracle(g, n):
    print(f"[tables] ft_r2c n={n}", flush=True)
    rng = np.random.default_rng(n)
    imgs = np.asfortranarray(np.stack([TL.embed_psf(_gauss(n // 2, 1.3, 0.25), n),
                                       rng.standard_normal((n, n))], axis=2))
    got = g.ft_r2c(imgs)
    for k in range(2):
        ref = OT.ft_r2c(imgs[:, :, k])
        # FP64 DFT by matrix products: error ~ sqrt(n) eps of the largest term
        assert _rel(got[:, :, k], ref) <= 1e-13, k
        # the completed half is the exact conjugate, as imcom_ft_r2c writes it
        r = n // 2 + 1
        assert np.array_equal(got[r, 1:, k], np.conj(got[n - r, :0:-1, k]))


@pytest.mark.parametrize("n,npad,nexp", [(8, 3, 2), (64, 3, 2), (64, 2, 3), (256, 3, 1)])
def test_tables_match_oracle(g, n, npad, nexp):
    print(f"[tables] build n={n} npad={npad} nexp={nexp}", flush=True)
    shifts = [(0.0, 0.0), (0.6, -0.35), (-0.2, 0.9)][:nexp]
    Gt = [OT.ft_r2c(TL.embed_psf(_gauss(n // 2, 1.1 + 0.1 * e, 0.25, shifts[e]), n))
          for e in range(nexp)]
    Gam = OT.ft_r2c(TL.embed_psf(_gauss(n // 2, 1.5, 0.25), n))
    psfconst = 1 if nexp == 1 else 0
    A = g.build_Alookup(psfconst, Gt, npad, out=True)
    B = g.build_Blookup(psfconst, Gam, Gt, npad, out=True)
    Aref = OT.build_Alookup(Gt, npad)
    Bref = OT.build_Blookup(Gam, Gt, npad)
    assert A.shape == Aref.shape and B.shape == Bref.shape
    for k in range(A.shape[2]):
        assert _rel(A[:, :, k], Aref[:, :, k]) <= 1e-13, ("A", k)
    for k in range(B.shape[2]):
        assert _rel(B[:, :, k], Bref[:, :, k]) <= 1e-13, ("B", k)
    C = g.calc_C(Gam)
    assert abs(C - OT.calc_C(Gam)) <= 1e-14 * C


def test_bad_grids_are_rejected(g):
    from imcom_amd.gpu import ImcomError
    G = np.ones((6, 6), dtype=complex)
    with pytest.raises(ImcomError, match="even side"):
        g.build_Alookup(1, [np.ones((7, 7), dtype=complex)], 3)
    with pytest.raises(ImcomError, match="nexppsf"):
        g.build_Alookup(1, [G, G], 3)
    with pytest.raises(ImcomError, match="square"):
        g.ft_r2c(np.ones((6, 8)))


def test_tile_from_device_tables_matches_host_tables(g):
    """The small workload's set-up with the tables built on the device: the
    A and B matrices interpolated from them agree with those from the host
    tables, and so do the solve's kappa and U."""
    print("[tables] small tile, device vs host tables", flush=True)
    t = synth.small()
    g.set_tile(t, tables="host")
    A_h = g.lookup_A(t.npoly, t.npad, out=True)
    B_h = g.lookup_B(t.npoly, t.npad, out=True)
    g.eigen_A(out=False)
    g.build_P()
    K_h, U_h, _ = g.bisect_kappa()
    g.set_tile(t, tables="device")
    A_d = g.lookup_A(t.npoly, t.npad, out=True)
    B_d = g.lookup_B(t.npoly, t.npad, out=True)
    g.eigen_A(out=False)
    g.build_P()
    K_d, U_d, _ = g.bisect_kappa()
    assert _rel(A_d, A_h) <= 1e-12
    assert _rel(B_d, B_h) <= 1e-12
    # U within the bisection's own tolerance; kappa identical or one flip at
    # |U - U_max| < US_tol
    assert np.max(np.abs(U_d - U_h)) <= 1e-12 + t.US_tol
    same = K_d == K_h
    assert same.mean() >= 0.95


@pytest.mark.parametrize("shape,theta", [((64, 64), 37.5), ((48, 40), 200.0), ((256, 256), 123.4)])
def test_rotate_image_matches_oracle(g, shape, theta):
    print(f"[tables] rotate {shape} {theta}", flush=True)
    n1, n2 = shape
    i = np.arange(n1)[:, None] - n1 // 2
    j = np.arange(n2)[None, :] - n2 // 2
    im = np.asfortranarray(np.exp(-0.5 * (i ** 2 / 7.0 ** 2 + j ** 2 / 4.0 ** 2)))
    for cf in (0, 1):
        got = g.rotate_image(im, theta, 7, cf)
        ref = OT.rotate_image(im, theta, 7, cf)
        # Lagrange form on the device against Neville's tableau on the host
        assert _rel(got, ref) <= 1e-13, cf
        assert np.all(got[ref == 0.0] == 0.0)  # positions outside the image


@pytest.mark.parametrize("shape,sat", [((16, 16, 4), 0.9), ((64, 64, 4), 0.97), ((5, 3, 2), 2.0)])
def test_assemble_inputs_matches_oracle(g, shape, sat):
    """Integer/ordering work: bit-exact against the oracle."""
    print(f"[setup] assemble {shape} sat={sat}", flush=True)
    rng = np.random.default_rng(shape[0])
    Im = np.asfortranarray(rng.uniform(0, 1, size=shape))
    ne = shape[2]
    sx, sy = rng.uniform(0.9, 1.1, ne), rng.uniform(0.9, 1.1, ne)
    dither = rng.uniform(0, 1, size=(ne, 2))
    got = g.assemble_inputs(Im, sat, sx, sy, dither)
    ref = OT.assemble_inputs(Im, sat, sx, sy, dither)
    for name, a, b in zip(("x", "y", "exp", "I"), got, ref):
        assert a.shape == b.shape, name
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, (name, bad.size, int(bad[0]), a[bad[0]], b[bad[0]])
