"""Regular-grid (USERXY=0) lookups through the C ABI: with the grids declared
(imcom_gpu_set_grid), lookup_A and lookup_B interpolate each distinct pixel
separation once and write A and B from those values.

  lattice vs per-entry kernels   |dA| <= 1e-13 max|A|, |dB| <= 1e-13 max|B|
                                 (the lattice argument u g + doff and the
                                 reference's x_j - x_i differ by the rounding
                                 of the coordinates only); A exactly symmetric
  lattice vs oracle              the same bound against the oracle's Neville
                                 lookups (imcom_matrix.f90:206-357)
  pitches in ratio P/Q           exposures of different pitches and an output
                                 pitch of 0.3 input px take lattices with
                                 P, Q > 1
  no small ratio                 B falls back to the per-entry kernel
                                 (bitwise the same B)
  inconsistent grid              set_grid refuses a dither that does not give
                                 the coordinates
"""
import numpy as np
import pytest

from imcom_amd import synth
from imcom_amd.gpu import ImcomError
from imcom_amd.tile import assemble_tile
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _lookups(g, t):
    A = g.lookup_A(t.npoly, t.npad, out=True)
    B = g.lookup_B(t.npoly, t.npad, out=True)
    return A, B


def _both_paths(g, t):
    g.set_tile(t, grid=False)
    assert g.lattice_state() == (False, False)
    A0, B0 = _lookups(g, t)
    g.set_grid(**t.grid)
    state = g.lattice_state()
    A1, B1 = _lookups(g, t)
    return A0, B0, A1, B1, state


def _pitch_tile(gimxscale, gimyscale, outscale, nexp=2, npix=10, nout=8, seed=11):
    """A Gaussian-PSF tile whose exposures and output grid have the given
    pitches (input px), the output grid inside the footprint."""
    rng = np.random.default_rng(seed)
    dither = rng.uniform(0.0, 1.0, size=(nexp, 2))
    gx, gy = np.asarray(gimxscale, float), np.asarray(gimyscale, float)
    psf = synth.gaussian_psf_image(64, 1.2, 0.25)
    gam = synth.gaussian_psf_image(64, 1.5, 0.25)
    xc = 0.5 * (npix - 1) * gx.min() + dither[:, 0].mean()
    yc = 0.5 * (npix - 1) * gy.min() + dither[:, 1].mean()
    I = np.asfortranarray(rng.normal(0.0, 1.0, size=(npix, npix, nexp)))
    return assemble_tile(
        I_unmasked=I, gimxscale=gx, gimyscale=gy, dither=dither, psf_img=psf,
        rotangdeg=np.zeros(nexp), gamma_img=gam, psfxscale=0.25, psfyscale=0.25,
        noise=np.full(nexp, 1e-6), n1out=nout, n2out=nout, outxscale=outscale,
        outyscale=outscale, outpos1=xc - 0.5 * outscale * (nout - 1),
        outpos2=yc - 0.5 * outscale * (nout - 1), meta=dict(name="pitches"))


def _check(A0, B0, A1, B1, t, oracle=True):
    dA = np.abs(A1 - A0).max()
    dB = np.abs(B1 - B0).max()
    print(f"[lattice] n={t.n} m={t.m} max|dA| {dA:.3e} of {np.abs(A0).max():.3e}, "
          f"max|dB| {dB:.3e} of {np.abs(B0).max():.3e}", flush=True)
    assert dA <= 1e-13 * np.abs(A0).max()
    assert dB <= 1e-13 * np.abs(B0).max()
    assert np.array_equal(A1, A1.T)
    if oracle:
        Ao, Bo = O.lookup_A(t), O.lookup_B(t)
        assert np.abs(A1 - Ao).max() <= 1e-13 * np.abs(Ao).max()
        assert np.abs(B1 - Bo).max() <= 1e-13 * np.abs(Bo).max()


@pytest.mark.parametrize("which", ["tiny", "small", "displaced"])
def test_lattice_matches_per_entry_and_oracle(gpu_ctx, which):
    print(f"[case] lattice lookups {which}", flush=True)
    t = {"tiny": synth.tiny, "small": synth.small, "displaced": synth.displaced_psf_tile}[which]()
    A0, B0, A1, B1, state = _both_paths(gpu_ctx, t)
    assert state == (True, True)
    _check(A0, B0, A1, B1, t)


def test_lattice_pitches_in_integer_ratios(gpu_ctx):
    """Exposure pitches 1 and 0.5 (x), 1 and 0.75 (y); output pitch 0.3:
    lattices with (P, Q) = (2, 1), (4, 3), (3, 10), (3, 5)."""
    print("[case] lattice lookups, pitches in ratios", flush=True)
    t = _pitch_tile([1.0, 0.5], [1.0, 0.75], 0.3)
    A0, B0, A1, B1, state = _both_paths(gpu_ctx, t)
    assert state == (True, True)
    _check(A0, B0, A1, B1, t)


def test_lattice_falls_back_without_small_ratio(gpu_ctx):
    """An output pitch of 1/pi input px has no small ratio to the input
    pitch: B takes the per-entry kernel (bitwise the same B), A the lattice."""
    print("[case] lattice lookups, output pitch 1/pi", flush=True)
    t = _pitch_tile([1.0, 1.0], [1.0, 1.0], 1.0 / np.pi)
    A0, B0, A1, B1, state = _both_paths(gpu_ctx, t)
    assert state == (True, False)
    assert np.array_equal(B1, B0)
    _check(A0, B0, A1, B1, t, oracle=False)


def test_set_grid_rejects_inconsistent_grid(gpu_ctx):
    print("[case] set_grid with a wrong dither", flush=True)
    t = synth.tiny()
    gpu_ctx.set_tile(t, grid=False)
    bad = dict(t.grid)
    bad["dither"] = t.grid["dither"] + np.array([[0.1, 0.0], [0.0, 0.0]])
    with pytest.raises(ImcomError, match="do not lie on the declared grids"):
        gpu_ctx.set_grid(**bad)
    assert gpu_ctx.lattice_state() == (False, False)
    gpu_ctx.set_grid(**t.grid)
    assert gpu_ctx.lattice_state() == (True, True)
