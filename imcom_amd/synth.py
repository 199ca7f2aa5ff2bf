"""Seeded synthetic IMCOM tiles for the benchmark workloads and tests.

The workloads (n = input pixels, m = output pixels):

* ``small``    4 exposures of 16x16 (n=1024), Gaussian PSF sigma=1.2 px on a
               128^2 image at 0.25 px, Gaussian target sigma=1.5, 24x24 output
               at 0.5 px (m=576), noise sigma 1e-3, U_max=1e-6.
* ``large``    4 exposures of 64x64 (n=16384), Airy PSF (obscuration 0.3,
               lambda/D = 1.0 px) convolved with the unit pixel on a 256^2
               image at 0.25 px, Gaussian target sigma=1.8, 96x96 output at
               0.25 px (m=9216, a 24 px square centred in the 64 px
               footprint), U_max=1e-6.  Both choices make U <= 1e-6 reachable
               at every output pixel: at lambda/D = 0.5 px the aliased Airy
               power leaves a leakage floor near 2e-4, and the Airy wings need
               about 20 px of inputs around an output pixel to bring the floor
               below 1e-6 (a 0.5 px output pitch would leave the grid's edge
               8 px from the footprint's).
* ``batch``    tiles with 2-6 exposures of 16-32 px squares (n in
               [1024, 6144]), m in [256, 2304], sigma_psf ~ U[1.0, 1.6],
               noise sigma log-uniform in [1e-4, 1e-2].
* ``tiny``     small shapes for unit tests.

Dithers are uniform in [0,1) px and PSF rotations uniform in [0,360) deg,
drawn from the seed; with USERXY=0 a rotation turns the PSF only, never the
pixel grid (imcom_proc.f90:68-73).
"""
from __future__ import annotations

import numpy as np

from .tile import Tile, assemble_tile

try:  # scipy is only needed for the Airy profile
    from scipy.special import j1 as _j1
except Exception:  # pragma: no cover
    _j1 = None


def gaussian_psf_image(npix, sigma_px, psfscale, shift=(0.0, 0.0)):
    """Circular Gaussian sampled at PSF pixel centres, peak at 1-based pixel
    npix/2+1 displaced by `shift` input pixels (first axis, second axis)."""
    c = npix // 2
    s = sigma_px / psfscale
    i = np.arange(npix) - c - shift[0] / psfscale
    j = np.arange(npix) - c - shift[1] / psfscale
    g = np.exp(-0.5 * (i[:, None] ** 2 + j[None, :] ** 2) / s ** 2)
    return np.asfortranarray(g)


def airy_pixel_psf_image(npix, lam_over_d_px, obscuration, psfscale):
    """Annular-pupil Airy intensity convolved with the unit square input
    pixel (a sinc multiply in Fourier space), sampled on PSF pixels."""
    if _j1 is None:
        raise RuntimeError("scipy.special is needed for the Airy PSF")
    c = npix // 2
    i = (np.arange(npix) - c) * psfscale
    r = np.sqrt(i[:, None] ** 2 + i[None, :] ** 2)
    x = np.pi * r / lam_over_d_px
    eps = obscuration
    with np.errstate(invalid="ignore", divide="ignore"):
        a1 = np.where(x > 0, 2.0 * _j1(x) / x, 1.0)
        a2 = np.where(x > 0, 2.0 * _j1(eps * x) / (eps * x), 1.0)
    amp = (a1 - eps ** 2 * a2) / (1.0 - eps ** 2)
    im = amp ** 2
    u = np.fft.fftfreq(npix)  # cycles per PSF pixel
    w = 1.0 / psfscale        # input pixel width in PSF pixels
    box = np.sinc(u * w)
    im = np.fft.ifft2(np.fft.fft2(np.fft.ifftshift(im)) * box[:, None] * box[None, :]).real
    return np.asfortranarray(np.fft.fftshift(im))


def _gauss_source(x, y, xc, yc, sigma, flux=1.0):
    r2 = (x - xc) ** 2 + (y - yc) ** 2
    return flux * np.exp(-0.5 * r2 / sigma ** 2) / (2 * np.pi * sigma ** 2)


def make_tile(*, nexp, npix_in, psf_kind="gauss", psf_sigma=1.2, gamma_sigma=1.5,
              psf_npix=128, psfscale=0.25, nout, outscale=0.5, noise_sigma=1e-3,
              US_max=1e-6, US_tol=1e-9, USB=True, seed=0, lam_over_d=0.5,
              obscuration=0.3, name="tile", host_tables=True) -> Tile:
    rng = np.random.default_rng(seed)
    dither = rng.uniform(0.0, 1.0, size=(nexp, 2))
    rot = 360.0 * rng.uniform(0.0, 1.0, size=nexp)
    if psf_kind == "gauss":
        psf = gaussian_psf_image(psf_npix, psf_sigma, psfscale)
        src_sigma = psf_sigma
    elif psf_kind == "airy":
        psf = airy_pixel_psf_image(psf_npix, lam_over_d, obscuration, psfscale)
        src_sigma = 0.45 * lam_over_d + 0.3
    else:
        raise ValueError(psf_kind)
    gam = gaussian_psf_image(psf_npix, gamma_sigma, psfscale)
    # footprint centre and an output grid centred inside it
    xc = 0.5 * (npix_in - 1) + dither[:, 0].mean()
    yc = 0.5 * (npix_in - 1) + dither[:, 1].mean()
    half = 0.5 * outscale * (nout - 1)
    I = np.empty((npix_in, npix_in, nexp), order="F")
    gx = np.arange(npix_in, dtype=np.float64)
    for e in range(nexp):
        X = gx[:, None] + dither[e, 0]
        Y = gx[None, :] + dither[e, 1]
        I[:, :, e] = _gauss_source(X, Y, xc, yc, src_sigma)
    I += rng.normal(0.0, noise_sigma, size=I.shape)
    return assemble_tile(
        I_unmasked=I, gimxscale=np.ones(nexp), gimyscale=np.ones(nexp), dither=dither,
        psf_img=psf, rotangdeg=rot, gamma_img=gam, psfxscale=psfscale, psfyscale=psfscale,
        noise=np.full(nexp, noise_sigma ** 2), n1out=nout, n2out=nout, outxscale=outscale,
        outyscale=outscale, outpos1=xc - half, outpos2=yc - half, USB=USB, US_max=US_max,
        US_tol=US_tol, meta=dict(name=name, seed=seed, psf_kind=psf_kind, psf_sigma=psf_sigma,
                                 gamma_sigma=gamma_sigma, noise_sigma=noise_sigma),
        host_tables=host_tables)


def small(seed=20261204, **kw) -> Tile:
    """4 x 16^2 inputs (n=1024), 24^2 output (m=576)."""
    args = dict(nexp=4, npix_in=16, psf_kind="gauss", psf_sigma=1.2, gamma_sigma=1.5,
                psf_npix=128, nout=24, noise_sigma=1e-3, seed=seed, name="small")
    args.update(kw)
    return make_tile(**args)


def large(seed=20261205, **kw) -> Tile:
    """4 x 64^2 inputs (n=16384), 96^2 output at 0.25 px (m=9216), Airy x pixel
    PSF with lambda/D = 1 input pixel."""
    args = dict(nexp=4, npix_in=64, psf_kind="airy", lam_over_d=1.0, gamma_sigma=1.8,
                psf_npix=256, nout=96, outscale=0.25, noise_sigma=1e-3, seed=seed, name="large")
    args.update(kw)
    return make_tile(**args)


def displaced_psf_tile(shift=(0.75, -0.5), seed=3, npix_in=8, nout=8, psf_sigma=1.2,
                       gamma_sigma=1.5, psfscale=0.25, psf_npix=64) -> Tile:
    """Two exposures with different PSF files (psfconst = 0, no rotation):
    exposure 1 sees a centred Gaussian, exposure 2 the same Gaussian
    displaced by `shift` input pixels.  The lookup results then have closed
    forms that are not point-symmetric, which fixes the sign and the axis of
    the table arguments x_j - x_i (A) and x_i - X_a (B)."""
    rng = np.random.default_rng(seed)
    nexp = 2
    dither = rng.uniform(0.0, 1.0, size=(nexp, 2))
    psfs = [gaussian_psf_image(psf_npix, psf_sigma, psfscale),
            gaussian_psf_image(psf_npix, psf_sigma, psfscale, shift)]
    gam = gaussian_psf_image(psf_npix, gamma_sigma, psfscale)
    xc = 0.5 * (npix_in - 1) + dither[:, 0].mean()
    yc = 0.5 * (npix_in - 1) + dither[:, 1].mean()
    half = 0.25 * (nout - 1)
    I = rng.normal(0.0, 1e-3, size=(npix_in, npix_in, nexp))
    return assemble_tile(
        I_unmasked=np.asfortranarray(I), gimxscale=np.ones(nexp), gimyscale=np.ones(nexp),
        dither=dither, psf_img=psfs, rotangdeg=np.zeros(nexp), gamma_img=gam,
        psfxscale=psfscale, psfyscale=psfscale, noise=np.full(nexp, 1e-6), n1out=nout,
        n2out=nout, outxscale=0.5, outyscale=0.5, outpos1=xc - half, outpos2=yc - half,
        psffile=["psf1.fits", "psf2.fits"],
        meta=dict(name="displaced_psf", shift=tuple(shift), psf_sigma=psf_sigma,
                  gamma_sigma=gamma_sigma, psfscale=psfscale))


def tiny(seed=7, nexp=2, npix_in=8, nout=8, psf_npix=64, **kw) -> Tile:
    args = dict(nexp=nexp, npix_in=npix_in, psf_kind="gauss", psf_sigma=1.2,
                gamma_sigma=1.5, psf_npix=psf_npix, nout=nout, seed=seed, name="tiny")
    args.update(kw)
    return make_tile(**args)


def batch_shapes(ntiles=64, seed=20261207):
    """Per-tile (nexp, npix_in, nout, psf_sigma, noise_sigma) for the batched
    workload: n = nexp*npix^2 in [1024, 6144], m = nout^2 in [256, 2304],
    the output grid kept inside the input footprint."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < ntiles:
        nexp = int(rng.integers(2, 7))
        npix = int(rng.integers(16, 33))
        n = nexp * npix * npix
        if not (1024 <= n <= 6144):
            continue
        nout_max = min(48, 2 * (npix - 4))
        nout = int(rng.integers(16, nout_max + 1))
        sig = float(rng.uniform(1.0, 1.6))
        noise = float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2))))
        out.append((nexp, npix, nout, sig, noise, int(rng.integers(1 << 30))))
    return out


def batch(ntiles=64, seed=20261207, psf_npix=64):
    tiles = []
    for (nexp, npix, nout, sig, noise, s) in batch_shapes(ntiles, seed):
        tiles.append(make_tile(nexp=nexp, npix_in=npix, psf_sigma=sig, gamma_sigma=1.5,
                               psf_npix=psf_npix, nout=nout, noise_sigma=noise, seed=s,
                               name="batch"))
    return tiles
