"""Host-side set-up of one IMCOM tile: everything the reference computes on the
CPU *before* its solve path, restated with numpy.

The reference keeps these stages on the host (coordinates, PSF padding and
rotation, PSF spectra, the FFT lookup tables, C, the noise diagonal and the
kappa range); their outputs are the inputs of the GPU solve path.  Each
function cites the Fortran routine it restates.  Arrays follow the reference's
Fortran layout: a 2-D array ``a[i1, i2]`` is stored column-major
(``order='F'``), pixel vectors are exposure-major, exposure ids are 1-based.
"""
from __future__ import annotations

import dataclasses
import numpy as np

# DLAMCH('E') in LAPACK 3.x with rounding: 2**-53 (imcom.f90:125, imcom_bisect.f90:187)
EPS = 2.0 ** -53


def xarr(n1: int, n2: int) -> np.ndarray:
    """x of pixel centres, 0 at the first pixel (imcom_proc.f90:29-40)."""
    return np.asfortranarray(np.repeat((np.arange(n1, dtype=np.float64))[:, None], n2, axis=1))


def yarr(n1: int, n2: int) -> np.ndarray:
    """y of pixel centres (imcom_proc.f90:44-52)."""
    return np.asfortranarray(np.repeat((np.arange(n2, dtype=np.float64))[None, :], n1, axis=0))


def build_xy(n1gim, n2gim, gimxscale, gimyscale, dither, mask_subs=None):
    """x_i, y_i for USERXY=0 regular grids (imcom_proc.f90:56-89).

    ``dither`` is (nexp, 2); the unmasked vector is the column-major reshape of
    (n1gim, n2gim, nexp) (imcom_proc.f90:80-83)."""
    nexp = dither.shape[0]
    xs = np.empty((n1gim, n2gim, nexp), order="F")
    ys = np.empty((n1gim, n2gim, nexp), order="F")
    for e in range(nexp):
        xs[:, :, e] = xarr(n1gim, n2gim) * gimxscale[e] + dither[e, 0]
        ys[:, :, e] = yarr(n1gim, n2gim) * gimyscale[e] + dither[e, 1]
    xu = xs.reshape(-1, order="F")
    yu = ys.reshape(-1, order="F")
    if mask_subs is None:
        return xu.copy(), yu.copy()
    return xu[mask_subs].copy(), yu[mask_subs].copy()


def build_XaYa(n1out, n2out, outxscale, outyscale, outpos1, outpos2):
    """Output pixel centres X_a, Y_a, column-major over (n1out, n2out)
    (imcom_proc.f90:93-119)."""
    Xa = xarr(n1out, n2out) * outxscale + outpos1
    Ya = yarr(n1out, n2out) * outyscale + outpos2
    return Xa.reshape(-1, order="F").copy(), Ya.reshape(-1, order="F").copy()


def mask_pixels(I_unmasked, saturation):
    """Saturation mask and 1-based exposure ids (imcom_io.f90:84-123).

    ``I_unmasked`` is (n1gim, n2gim, nexp).  Returns (mask_subs, I_i, exp_i)."""
    nexp = I_unmasked.shape[2]
    Iu = I_unmasked.reshape(-1, order="F")
    ex = np.empty(I_unmasked.shape, dtype=np.int32, order="F")
    for e in range(nexp):
        ex[:, :, e] = e + 1
    exu = ex.reshape(-1, order="F")
    mask_subs = np.nonzero(Iu < saturation)[0]
    return mask_subs, Iu[mask_subs].copy(), exu[mask_subs].copy()


def pad_size(x_i, y_i, X_a, Y_a, psfxscale, psfyscale, n1img, n2img):
    """n1pad = n2pad from the coordinate span (imcom_io.f90:279-286).  Fortran
    nint rounds half away from zero."""
    def nint(v):
        return int(np.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)
    n1pad = 2 * nint((max(x_i.max(), X_a.max()) - min(x_i.min(), X_a.min())) / psfxscale)
    n2pad = 2 * nint((max(y_i.max(), Y_a.max()) - min(y_i.min(), Y_a.min())) / psfyscale)
    n1pad = max(n1pad, n2pad)
    if n1pad < 2 * n1img or n1pad < 2 * n2img:
        n1pad = max(2 * n1img, 2 * n2img)
    return n1pad


def embed_psf(img, n1pad):
    """Zero-pad a PSF image into n1pad x n1pad and normalise to unit sum
    (imcom_io.f90:306-329)."""
    n1, n2 = img.shape
    out = np.zeros((n1pad, n1pad), order="F")
    i0 = (n1pad - n1) // 2
    j0 = (n1pad - n2) // 2
    out[i0:i0 + n1, j0:j0 + n2] = img / img.sum()
    return out


def _lagrange_weights(x, lo, cnt, width=8):
    """Weights of the polynomial through integer nodes lo..lo+cnt-1 evaluated
    at x (vectorised).  Returns (N, width) with zeros past cnt."""
    N = x.shape[0]
    w = np.zeros((N, width))
    for k in range(width):
        num = np.ones(N)
        den = np.ones(N)
        valid_k = k < cnt
        for l in range(width):
            if l == k:
                continue
            use = l < cnt
            num = np.where(use, num * (x - (lo + l)), num)
            den = np.where(use, den * float(k - l), den)
        w[:, k] = np.where(valid_k, num / den, 0.0)
    return w


def rotate_image(im_in, theta_deg, npoly=7, conserve_flux=1):
    """Rotate about ((n+1)/2, (n+1)/2) in 1-based pixels by polynomial
    interpolation (imcom_proc.f90:184-250).  Same stencil and clipping rules
    as the reference; the polynomial is evaluated in Lagrange form instead of
    Neville's tableau (same polynomial, last-bit rounding differences)."""
    if npoly % 2 == 0:
        npoly -= 1
    npm = npoly // 2
    nim1, nim2 = im_in.shape
    th = np.mod(theta_deg, 360.0) * 3.141592653589793 / 180.0
    c, s = np.cos(th), np.sin(th)
    xc = (nim1 + 1) / 2.0
    yc = (nim2 + 1) / 2.0
    io, jo = np.meshgrid(np.arange(1, nim1 + 1, dtype=np.float64),
                         np.arange(1, nim2 + 1, dtype=np.float64), indexing="ij")
    xo = (io - xc).ravel(order="F")
    yo = (jo - yc).ravel(order="F")
    xs = xo * c + yo * s + xc
    ys = -xo * s + yo * c + yc
    ifl, jfl = np.floor(xs).astype(np.int64), np.floor(ys).astype(np.int64)
    ice, jce = np.ceil(xs).astype(np.int64), np.ceil(ys).astype(np.int64)
    ok = (ifl >= 1) & (jfl >= 1) & (ice <= nim1) & (jce <= nim2)
    imin = np.maximum(1, ifl - npm)
    jmin = np.maximum(1, jfl - npm)
    imax = np.minimum(nim1, ice + npm)
    jmax = np.minimum(nim2, jce + npm)
    ci = np.where(ok, imax - imin + 1, 1)
    cj = np.where(ok, jmax - jmin + 1, 1)
    wx = _lagrange_weights(xs, imin, ci)
    wy = _lagrange_weights(ys, jmin, cj)
    out = np.zeros(xs.shape[0])
    for k in range(8):
        ii = np.clip(imin + k, 1, nim1) - 1
        for l in range(8):
            jj = np.clip(jmin + l, 1, nim2) - 1
            out += wx[:, k] * wy[:, l] * im_in[ii, jj]
    out = np.where(ok, out, 0.0).reshape((nim1, nim2), order="F")
    if conserve_flux != 0:
        out = im_in.sum() * out / out.sum()
    return np.asfortranarray(out)


def ft_r2c(G):
    """Full complex forward DFT of a real image (imcom_proc.f90:433-475: FFTW
    r2c plus Hermitian completion equals the full 2-D DFT)."""
    return np.fft.fft2(G)


def _pad_spectrum(prod, npad):
    """Place an n1psf^2 spectrum into an (npad*n1psf)^2 grid with the
    reference's pair of circular shifts (imcom_matrix.f90:403-410)."""
    n1psf, n2psf = prod.shape
    n1big, n2big = n1psf * npad, n2psf * npad
    n1min = (n1big - n1psf) // 2
    n2min = (n2big - n2psf) // 2
    u = np.zeros((n1big, n2big), dtype=np.complex128)
    # Fortran cshift(a, s, dim) == np.roll(a, -s, axis)
    u[n1min:n1min + n1psf, n2min:n2min + n2psf] = np.roll(np.roll(prod, -(n2psf // 2), 1),
                                                        -(n1psf // 2), 0)
    return np.roll(np.roll(u, -(n2big // 2), 1), -(n1big // 2), 0)


def _invft_shift(u):
    """Unnormalised inverse DFT (FFTW_BACKWARD), real part, recentred so the
    origin sits at 1-based pixel (n/2+1, n/2+1) (imcom_matrix.f90:411-414)."""
    n1big, n2big = u.shape
    a = np.fft.ifft2(u).real * (n1big * n2big)
    return np.roll(np.roll(a, -(n2big // 2), 1), -(n1big // 2), 0)


def build_Alookup(Gt_rot, npad):
    """A lookup tables, packed upper-triangular over exposure pairs
    ial = iexp + jexp*(jexp-1)/2 (imcom_matrix.f90:361-430).

    ``Gt_rot`` is a list of nexppsf spectra (n1psf x n1psf).  Returns an
    (n1big, n2big, nexppsf*(nexppsf+1)/2) Fortran-ordered array."""
    k = len(Gt_rot)
    n1psf, n2psf = Gt_rot[0].shape
    out = np.zeros((n1psf * npad, n2psf * npad, k * (k + 1) // 2), order="F")
    for iexp in range(1, k + 1):
        for jexp in range(iexp, k + 1):
            ial = iexp + jexp * (jexp - 1) // 2
            prod = np.conj(Gt_rot[iexp - 1]) * Gt_rot[jexp - 1]
            out[:, :, ial - 1] = _invft_shift(_pad_spectrum(prod, npad))
    out /= float(n1psf) * float(n2psf)
    return out


def build_Blookup(Gammat, Gt_rot, npad):
    """B lookup tables, -2 x the Gamma--G cross-correlation
    (imcom_matrix.f90:434-493)."""
    k = len(Gt_rot)
    n1psf, n2psf = Gammat.shape
    out = np.zeros((n1psf * npad, n2psf * npad, k), order="F")
    for iexp in range(k):
        prod = np.conj(Gammat) * Gt_rot[iexp]
        out[:, :, iexp] = -2.0 * _invft_shift(_pad_spectrum(prod, npad))
    out /= float(n1psf) * float(n2psf)
    return out


def calc_C(Gammat):
    """C = sum |Gamma~|^2 / (n1psf n2psf) (imcom_matrix.f90:546-552)."""
    n1, n2 = Gammat.shape
    return float(np.sum((Gammat * np.conj(Gammat)).real) / float(n1) / float(n2))


def build_Ndiag(n, noise):
    """Noise diagonal split into nexp equal chunks ignoring the mask
    (imcom_matrix.f90:556-582)."""
    nexp = len(noise)
    Nd = np.zeros(n)
    for i in range(1, nexp + 1):
        n_i = 1 + (i - 1) * (n // nexp)
        n_f = i * (n // nexp) if i != nexp else n
        Nd[n_i - 1:n_f] = noise[i - 1]
    return Nd


def test_psfconst(psffile, rotangdeg):
    """1 when every exposure has the same PSF file and (sic) rotation angles
    that all DIFFER from the first one by more than 1e-15; 0 otherwise
    (imcom_proc.f90:582-599, inherited as is)."""
    for i in range(1, len(psffile)):
        if psffile[i] != psffile[0] or abs(rotangdeg[i] - rotangdeg[0]) <= 1e-15:
            return 0
    return 1


@dataclasses.dataclass
class Tile:
    """All inputs of the solve path for one tile, in the reference's layout
    (the imcom_data globals the get_* procedures read, imcom_data.f90:24-132)."""
    n: int
    m: int
    x_i: np.ndarray
    y_i: np.ndarray
    exp_i: np.ndarray          # int32, 1-based
    X_a: np.ndarray
    Y_a: np.ndarray
    I_i: np.ndarray
    Ndiag: np.ndarray
    C_a: float
    kappa_min: float
    kappa_max: float
    Alookup: np.ndarray        # (n1al, n2al, nA) Fortran order
    Blookup: np.ndarray        # (n1bl, n2bl, nexppsf) Fortran order
    psfconst: int
    nexppsf: int
    npad: int = 3
    npoly: int = 7
    psfxscale: float = 0.25
    psfyscale: float = 0.25
    n1out: int = 0
    n2out: int = 0
    USB: bool = True           # True: bisect by U (leakage), False: by S (noise)
    US_max: float = 1e-6
    US_tol: float = 1e-9
    maxNbis: int = 53
    meta: dict = dataclasses.field(default_factory=dict)
    # USERXY=0 regular grids, the arguments of ImcomGPU.set_grid (None for
    # USERXY=1 tiles, whose coordinates come from files)
    grid: dict = None
    Gt_rot: list = None        # PSF spectra (n1pad^2 complex, nexppsf of them)
    Gammat: np.ndarray = None  # target spectrum


def assemble_tile(*, I_unmasked, gimxscale=None, gimyscale=None, dither=None, psf_img, rotangdeg,
                  gamma_img, psfxscale, psfyscale, noise, n1out=0, n2out=0, outxscale=None,
                  outyscale=None, outpos1=None, outpos2=None, psffile=None, npad=3, npoly=7,
                  saturation=np.finfo(np.float64).max, USB=True, US_max=1e-6,
                  US_tol=1e-9, meta=None, host_tables=True, xy_unmasked=None,
                  XaYa_grid=None) -> Tile:
    """Run the reference's CPU set-up (imcom.f90:55-127): masking,
    coordinates, psfconst, padding, rotation, spectra, lookup tables, Ndiag,
    C and the kappa range.  USERXY=0 inputs give the grid scales, dithers
    and output grid; USERXY=1 inputs give xy_unmasked = (x, y), each
    (n1gim, n2gim, nexp) (imcom_alloc_xy, imcom_io.f90:133-197), and
    XaYa_grid = (Xa, Ya), each (n1out, n2out) (imcom_alloc_XaYa,
    imcom_io.f90:201-240).  host_tables=False leaves the lookup tables to the
    device (ImcomGPU.set_tile(tile, tables="device")); the spectra are kept
    on the tile either way."""
    n1gim, n2gim, nexp = I_unmasked.shape
    mask_subs, I_i, exp_i = mask_pixels(I_unmasked, saturation)
    if xy_unmasked is None:
        x_i, y_i = build_xy(n1gim, n2gim, gimxscale, gimyscale, dither, mask_subs)
    else:
        x_i = np.asarray(xy_unmasked[0]).reshape(-1, order="F")[mask_subs].copy()
        y_i = np.asarray(xy_unmasked[1]).reshape(-1, order="F")[mask_subs].copy()
    if XaYa_grid is None:
        X_a, Y_a = build_XaYa(n1out, n2out, outxscale, outyscale, outpos1, outpos2)
    else:
        n1out, n2out = XaYa_grid[0].shape
        X_a = np.asarray(XaYa_grid[0]).reshape(-1, order="F").copy()
        Y_a = np.asarray(XaYa_grid[1]).reshape(-1, order="F").copy()
    if psffile is None:
        psffile = ["psf.fits"] * nexp
    psfconst = test_psfconst(psffile, rotangdeg)
    nexppsf = 1 if psfconst == 1 else nexp
    imgs = psf_img if isinstance(psf_img, list) else [psf_img] * nexp
    n1pad = pad_size(x_i, y_i, X_a, Y_a, psfxscale, psfyscale, *imgs[0].shape)
    Gt = []
    for e in range(nexppsf):
        G = embed_psf(imgs[e], n1pad)
        if abs(rotangdeg[e]) > 2.0 * 3.141592653589793 * EPS:
            G = rotate_image(G, rotangdeg[e], npoly, 1)
        Gt.append(ft_r2c(G))
    Gammat = ft_r2c(embed_psf(gamma_img, n1pad))
    n = x_i.shape[0]
    grid = None
    if xy_unmasked is None and XaYa_grid is None:
        dth = np.asarray(dither, dtype=np.float64).reshape(nexp, 2)
        grid = dict(n1gim=n1gim, n2gim=n2gim, nexp=nexp,
                    gimxscale=np.asarray(gimxscale, dtype=np.float64).copy(),
                    gimyscale=np.asarray(gimyscale, dtype=np.float64).copy(),
                    dither=dth.copy(), mask_subs_i=(mask_subs + 1).astype(np.int32),
                    n1out=n1out, n2out=n2out, outxscale=float(outxscale),
                    outyscale=float(outyscale), outpos1=float(outpos1), outpos2=float(outpos2))
    Ndiag = build_Ndiag(n, noise)
    C_a = calc_C(Gammat)
    return Tile(n=n, m=X_a.shape[0], x_i=x_i, y_i=y_i, exp_i=exp_i.astype(np.int32),
                X_a=X_a, Y_a=Y_a, I_i=I_i, Ndiag=Ndiag, C_a=C_a,
                kappa_min=C_a * EPS, kappa_max=C_a / EPS,
                Alookup=build_Alookup(Gt, npad) if host_tables else None,
                Blookup=build_Blookup(Gammat, Gt, npad) if host_tables else None,
                psfconst=psfconst, nexppsf=nexppsf, npad=npad, npoly=npoly,
                psfxscale=psfxscale, psfyscale=psfyscale, n1out=n1out, n2out=n2out,
                USB=USB, US_max=US_max, US_tol=US_tol, meta=dict(meta or {}, n1pad=n1pad),
                grid=grid, Gt_rot=Gt, Gammat=Gammat)
