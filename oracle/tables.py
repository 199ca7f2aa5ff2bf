"""CPU oracle of the lookup-table set-up -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, as the reference the device table build is compared
with; the GPU library never calls it.

Pieces and what they follow (file:line):
  assemble_inputs  the masking of imcom_alloc_gims (imcom_io.f90:84-123) and
                imcom_build_xy (imcom_proc.f90:56-89), as loops over the
                column-major pixel order
  rotate_image  imcom_rotate_image (imcom_proc.f90:184-250), in C
                (imcom_oracle.c: or_rotate_image) over the oracle's Neville
                routine; the angle's cos/sin from the host libm
  ft_r2c        imcom_ft_r2c (imcom_proc.f90:433-475): the FFTW r2c half
                spectrum (first axis 0..n1/2, FFTW_FORWARD, unnormalised),
                then the Hermitian-completion statements as written
  build_Alookup imcom_build_Alookup (imcom_matrix.f90:361-430): ufunc zeroed,
                the inner block set from cshift(cshift(conj(G_i) G_j, n2/2, 2),
                n1/2, 1), the outer cshift pair by N/2, the FFTW backward
                transform (unnormalised), real part, cshift pair by N/2, and
                the final division by n1psf n2psf; packed upper triangle
                ial = iexp + jexp (jexp - 1) / 2
  build_Blookup imcom_build_Blookup (imcom_matrix.f90:434-493): the same with
                conj(Gamma~) G_i and the factor -2
  calc_C        imcom_calc_C (imcom_matrix.f90:546-552)

How each piece is pinned (tests/test_tables_oracle.py):
  assemble_inputs  a hand-worked 3 x 2 x 2 case (saturated pixels, the
                strict comparison, exposure ids, dithers); agreement with the
                host set-up (imcom_amd/tile.py)
  rotate_image  identity at 0 degrees; exact reproduction of a cubic image
                at the rotated positions (rotation direction and centre);
                a quarter turn as a pixel permutation; flux conservation
  ft_r2c        the DFT sum written out on small images; Hermitian symmetry
  tables        the DFT sums written out with the reference's 1-based index
                ranges on tiny grids (n = 4, 6; npad = 1..3); the Gaussian
                closed forms (autocorrelation and -2 x the cross-correlation,
                Parseval for C); agreement with the host set-up
                (imcom_amd/tile.py)
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import oracle as _O

_dp = ctypes.POINTER(ctypes.c_double)


def assemble_inputs(Im_unmasked, saturation, gimxscale, gimyscale, dither):
    """x_i, y_i, exp_i, I_i of the pixels with I < saturation."""
    n1, n2, nexp = Im_unmasked.shape
    xs, ys, es, Is = [], [], [], []
    for e in range(nexp):                     # column-major (i1, i2, e) order
        for i2 in range(n2):
            for i1 in range(n1):
                v = Im_unmasked[i1, i2, e]
                if v < saturation:
                    xs.append(float(i1) * gimxscale[e] + dither[e][0])
                    ys.append(float(i2) * gimyscale[e] + dither[e][1])
                    es.append(e + 1)
                    Is.append(v)
    return np.array(xs), np.array(ys), np.array(es, dtype=np.int32), np.array(Is)


def rotate_image(im_in, theta_deg, npoly=7, conserve_flux=1):
    """imcom_rotate_image: the rotated image (same shape)."""
    a = np.asfortranarray(im_in, dtype=np.float64)
    n1, n2 = a.shape
    out = np.empty((n1, n2), order="F")
    th = math.fmod(theta_deg, 360.0) * 3.141592653589793 / 180.0
    L = _O.lib()
    L.or_rotate_image.restype = None
    L.or_rotate_image.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]
    L.or_rotate_image(n1, n2, a.ctypes.data_as(_dp), math.cos(th), math.sin(th), npoly,
                      conserve_flux, out.ctypes.data_as(_dp))
    return out


def cshift(a, shift, dim):
    """Fortran CSHIFT(a, shift, dim): result(i) = a(i + shift), circular."""
    return np.roll(a, -shift, axis=dim - 1)


def ft_r2c(F):
    """imcom_ft_r2c: complex n1 x n2 spectrum of a real image."""
    F = np.asarray(F, dtype=np.float64)
    n1, n2 = F.shape
    # FFTW r2c over a column-major n1 x n2 array halves the first axis
    Ft_tmp = np.fft.rfft2(F.T).T            # (n1//2 + 1, n2), exp(-i...)
    Ft_conj = np.conj(Ft_tmp)
    Ft = np.zeros((n1, n2), dtype=np.complex128)
    Ft[0:n1 // 2 + 1, :] = Ft_tmp[0:n1 // 2 + 1, :]
    # Ft_dum(n1/2+1 : n1-1, 0) = Ft_conj((n1-1)/2 : 1 : -1, 0)
    rows = np.arange(n1 // 2 + 1, n1)
    Ft[rows, 0] = Ft_conj[(n1 - 1) // 2 - (rows - (n1 // 2 + 1)), 0]
    for i in range(1, (n1 - 1) // 2 + 1):
        src = n1 // 2 - i if n1 % 2 == 0 else n1 // 2 + 1 - i
        Ft[n1 // 2 + i, 1:n2] = Ft_conj[src, n2 - 1:0:-1]
    return Ft


def _pad_and_invert(prod, npad):
    n1psf, n2psf = prod.shape
    n1big, n2big = n1psf * npad, n2psf * npad
    n1min = (n1big - n1psf) // 2 + 1         # 1-based, as in the reference
    n2min = (n2big - n2psf) // 2 + 1
    n1max = (n1big + n1psf) // 2
    n2max = (n2big + n2psf) // 2
    ufunc = np.zeros((n1big, n2big), dtype=np.complex128)
    ufunc[n1min - 1:n1max, n2min - 1:n2max] = cshift(cshift(prod, n2psf // 2, 2), n1psf // 2, 1)
    ufunc = cshift(cshift(ufunc, n2big // 2, 2), n1big // 2, 1)
    # FFTW_BACKWARD, unnormalised: sum_k u(k) exp(+2 pi i k x / N)
    tmp = np.fft.ifft2(ufunc) * (n1big * n2big)
    return cshift(cshift(tmp.real, n2big // 2, 2), n1big // 2, 1)


def build_Alookup(Gt_rot, npad):
    """(n1big, n2big, nexppsf (nexppsf + 1) / 2) A tables."""
    k = len(Gt_rot)
    n1psf, n2psf = Gt_rot[0].shape
    out = np.zeros((n1psf * npad, n2psf * npad, k * (k + 1) // 2), order="F")
    for iexp in range(1, k + 1):
        for jexp in range(iexp, k + 1):
            ial = iexp + jexp * (jexp - 1) // 2
            prod = np.conj(Gt_rot[iexp - 1]) * Gt_rot[jexp - 1]
            out[:, :, ial - 1] = _pad_and_invert(prod, npad)
    return out / float(n1psf) / float(n2psf)


def build_Blookup(Gammat, Gt_rot, npad):
    """(n1big, n2big, nexppsf) B tables."""
    k = len(Gt_rot)
    n1psf, n2psf = Gammat.shape
    out = np.zeros((n1psf * npad, n2psf * npad, k), order="F")
    for iexp in range(k):
        out[:, :, iexp] = -2.0 * _pad_and_invert(np.conj(Gammat) * Gt_rot[iexp], npad)
    return out / float(n1psf) / float(n2psf)


def calc_C(Gammat):
    n1, n2 = Gammat.shape
    return float(np.sum((Gammat * np.conj(Gammat)).real) / float(n1) / float(n2))
