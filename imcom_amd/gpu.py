"""Python host mirror of the reference's solve-path procedures over the C ABI
of libimcom_hip.so (include/imcom_gpu.h).

``ImcomGPU`` keeps the reference's procedure names and argument meaning
(imcom_get_A / imcom_lookup_A, imcom_get_QL / imcom_eigen_A, imcom_get_B /
imcom_lookup_B, imcom_get_P / imcom_build_P, imcom_get_T /
imcom_bisect_kappa + imcom_build_T/S/U, imcom_build_image); errors raise
``ImcomError`` carrying the library's message, as the reference prints
"IMCOM ERROR: ..." and stops.  There is no CPU fallback: if the library is
missing or no GPU is visible, construction fails.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimcom_hip.so")

IMCOM_PREC_FP64 = 0
IMCOM_PREC_FP32X3 = 1
IMCOM_PREC_BF16X6 = 2

# every symbol declared in include/imcom_gpu.h
# the library's default reduction (imcom_gpu_set_reduction_stages)
REDUCTION_STAGES = 2
TWOSTAGE_MIN = 7168

EXPORTS = [
    "imcom_gpu_init", "imcom_gpu_finalize", "imcom_gpu_last_error", "imcom_gpu_synchronize",
    "imcom_gpu_set_tile", "imcom_gpu_set_grid", "imcom_gpu_lattice_state",
    "imcom_gpu_set_tables", "imcom_gpu_assemble_inputs",
    "imcom_gpu_rotate_image", "imcom_gpu_ft_r2c", "imcom_gpu_build_Alookup",
    "imcom_gpu_build_Blookup", "imcom_gpu_calc_C", "imcom_gpu_lookup_A", "imcom_gpu_set_A",
    "imcom_gpu_eigen_A", "imcom_gpu_tridiagonalize", "imcom_gpu_tridiag_eig",
    "imcom_gpu_set_QL", "imcom_gpu_lookup_B", "imcom_gpu_set_B", "imcom_gpu_build_P",
    "imcom_gpu_set_P", "imcom_gpu_set_T", "imcom_gpu_bisect_kappa", "imcom_gpu_build_T", "imcom_gpu_build_S",
    "imcom_gpu_build_U", "imcom_gpu_build_image", "imcom_gpu_set_profiling",
    "imcom_gpu_stage_times", "imcom_gpu_mlp_nparams", "imcom_gpu_mlp_set_params",
    "imcom_gpu_mlp_get_params", "imcom_gpu_kappa_features", "imcom_gpu_mlp_infer",
    "imcom_gpu_mlp_grad", "imcom_gpu_mlp_train", "imcom_gpu_bisect_kappa_learned",
    "imcom_gpu_batch_set", "imcom_gpu_batch_solve", "imcom_gpu_batch_reserve",
    "imcom_gpu_batch_result",
    "imcom_gpu_set_tuning", "imcom_gpu_set_reduction_tail", "imcom_gpu_set_reduction_stages",
    "imcom_gpu_band_reduce", "imcom_gpu_fetch_async", "imcom_gpu_fetch_wait",
    "imcom_gpu_host_alloc", "imcom_gpu_host_free", "imcom_gpu_bisect_kappa_learned_noalloc",
]

# stage outputs of fetch_async (IMCOM_FETCH_*)
FETCH = {"A": 0, "Q": 1, "L": 2, "B": 3, "P": 4, "T": 5}


class TileDesc(ctypes.Structure):
    """struct imcom_gpu_tile (include/imcom_gpu.h)."""
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64),
                ("x_i", ctypes.c_void_p), ("y_i", ctypes.c_void_p), ("exp_i", ctypes.c_void_p),
                ("X_a", ctypes.c_void_p), ("Y_a", ctypes.c_void_p), ("Ndiag", ctypes.c_void_p),
                ("I_i", ctypes.c_void_p),
                ("C_a", ctypes.c_double), ("kappa_min", ctypes.c_double),
                ("kappa_max", ctypes.c_double), ("US_max", ctypes.c_double),
                ("US_tol", ctypes.c_double),
                ("UnotS", ctypes.c_int), ("maxnits", ctypes.c_int),
                ("psfconst", ctypes.c_int), ("nexppsf", ctypes.c_int), ("n1", ctypes.c_int),
                ("n2", ctypes.c_int),
                ("Alookup", ctypes.c_void_p), ("Blookup", ctypes.c_void_p),
                ("npoly", ctypes.c_int), ("npad", ctypes.c_int),
                ("psfxscale", ctypes.c_double), ("psfyscale", ctypes.c_double)]

STAGE_SLOTS = ["lookup_A", "sytrd", "stedc", "ormtr", "lookup_B", "build_P", "bisect",
               "build_T", "maps", "symv", "syr2k", "tables", "bulge", "q2"]


class ImcomError(RuntimeError):
    pass


_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double

_LIB = None


def load_library(path: str = LIB_PATH):
    """Load libimcom_hip.so and declare the C signatures.  Loading works
    without a GPU; calls need one."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ImcomError(f"{path} is missing: build it with imcom_amd.build.build()")
    L = ctypes.CDLL(path)
    sig = {
        "imcom_gpu_init": (_int, [ctypes.POINTER(_vp), _int]),
        "imcom_gpu_finalize": (None, [_vp]),
        "imcom_gpu_last_error": (ctypes.c_char_p, [_vp]),
        "imcom_gpu_synchronize": (_int, [_vp]),
        "imcom_gpu_set_tile": (_int, [_vp, _i64, _i64, _dp, _dp, _ip, _dp, _dp, _dp, _dbl]),
        "imcom_gpu_set_grid": (_int, [_vp, _int, _int, _int, _dp, _dp, _dp, _ip, _int, _int,
                                       _dbl, _dbl, _dbl, _dbl]),
        "imcom_gpu_lattice_state": (_int, [_vp, _ip, _ip]),
        "imcom_gpu_set_tables": (_int, [_vp, _int, _int, _int, _int, _dp, _dp]),
        "imcom_gpu_lookup_A": (_int, [_vp, _int, _int, _dbl, _dbl, _dp]),
        "imcom_gpu_set_A": (_int, [_vp, _dp]),
        "imcom_gpu_eigen_A": (_int, [_vp, _dp, _dp, _dp]),
        "imcom_gpu_tridiagonalize": (_int, [_vp, _dp, _dp]),
        "imcom_gpu_tridiag_eig": (_int, [_vp, _i64, _dp, _dp, _dp, _dp]),
        "imcom_gpu_set_QL": (_int, [_vp, _dp, _dp]),
        "imcom_gpu_lookup_B": (_int, [_vp, _int, _int, _dbl, _dbl, _dp]),
        "imcom_gpu_set_B": (_int, [_vp, _dp]),
        "imcom_gpu_build_P": (_int, [_vp, _int, _dp]),
        "imcom_gpu_set_P": (_int, [_vp, _dp]),
        "imcom_gpu_bisect_kappa": (_int, [_vp, _int, _dbl, _dbl, _dbl, _dbl, _int, _fp, _dp, _dp, _ip]),
        "imcom_gpu_build_T": (_int, [_vp, _int, _dp]),
        "imcom_gpu_set_T": (_int, [_vp, _dp]),
        "imcom_gpu_build_S": (_int, [_vp, _dp]),
        "imcom_gpu_build_U": (_int, [_vp, _dp]),
        "imcom_gpu_build_image": (_int, [_vp, _dp, _dp]),
        "imcom_gpu_set_profiling": (_int, [_vp, _int]),
        "imcom_gpu_stage_times": (_int, [_vp, _dp, ctypes.POINTER(_i64), _int]),
        "imcom_gpu_mlp_nparams": (_int, []),
        "imcom_gpu_mlp_set_params": (_int, [_vp, _fp]),
        "imcom_gpu_mlp_get_params": (_int, [_vp, _fp]),
        "imcom_gpu_kappa_features": (_int, [_vp, _dbl, _fp]),
        "imcom_gpu_mlp_infer": (_int, [_vp, _fp, _i64, _fp]),
        "imcom_gpu_mlp_grad": (_int, [_vp, _fp, _fp, _i64, _fp, _fp]),
        "imcom_gpu_mlp_train": (_int, [_vp, _fp, _fp, _i64, _int, _int, ctypes.c_float,
                                       ctypes.c_uint64, _fp]),
        "imcom_gpu_bisect_kappa_learned": (_int, [_vp, _int, _dbl, _dbl, _dbl, _dbl, _int,
                                                  ctypes.c_float, _dp, _dp, _ip, _ip]),
        "imcom_gpu_batch_set": (_int, [_vp, _int, ctypes.POINTER(TileDesc)]),
        "imcom_gpu_batch_solve": (_int, [_vp, _int, _int, _int, ctypes.c_float, _int]),
        "imcom_gpu_batch_reserve": (_int, [_vp, _int, _int]),
        "imcom_gpu_batch_result": (_int, [_vp, _int, _dp, _dp, _dp, _dp, _ip, _ip, _dp]),
        "imcom_gpu_set_tuning": (_int, [_vp, _int, _int]),
        "imcom_gpu_set_reduction_tail": (_int, [_vp, _int]),
        "imcom_gpu_set_reduction_stages": (_int, [_vp, _int, _i64]),
        "imcom_gpu_band_reduce": (_int, [_vp, _dp]),
        "imcom_gpu_assemble_inputs": (_int, [_vp, _int, _int, _int, _dp, _dbl, _dp, _dp, _dp,
                                              ctypes.POINTER(_i64), _dp, _dp, _ip, _dp]),
        "imcom_gpu_rotate_image": (_int, [_vp, _int, _int, _dp, _dbl, _int, _int, _dp]),
        "imcom_gpu_ft_r2c": (_int, [_vp, _int, _int, _int, _dp, _dp]),
        "imcom_gpu_build_Alookup": (_int, [_vp, _int, _int, _int, _int, _int, _dp, _dp]),
        "imcom_gpu_build_Blookup": (_int, [_vp, _int, _int, _int, _int, _int, _dp, _dp, _dp]),
        "imcom_gpu_calc_C": (_int, [_vp, _int, _int, _dp, _dp]),
        "imcom_gpu_fetch_async": (_int, [_vp, _int, _dp]),
        "imcom_gpu_fetch_wait": (_int, [_vp, _int]),
        "imcom_gpu_host_alloc": (_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
        "imcom_gpu_host_free": (_int, [_vp]),
        "imcom_gpu_bisect_kappa_learned_noalloc": (_int, [_vp, _vp, _i64, _i64, _vp, _vp, _i64, _vp,
                                                          _dbl, _vp, _int, _dbl, _dbl, _dbl, _dbl,
                                                          _int, ctypes.c_float, _vp, _vp, _vp,
                                                          _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def host_empty(shape, order="F") -> np.ndarray:
    """An uninitialised float64 array in page-locked host memory
    (imcom_gpu_host_alloc), released when the last view of it is gone.
    Device-to-host copies into it run asynchronously (fetch_async)."""
    L = load_library()
    count = int(np.prod(shape))
    p = _vp()
    rc = L.imcom_gpu_host_alloc(max(count, 1) * 8, ctypes.byref(p))
    if rc != 0:
        raise ImcomError(f"host_alloc of {count * 8} bytes failed (code {rc})")
    flat = np.ctypeslib.as_array(ctypes.cast(p, _dp), shape=(max(count, 1),))
    weakref.finalize(flat, L.imcom_gpu_host_free, p.value)
    return flat[:count].reshape(shape, order=order)


def _ptr(a, kind=_dp):
    return None if a is None else a.ctypes.data_as(kind)


def _vec(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _mat(a):
    return np.asfortranarray(a, dtype=np.float64)


class ImcomGPU:
    """One tile's solve path on one GPU (a context of the C ABI)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _vp()
        rc = self.lib.imcom_gpu_init(ctypes.byref(h), device)
        if rc != 0:
            raise ImcomError(f"imcom_gpu_init failed (code {rc}): no usable GPU {device}")
        self.h = h
        self.n = self.m = 0
        self._inflight = {}  # output name -> host arrays with a copy in flight

    def close(self):
        if getattr(self, "h", None):
            self.lib.imcom_gpu_finalize(self.h)  # waits for every copy in flight
            self.h = None
            self._inflight = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc != 0:
            msg = self.lib.imcom_gpu_last_error(self.h).decode(errors="replace")
            raise ImcomError(f"IMCOM ERROR: {what}: {msg} (code {rc})")

    # ---- inputs ---------------------------------------------------------
    def set_tile(self, tile, tables="host", grid=True):
        """Hand over the imcom_data globals of one tile.  tables="host"
        uploads the tile's Alookup/Blookup; "device" builds them on the GPU
        from the tile's spectra Gt_rot and Gammat (build_Alookup/Blookup);
        "none" hands over no tables (a T read from TFILE needs none).
        grid=True also declares a USERXY=0 tile's regular grids (set_grid),
        so that lookup_A/B interpolate each pixel separation once."""
        self.n, self.m = tile.n, tile.m
        self._keep = [_vec(tile.x_i), _vec(tile.y_i), _vec(tile.exp_i, np.int32),
                      _vec(tile.X_a), _vec(tile.Y_a), _vec(tile.Ndiag)]
        x, y, e, X, Y, N = self._keep
        self._ck(self.lib.imcom_gpu_set_tile(self.h, tile.n, tile.m, _ptr(x), _ptr(y), _ptr(e, _ip),
                                             _ptr(X), _ptr(Y), _ptr(N), tile.C_a), "set_tile")
        self._inflight = {}  # set_tile has waited for every copy
        g = getattr(tile, "grid", None) if grid else None
        if g:
            if len(g["mask_subs_i"]) != tile.n or g["n1out"] * g["n2out"] != tile.m:
                raise ImcomError("set_tile: tile.grid does not describe this tile's pixels")
            self.set_grid(**g)
        if tables == "device":
            if tile.Gt_rot is None or tile.Gammat is None:
                raise ImcomError("set_tile(tables='device') needs the tile's spectra")
            self.build_Alookup(tile.psfconst, tile.Gt_rot, tile.npad)
            self.build_Blookup(tile.psfconst, tile.Gammat, tile.Gt_rot, tile.npad)
        elif tables == "none":
            pass
        elif tables == "host":
            A, B = _mat(tile.Alookup), _mat(tile.Blookup)
            self._ck(self.lib.imcom_gpu_set_tables(self.h, tile.psfconst, tile.nexppsf, A.shape[0],
                                                   A.shape[1], _ptr(A), _ptr(B)), "set_tables")
        else:
            raise ValueError("tables must be 'host', 'device' or 'none'")
        self.tile = tile

    def set_grid(self, n1gim, n2gim, nexp, gimxscale, gimyscale, dither, mask_subs_i, n1out,
                 n2out, outxscale, outyscale, outpos1, outpos2):
        """The regular grids of a USERXY=0 tile (imcom_build_xy /
        imcom_build_XaYa, imcom_proc.f90:56-119): dither is (nexp, 2),
        mask_subs_i the 1-based unmasked indices.  n1gim = 0 clears them."""
        gx, gy = _vec(gimxscale), _vec(gimyscale)
        dth = np.asfortranarray(np.asarray(dither, dtype=np.float64).reshape(nexp, 2))
        ms = _vec(mask_subs_i, np.int32)
        self._grid_keep = [gx, gy, dth, ms]
        self._ck(self.lib.imcom_gpu_set_grid(self.h, n1gim, n2gim, nexp, _ptr(gx), _ptr(gy),
                                             _ptr(dth), _ptr(ms, _ip), n1out, n2out, outxscale,
                                             outyscale, outpos1, outpos2), "set_grid")

    def clear_grid(self):
        """Per-entry lookups from here on (the grids are forgotten)."""
        self._ck(self.lib.imcom_gpu_set_grid(self.h, 0, 0, 0, None, None, None, None, 0, 0, 0.0,
                                             0.0, 0.0, 0.0), "set_grid")

    def lattice_state(self):
        """(A, B): whether lookup_A / lookup_B take the grid lattices."""
        a, b = ctypes.c_int(0), ctypes.c_int(0)
        self._ck(self.lib.imcom_gpu_lattice_state(self.h, ctypes.byref(a), ctypes.byref(b)),
                 "lattice_state")
        return bool(a.value), bool(b.value)

    # ---- lookup-table set-up (imcom_proc.f90:184-250,433-475, imcom_matrix.f90:361-552)
    def assemble_inputs(self, I_unmasked, saturation, gimxscale, gimyscale, dither):
        """Masked input vectors (imcom_io.f90:84-123, imcom_proc.f90:56-89):
        returns x_i, y_i, exp_i, I_i."""
        Im = np.asfortranarray(I_unmasked, dtype=np.float64)
        n1, n2, ne = Im.shape
        sx, sy = _vec(gimxscale), _vec(gimyscale)
        d = np.asfortranarray(dither, dtype=np.float64)
        tot = n1 * n2 * ne
        x, y, I = np.empty(tot), np.empty(tot), np.empty(tot)
        e = np.empty(tot, dtype=np.int32)
        n = _i64(0)
        self._ck(self.lib.imcom_gpu_assemble_inputs(self.h, n1, n2, ne, _ptr(Im), float(saturation),
                                                    _ptr(sx), _ptr(sy), _ptr(d), ctypes.byref(n),
                                                    _ptr(x), _ptr(y), _ptr(e, _ip), _ptr(I)),
                 "assemble_inputs")
        k = n.value
        return x[:k].copy(), y[:k].copy(), e[:k].copy(), I[:k].copy()

    def rotate_image(self, im, theta_deg, npoly=7, conserve_flux=1):
        """imcom_rotate_image on the device."""
        a = np.asfortranarray(im, dtype=np.float64)
        out = np.empty(a.shape, order="F")
        self._ck(self.lib.imcom_gpu_rotate_image(self.h, a.shape[0], a.shape[1], _ptr(a),
                                                 float(theta_deg), npoly, conserve_flux,
                                                 _ptr(out)), "rotate_image")
        return out

    def ft_r2c(self, imgs):
        """Full complex DFT of real square images (n x n or n x n x k,
        Fortran order) as imcom_ft_r2c returns it; complex128 of the same shape."""
        a = np.asfortranarray(imgs, dtype=np.float64)
        k = 1 if a.ndim == 2 else a.shape[2]
        out = np.empty(a.shape, dtype=np.complex128, order="F")
        self._ck(self.lib.imcom_gpu_ft_r2c(self.h, a.shape[0], a.shape[1], k, _ptr(a),
                                           out.ctypes.data_as(_dp)), "ft_r2c")
        return out

    @staticmethod
    def _spectra(Gt_rot):
        Gs = [np.asarray(g) for g in Gt_rot] if isinstance(Gt_rot, (list, tuple)) else \
            [Gt_rot[:, :, i] for i in range(Gt_rot.shape[2])] if np.ndim(Gt_rot) == 3 else [Gt_rot]
        return np.asfortranarray(np.stack(Gs, axis=2).astype(np.complex128)), len(Gs)

    def build_Alookup(self, psfconst, Gt_rot, npad=3, out=False):
        """imcom_build_Alookup on the device; the tables stay there for lookup_A."""
        S, k = self._spectra(Gt_rot)
        n = S.shape[0]
        A = np.empty((n * npad, n * npad, k * (k + 1) // 2), order="F") if out else None
        self._ck(self.lib.imcom_gpu_build_Alookup(self.h, psfconst, k, n, S.shape[1], npad,
                                                  S.ctypes.data_as(_dp), _ptr(A)), "build_Alookup")
        return A

    def build_Blookup(self, psfconst, Gammat, Gt_rot, npad=3, out=False):
        """imcom_build_Blookup on the device; the tables stay there for lookup_B."""
        S, k = self._spectra(Gt_rot)
        G = np.asfortranarray(Gammat, dtype=np.complex128)
        n = S.shape[0]
        B = np.empty((n * npad, n * npad, k), order="F") if out else None
        self._ck(self.lib.imcom_gpu_build_Blookup(self.h, psfconst, k, n, S.shape[1], npad,
                                                  G.ctypes.data_as(_dp), S.ctypes.data_as(_dp),
                                                  _ptr(B)), "build_Blookup")
        return B

    def calc_C(self, Gammat):
        """imcom_calc_C on the device."""
        G = np.asfortranarray(Gammat, dtype=np.complex128)
        C = np.zeros(1)
        self._ck(self.lib.imcom_gpu_calc_C(self.h, G.shape[0], G.shape[1], G.ctypes.data_as(_dp),
                                           _ptr(C)), "calc_C")
        return float(C[0])

    # ---- imcom_matrix ---------------------------------------------------
    def lookup_A(self, npoly=7, npad=3, out=False):
        t = self.tile
        A = np.empty((self.n, self.n), order="F") if out else None
        self._ck(self.lib.imcom_gpu_lookup_A(self.h, npoly, npad, t.psfxscale, t.psfyscale,
                                             _ptr(A)), "lookup_A")
        return A

    def set_A(self, A):
        A = _mat(A)
        self._ck(self.lib.imcom_gpu_set_A(self.h, _ptr(A)), "set_A")

    def eigen_A(self, out=True, out_Q=None, out_L=None):
        n = self.n
        out_Q = out if out_Q is None else out_Q
        out_L = out if out_L is None else out_L
        Q = np.empty((n, n), order="F") if out_Q else None
        L = np.empty(n) if out_L else None
        cond = ctypes.c_double(0.0)
        self._ck(self.lib.imcom_gpu_eigen_A(self.h, _ptr(Q), _ptr(L), ctypes.byref(cond)),
                 "eigen_A")
        return Q, L, cond.value

    def tridiagonalize(self):
        n = self.n
        d = np.empty(n)
        e = np.empty(max(n - 1, 1))
        self._ck(self.lib.imcom_gpu_tridiagonalize(self.h, _ptr(d), _ptr(e)), "tridiagonalize")
        return d, e[:n - 1]

    def tridiag_eig(self, d, e):
        d = _vec(d)
        e = _vec(e if len(e) else np.zeros(1))
        n = d.shape[0]
        lam = np.empty(n)
        Z = np.empty((n, n), order="F")
        self._ck(self.lib.imcom_gpu_tridiag_eig(self.h, n, _ptr(d), _ptr(e), _ptr(lam), _ptr(Z)),
                 "tridiag_eig")
        return lam, Z

    def set_QL(self, Q, L):
        Q, L = _mat(Q), _vec(L)
        self._ck(self.lib.imcom_gpu_set_QL(self.h, _ptr(Q), _ptr(L)), "set_QL")

    def lookup_B(self, npoly=7, npad=3, out=False):
        t = self.tile
        B = np.empty((self.n, self.m), order="F") if out else None
        self._ck(self.lib.imcom_gpu_lookup_B(self.h, npoly, npad, t.psfxscale, t.psfyscale,
                                             _ptr(B)), "lookup_B")
        return B

    def set_B(self, B):
        B = _mat(B)
        self._ck(self.lib.imcom_gpu_set_B(self.h, _ptr(B)), "set_B")

    def build_P(self, precision=IMCOM_PREC_FP64, out=False):
        P = np.empty((self.n, self.m), order="F") if out else None
        self._ck(self.lib.imcom_gpu_build_P(self.h, precision, _ptr(P)), "build_P")
        return P

    def set_P(self, P):
        P = _mat(P)
        self._ck(self.lib.imcom_gpu_set_P(self.h, _ptr(P)), "set_P")

    # ---- stage files ----------------------------------------------------
    def fetch_async(self, what, out=None):
        """Start copying the stage output `what` ("A", "Q", "L", "B", "P"
        or "T"), as it stands after the work issued so far, into `out` (a
        new page-locked array by default) and return that array; it holds
        the output once fetch_wait(what) has returned.  Later stages that
        overwrite the device buffer wait for the copy on the device."""
        n, m = self.n, self.m
        shape = {"A": (n, n), "Q": (n, n), "L": (n,), "B": (n, m), "P": (n, m), "T": (n, m)}[what]
        if out is None:
            out = host_empty(shape)
        elif out.shape != shape or out.dtype != np.float64 or not out.flags.f_contiguous:
            raise ValueError(f"fetch_async: out must be a Fortran-ordered float64 {shape} array")
        self._ck(self.lib.imcom_gpu_fetch_async(self.h, FETCH[what], _ptr(out)), f"fetch_async({what})")
        # the copy writes into `out` until fetch_wait: keep it alive until then
        self._inflight.setdefault(what, []).append(out)
        return out

    def fetch_wait(self, what=None):
        """Wait for the copy of `what` (every output when None)."""
        self._ck(self.lib.imcom_gpu_fetch_wait(self.h, -1 if what is None else FETCH[what]),
                 f"fetch_wait({what})")
        if what is None:
            self._inflight.clear()
        else:
            self._inflight.pop(what, None)

    # ---- imcom_bisect ---------------------------------------------------
    def bisect_kappa(self, UnotS=None, US=None, k_min=None, k_max=None, tol=None, maxnits=None,
                     bracket=None):
        t = self.tile
        UnotS = t.USB if UnotS is None else UnotS
        US = t.US_max if US is None else US
        k_min = t.kappa_min if k_min is None else k_min
        k_max = t.kappa_max if k_max is None else k_max
        tol = t.US_tol if tol is None else tol
        maxnits = t.maxNbis if maxnits is None else maxnits
        m = self.m
        K, USa, ev = np.empty(m), np.empty(m), np.empty(m, dtype=np.int32)
        br = None if bracket is None else np.ascontiguousarray(bracket, dtype=np.float32)
        self._ck(self.lib.imcom_gpu_bisect_kappa(self.h, 1 if UnotS else 0, US, k_min, k_max, tol,
                                                 maxnits, _ptr(br, _fp), _ptr(K), _ptr(USa),
                                                 _ptr(ev, _ip)), "bisect_kappa")
        return K, USa, ev

    def build_T(self, precision=IMCOM_PREC_FP64, out=True):
        T = np.empty((self.n, self.m), order="F") if out else None
        self._ck(self.lib.imcom_gpu_build_T(self.h, precision, _ptr(T)), "build_T")
        return T

    def set_T(self, T):
        """imcom_get_T cache hit: T read from TFILE (imcom_bisect.f90:47-62)."""
        T = _mat(T)
        self._ck(self.lib.imcom_gpu_set_T(self.h, _ptr(T)), "set_T")

    def build_S(self):
        S = np.empty(self.m)
        self._ck(self.lib.imcom_gpu_build_S(self.h, _ptr(S)), "build_S")
        return S

    def build_U(self):
        U = np.empty(self.m)
        self._ck(self.lib.imcom_gpu_build_U(self.h, _ptr(U)), "build_U")
        return U

    def build_image(self, I_i=None):
        I_i = _vec(self.tile.I_i if I_i is None else I_i)
        H = np.empty(self.m)
        self._ck(self.lib.imcom_gpu_build_image(self.h, _ptr(I_i), _ptr(H)), "build_image")
        return H

    def set_profiling(self, level):
        self._ck(self.lib.imcom_gpu_set_profiling(self.h, level), "set_profiling")

    def stage_times(self):
        k = len(STAGE_SLOTS)
        ms = np.zeros(k)
        cnt = np.zeros(k, dtype=np.int64)
        self._ck(self.lib.imcom_gpu_stage_times(self.h, _ptr(ms), cnt.ctypes.data_as(ctypes.POINTER(_i64)), k),
                 "stage_times")
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(STAGE_SLOTS)}

    def set_tuning(self, nb=32, nbb=128):
        self._ck(self.lib.imcom_gpu_set_tuning(self.h, nb, nbb), "set_tuning")

    def set_reduction_tail(self, max_rows=2048):
        """Largest trailing size whose columns the persistent kernel reduces
        (the library's default 2048; 0: every column on the blocked kernels)."""
        self._ck(self.lib.imcom_gpu_set_reduction_tail(self.h, max_rows), "set_reduction_tail")

    def set_reduction_stages(self, stages=REDUCTION_STAGES, min_n=TWOSTAGE_MIN):
        """Reduction to tridiagonal form: 1 the one-stage blocked kernels, 2
        the two-stage reduction (dense -> band -> tridiagonal) for n >= min_n
        (the library's default: 2 from n = 7168 up, where it is faster)."""
        self._ck(self.lib.imcom_gpu_set_reduction_stages(self.h, stages, min_n),
                 "set_reduction_stages")

    def band_reduce(self):
        """First stage of the two-stage reduction alone (A consumed): the n x n
        result, band matrix in the lower triangle, the panels' Householder
        vectors in the upper."""
        out = np.empty((self.n, self.n), order="F")
        self._ck(self.lib.imcom_gpu_band_reduce(self.h, _ptr(out)), "band_reduce")
        return out

    # ---- batched tiles ----------------------------------------------------
    def batch_set(self, tiles):
        """Upload many tiles (each a tile.Tile) for one batched solve."""
        self._batch_keep = []
        descs = (TileDesc * len(tiles))()
        for d, t in zip(descs, tiles):
            arrs = dict(x_i=_vec(t.x_i), y_i=_vec(t.y_i), exp_i=_vec(t.exp_i, np.int32),
                        X_a=_vec(t.X_a), Y_a=_vec(t.Y_a), Ndiag=_vec(t.Ndiag), I_i=_vec(t.I_i),
                        Alookup=_mat(t.Alookup), Blookup=_mat(t.Blookup))
            self._batch_keep.append(arrs)
            for k, a in arrs.items():
                setattr(d, k, a.ctypes.data)
            d.n, d.m = t.n, t.m
            d.C_a, d.kappa_min, d.kappa_max = t.C_a, t.kappa_min, t.kappa_max
            d.US_max, d.US_tol = t.US_max, t.US_tol
            d.UnotS, d.maxnits = (1 if t.USB else 0), t.maxNbis
            d.psfconst, d.nexppsf = t.psfconst, t.nexppsf
            d.n1, d.n2 = t.Alookup.shape[0], t.Alookup.shape[1]
            d.npoly, d.npad = t.npoly, t.npad
            d.psfxscale, d.psfyscale = t.psfxscale, t.psfyscale
        self._ck(self.lib.imcom_gpu_batch_set(self.h, len(tiles), descs), "batch_set")
        self.batch_tiles = list(tiles)

    def batch_solve(self, precision_P=IMCOM_PREC_FP64, precision_T=IMCOM_PREC_FP64,
                    learned=False, zsig=3.0, use_graph=True):
        self._ck(self.lib.imcom_gpu_batch_solve(self.h, precision_P, precision_T,
                                                1 if learned else 0, zsig, 1 if use_graph else 0),
                 "batch_solve")

    def batch_reserve(self, precision_P=IMCOM_PREC_FP64, precision_T=IMCOM_PREC_FP64):
        """Allocate batch_solve's workspace for these precisions now."""
        self._ck(self.lib.imcom_gpu_batch_reserve(self.h, precision_P, precision_T),
                 "batch_reserve")

    def batch_result(self, i, want_T=False):
        t = self.batch_tiles[i]
        m = t.m
        K, US, SU, H = (np.empty(m) for _ in range(4))
        ev, hits = np.empty(m, dtype=np.int32), np.empty(m, dtype=np.int32)
        T = np.empty((t.n, m), order="F") if want_T else None
        self._ck(self.lib.imcom_gpu_batch_result(self.h, i, _ptr(K), _ptr(US), _ptr(SU), _ptr(H),
                                                 _ptr(ev, _ip), _ptr(hits, _ip), _ptr(T)),
                 "batch_result")
        U, S = (US, SU) if t.USB else (SU, US)
        return dict(K=K, U=U, S=S, H=H, evals=ev, hits=hits, T=T)

    # ---- learned kappa bracket --------------------------------------------
    def mlp_nparams(self):
        return int(self.lib.imcom_gpu_mlp_nparams())

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape[0] == self.mlp_nparams()
        self._ck(self.lib.imcom_gpu_mlp_set_params(self.h, _ptr(p, _fp)), "mlp_set_params")

    def mlp_get_params(self):
        p = np.empty(self.mlp_nparams(), dtype=np.float32)
        self._ck(self.lib.imcom_gpu_mlp_get_params(self.h, _ptr(p, _fp)), "mlp_get_params")
        return p

    def kappa_features(self, US_target=None):
        t = self.tile
        US_target = t.US_max if US_target is None else US_target
        X = np.empty((self.m, 32), dtype=np.float32)
        self._ck(self.lib.imcom_gpu_kappa_features(self.h, US_target, _ptr(X, _fp)), "kappa_features")
        return X

    def mlp_infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty((X.shape[0], 2), dtype=np.float32)
        self._ck(self.lib.imcom_gpu_mlp_infer(self.h, _ptr(X, _fp), X.shape[0], _ptr(out, _fp)),
                 "mlp_infer")
        return out

    def mlp_grad(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(6402, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        self._ck(self.lib.imcom_gpu_mlp_grad(self.h, _ptr(X, _fp), _ptr(y, _fp), X.shape[0],
                                             _ptr(g, _fp), ctypes.byref(loss)), "mlp_grad")
        return g, float(loss.value)

    def mlp_train(self, X, y, epochs=10, batch=1024, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        self._ck(self.lib.imcom_gpu_mlp_train(self.h, _ptr(X, _fp), _ptr(y, _fp), X.shape[0], epochs,
                                              batch, lr, seed, _ptr(losses, _fp)), "mlp_train")
        return losses[:epochs]

    def bisect_kappa_learned(self, zsig=3.0, UnotS=None, US=None, k_min=None, k_max=None,
                             tol=None, maxnits=None):
        t = self.tile
        UnotS = t.USB if UnotS is None else UnotS
        US = t.US_max if US is None else US
        k_min = t.kappa_min if k_min is None else k_min
        k_max = t.kappa_max if k_max is None else k_max
        tol = t.US_tol if tol is None else tol
        maxnits = t.maxNbis if maxnits is None else maxnits
        m = self.m
        K, USa = np.empty(m), np.empty(m)
        ev, hits = np.empty(m, dtype=np.int32), np.empty(m, dtype=np.int32)
        self._ck(self.lib.imcom_gpu_bisect_kappa_learned(self.h, 1 if UnotS else 0, US, k_min, k_max,
                                                         tol, maxnits, zsig, _ptr(K), _ptr(USa),
                                                         _ptr(ev, _ip), _ptr(hits, _ip)),
                 "bisect_kappa_learned")
        return K, USa, ev, hits

    def bisect_kappa_learned_noalloc(self, n, m, L, P, ldp, Ndiag, C_a, params, K, US,
                                     evals=0, hits=0, stream=0, zsig=3.0, UnotS=True,
                                     US_max=1e-6, k_min=None, k_max=None, tol=1e-9, maxnits=53):
        """imcom_gpu_bisect_kappa_learned_noalloc: every array argument is a
        device address (int), as from torch's data_ptr(); the call only
        enqueues the work (on `stream`, 0 for the context's stream)."""
        k_min = C_a * 2.0 ** -53 if k_min is None else k_min
        k_max = C_a * 2.0 ** 53 if k_max is None else k_max
        self._ck(self.lib.imcom_gpu_bisect_kappa_learned_noalloc(
            self.h, stream or None, n, m, L, P, ldp, Ndiag or None, C_a, params,
            1 if UnotS else 0, US_max, k_min, k_max, tol, maxnits, zsig, K, US, evals or None,
            hits or None), "bisect_kappa_learned_noalloc")

    def synchronize(self):
        self._ck(self.lib.imcom_gpu_synchronize(self.h), "synchronize")

    # ---- the driver's sequence (imcom.f90:151-169) ----------------------
    def solve(self, tile=None, precision_P=IMCOM_PREC_FP64, precision_T=IMCOM_PREC_FP64,
              bracket=None, want_T=True):
        """get_A, get_QL, get_B, get_P, get_T (bisect, T, S or U), build_image."""
        if tile is not None:
            self.set_tile(tile)
        t = self.tile
        self.lookup_A(t.npoly, t.npad)
        _, L, cond = self.eigen_A(out=False)
        self.lookup_B(t.npoly, t.npad)
        self.build_P(precision_P)
        K, US, ev = self.bisect_kappa(bracket=bracket)
        T = self.build_T(precision_T, out=want_T)
        if t.USB:
            U, S = US, self.build_S()
        else:
            S, U = US, self.build_U()
        H = self.build_image()
        return dict(K=K, U=U, S=S, T=T, H=H, evals=ev, cond=cond)
