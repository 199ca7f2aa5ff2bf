"""The reference's driver program (imcom.f90) over the GPU solve path.

    python -m imcom_amd.driver <config_file> <U|S> <US_max> <US_tol> [forceT [forceSys [saturation]]]
    python -m imcom_amd.driver --batch <U|S> <US_max> <US_tol> <config_file> [<config_file> ...]

The first form takes the reference's command line (imcom_get_cmdline,
imcom_io.f90:1006-1051) and configuration files (imcom_read_config,
imcom_read_inconfigs, imcom_read_outconfig, imcom_io.f90:499-971), reads the
input, PSF and Gamma images from FITS, runs the tile set-up (masking,
coordinates, padding, rotation, spectra, lookup tables: tile.assemble_tile
with the lookup tables built on the device) and the solve path through
ImcomGPU, and writes the reference's outputs: H (OUTFILE), K, T, U and S.
The A, B, Q/L and P stage files are read back when they exist and forceSys
is 0 and written when their name is not None (imcom_get_A/QL/B/P,
imcom_matrix.f90:41-202).  An existing TFILE with forceT = 0 is read back
instead of solved (imcom_get_T, imcom_bisect.f90:47-64): then only H is
formed and written, and the system (A, Q/L, B, P) is not built unless
forceSys is set (imcom.f90:146-160), as in the reference.

The second form solves several tiles (one configuration file each) as one
batched launch sequence (imcom_gpu_batch_*) and writes every tile's outputs;
it does not read or write the stage files.

File names inside a configuration file are relative to the working directory,
as in the reference; in the batched form, to that configuration file's own
directory (as if the reference ran once from each).
"""
from __future__ import annotations

import dataclasses
import os
import sys

import numpy as np

from . import fits
from . import tile as TL

DEFAULT_STAGE_FILES = {"A": "A.fits", "B": "B.fits", "Q": "Q.fits", "L": "L.fits", "P": "P.fits"}


@dataclasses.dataclass
class Exposure:
    psffile: str
    gimfile: str
    rotangdeg: float
    noise: float
    gimxscale: float = 1.0
    gimyscale: float = 1.0
    dither: tuple = (0.0, 0.0)
    gimxfile: str = ""
    gimyfile: str = ""


@dataclasses.dataclass
class Config:
    psfxscale: float
    psfyscale: float
    nexp: int
    userxy: int
    exposures: list
    gamfile: str
    Hfile: str
    Kfile: str
    Tfile: str
    Sfile: str
    Ufile: str
    outxscale: float = 0.0
    outyscale: float = 0.0
    outpos: tuple = (0.0, 0.0)
    nout: tuple = (0, 0)
    outxfile: str = ""
    outyfile: str = ""
    stage_files: dict = dataclasses.field(default_factory=lambda: dict(DEFAULT_STAGE_FILES))


class ConfigError(ValueError):
    pass


def _list_directed(line: str, what: str, n: int, conv=float):
    """A list-directed read of 'KEY v1 [v2]' (FMT=*): the values after the
    first token.  Fortran D exponents are accepted."""
    tok = line.split()
    if len(tok) < 1 + n:
        raise ConfigError(f"IMCOM ERROR: Expecting: {what}")
    try:
        vals = [conv(v.replace("D", "E").replace("d", "e")) if conv is float else conv(v)
                for v in tok[1:1 + n]]
    except ValueError as e:
        raise ConfigError(f"IMCOM ERROR: Expecting: {what}") from e
    return vals if n > 1 else vals[0]


def _a12(line: str) -> str:
    """FMT='(A12, A)': a 12-character field, then the file name."""
    return line[12:].strip()


def _lines(path: str):
    try:
        with open(path) as f:
            return [ln.rstrip("\n") for ln in f]
    except OSError as e:
        raise ConfigError(f"IMCOM ERROR: Cannot open {path}") from e


def read_config(path: str) -> Config:
    """imcom_read_config + imcom_read_inconfigs + imcom_read_outconfig."""
    lines = [ln for ln in _lines(path)]
    it = iter(lines)

    def nxt(what):
        try:
            return next(it)
        except StopIteration:
            raise ConfigError(f"IMCOM ERROR: Incorrect line format for {path}: expecting {what}")
    psfxscale = _list_directed(nxt("PSFXSCALE"), "PSFXSCALE  <psfxscale>", 1)
    psfyscale = _list_directed(nxt("PSFYSCALE"), "PSFYSCALE  <psfyscale>", 1)
    nexp = int(_list_directed(nxt("NEXP"), "NEXP      <integer>", 1, int))
    userxy = int(_list_directed(nxt("USERXY"), "USERXY    <integer>", 1, int))
    inconfigs = [_a12(nxt("INCONFIG  <filename>")) for _ in range(nexp)]
    exps = [_read_inconfig(f, userxy) for f in inconfigs]
    out = _read_outconfig(_a12(nxt("OUTCONFIG <filename>")), userxy)
    stage = dict(DEFAULT_STAGE_FILES)
    rest = [ln for ln in it]
    for key, ln in zip(("A", "B", "Q", "L", "P"), rest):   # optional; defaults at EOF
        stage[key] = _a12(ln)
    return Config(psfxscale=psfxscale, psfyscale=psfyscale, nexp=nexp, userxy=userxy,
                  exposures=exps, stage_files=stage, **out)


def _read_inconfig(path: str, userxy: int) -> Exposure:
    ls = _lines(path)
    if len(ls) < (6 if userxy == 1 else 7):
        raise ConfigError(f"IMCOM ERROR: Incorrect line format for {path}")
    e = Exposure(psffile=_a12(ls[0]), gimfile=_a12(ls[1]),
                 rotangdeg=_list_directed(ls[2], "ROTANGDEG <float>", 1),
                 noise=_list_directed(ls[3], "NOISE     <float>", 1))
    if userxy == 1:
        e.gimxfile, e.gimyfile = _a12(ls[4]), _a12(ls[5])
    else:
        e.gimxscale = _list_directed(ls[4], "GIMXSCALE <float>", 1)
        e.gimyscale = _list_directed(ls[5], "GIMYSCALE <float>", 1)
        e.dither = tuple(_list_directed(ls[6], "DITHER    <float> <float>", 2))
    return e


def _read_outconfig(path: str, userxy: int) -> dict:
    ls = _lines(path)
    if len(ls) < (8 if userxy == 1 else 10):
        raise ConfigError(f"IMCOM ERROR: Incorrect line format for {path}")
    d = dict(gamfile=_a12(ls[0]), Hfile=_a12(ls[1]), Kfile=_a12(ls[2]), Tfile=_a12(ls[3]),
             Sfile=_a12(ls[4]), Ufile=_a12(ls[5]))
    if userxy == 1:
        d.update(outxfile=_a12(ls[6]), outyfile=_a12(ls[7]))
    else:
        d.update(outxscale=_list_directed(ls[6], "OUTXSCALE <float>", 1),
                 outyscale=_list_directed(ls[7], "OUTYSCALE <float>", 1),
                 outpos=tuple(_list_directed(ls[8], "OUTPOS    <float> <float>", 2)),
                 nout=tuple(_list_directed(ls[9], "NOUT      <int> <int>", 2, int)))
    return d


def build_tile(cfg: Config, USB: bool, US_max: float, US_tol: float,
               saturation: float = np.finfo(np.float64).max, npoly: int = 7, npad: int = 3,
               host_tables: bool = False):
    """The set-up of imcom.f90:55-127 for one configuration."""
    ims = [fits.readfits(e.gimfile) for e in cfg.exposures]
    if any(im.shape != ims[0].shape for im in ims):
        raise ConfigError("IMCOM ERROR: Variable size input images not supported in this version")
    I_unmasked = np.asfortranarray(np.stack(ims, axis=2))
    psfs = [fits.readfits(e.psffile) for e in cfg.exposures]
    if any(p.shape != psfs[0].shape for p in psfs):
        raise ConfigError("IMCOM ERROR: PSF images must all be the same size")
    gamma = fits.readfits(cfg.gamfile)
    if gamma.shape != psfs[0].shape:
        raise ConfigError("IMCOM ERROR: Gamma PSF must be the same size as the PSFs")
    kw = dict(I_unmasked=I_unmasked, psf_img=psfs, rotangdeg=np.array([e.rotangdeg for e in cfg.exposures]),
              gamma_img=gamma, psfxscale=cfg.psfxscale, psfyscale=cfg.psfyscale,
              noise=np.array([e.noise for e in cfg.exposures]), psffile=[e.psffile for e in cfg.exposures],
              npad=npad, npoly=npoly, saturation=saturation, USB=USB, US_max=US_max, US_tol=US_tol,
              host_tables=host_tables)
    if cfg.userxy == 1:
        xs = np.stack([fits.readfits(e.gimxfile) for e in cfg.exposures], axis=2)
        ys = np.stack([fits.readfits(e.gimyfile) for e in cfg.exposures], axis=2)
        if xs.shape != I_unmasked.shape or ys.shape != I_unmasked.shape:
            raise ConfigError("IMCOM ERROR: X/Y coordinate images must match the input images")
        Xa, Ya = fits.readfits(cfg.outxfile), fits.readfits(cfg.outyfile)
        if Xa.shape != Ya.shape:
            raise ConfigError("IMCOM ERROR: X_a array does not match Y_a array")
        kw.update(xy_unmasked=(xs, ys), XaYa_grid=(Xa, Ya))
    else:
        kw.update(gimxscale=np.array([e.gimxscale for e in cfg.exposures]),
                  gimyscale=np.array([e.gimyscale for e in cfg.exposures]),
                  dither=np.array([e.dither for e in cfg.exposures]), n1out=cfg.nout[0],
                  n2out=cfg.nout[1], outxscale=cfg.outxscale, outyscale=cfg.outyscale,
                  outpos1=cfg.outpos[0], outpos2=cfg.outpos[1])
    return TL.assemble_tile(**kw)


class _Writer:
    """FITS writes on a background thread, so that the multi-GB stage files
    (A, Q, B, P, T) go to disk while the device runs the next stages.  The
    arrays handed over are host copies the caller no longer touches."""

    def __init__(self):
        from concurrent.futures import ThreadPoolExecutor
        self.pool = ThreadPoolExecutor(max_workers=1)
        self.jobs = []

    def put(self, path, arr):
        self.jobs.append(self.pool.submit(fits.writefits, path, arr))

    def close(self):
        self.pool.shutdown(wait=True)
        for j in self.jobs:
            j.result()   # re-raise a failed write


def _cached(name: str, forceSys: int) -> bool:
    return name != "None" and forceSys == 0 and os.path.exists(name)


def run(config: str, mode: str, US_max: float, US_tol: float, forceT: int = 0, forceSys: int = 0,
        saturation: float = np.finfo(np.float64).max, device: int = 0, log=print):
    """One tile, as the reference's main program; returns the outputs.  The
    stage outputs are copied to page-locked host memory on a second stream
    while the next stages run (fetch_async), and the FITS files are written
    on a background thread."""
    from .gpu import ImcomGPU
    if mode not in ("U", "u", "S", "s"):
        raise ConfigError("IMCOM ERROR: mode must be U or S")
    USB = mode in ("U", "u")
    cfg = read_config(config)
    t = build_tile(cfg, USB, US_max, US_tol, saturation)
    sf = cfg.stage_files
    n, m = t.n, t.m
    w = _Writer()
    inflight = []  # (output, file name, host array, file shape): copies not yet handed over

    def fetch(g, what, name, shape):
        inflight.append((what, name, g.fetch_async(what), shape))

    def hand_over(g, *whats):
        """Wait for the named copies and queue their files."""
        for item in [f for f in inflight if f[0] in whats]:
            g.fetch_wait(item[0])
            w.put(item[1], item[2].reshape(item[3], order="F"))
            inflight.remove(item)

    Texists = os.path.exists(cfg.Tfile)
    read_T = Texists and forceT == 0                      # imcom_get_T's cache hit
    build_system = forceSys != 0 or forceT != 0 or not Texists   # imcom.f90:151
    K = U = S = ev = None
    try:
        with ImcomGPU(device) as g:
            g.set_tile(t, tables="device" if build_system else "none")
            if build_system:
                _system(g, t, sf, forceSys, fetch, hand_over, log)
            if read_T:
                log(f"IMCOM: Reading T matrix from {cfg.Tfile}")
                T = _sized(fits.readfits(cfg.Tfile), (n, m), cfg.Tfile)
                g.set_T(T)
                H = g.build_image()
                hand_over(g, "A", "Q", "L", "B", "P")
            else:
                K, US, ev = g.bisect_kappa()
                hand_over(g, "Q", "L", "B")
                g.build_T(out=False)
                T = g.fetch_async("T")
                if USB:
                    U, S = US, g.build_S()
                else:
                    S, U = US, g.build_U()
                H = g.build_image()
                hand_over(g, "A", "P")
                g.fetch_wait("T")
        grid = (t.n1out, t.n2out)
        if not read_T:
            w.put(cfg.Kfile, K.reshape(grid, order="F"))
            w.put(cfg.Ufile, U.reshape(grid, order="F"))
            w.put(cfg.Sfile, S.reshape(grid, order="F"))
            w.put(cfg.Tfile, T)
        w.put(cfg.Hfile, H.reshape(grid, order="F"))
    finally:
        w.close()
    if read_T:
        log(f"IMCOM: wrote {cfg.Hfile} (T read from {cfg.Tfile})")
    else:
        log(f"IMCOM: wrote {cfg.Hfile}, {cfg.Kfile}, {cfg.Ufile}, {cfg.Sfile}, {cfg.Tfile}")
    return dict(K=K, U=U, S=S, T=T, H=H, evals=ev, tile=t, T_read=read_T)


def _system(g, t, sf, forceSys, fetch, hand_over, log):
    """imcom_get_A / imcom_get_QL / imcom_get_B / imcom_get_P
    (imcom_matrix.f90:41-202); each stage file is copied out on a second
    stream while the next stages run."""
    n, m = t.n, t.m
    if _cached(sf["A"], forceSys):
        log(f"IMCOM: Reading A matrix from {sf['A']}")
        g.set_A(_sized(fits.readfits(sf["A"]), (n, n), sf["A"]))
    else:
        g.lookup_A(t.npoly, t.npad)
        if sf["A"] != "None":
            fetch(g, "A", sf["A"], (n, n))
    if _cached(sf["Q"], forceSys) and _cached(sf["L"], forceSys):
        log(f"IMCOM: Reading Q, L from {sf['Q']}, {sf['L']}")
        g.set_QL(_sized(fits.readfits(sf["Q"]), (n, n), sf["Q"]),
                 _sized(fits.readfits(sf["L"]), (n, 1), sf["L"]).ravel(order="F"))
    else:
        _, _, cond = g.eigen_A(out_Q=False, out_L=False)
        log(f"IMCOM: A matrix condition number = {cond:24.16E}")
        if sf["Q"] != "None" and sf["L"] != "None":
            fetch(g, "Q", sf["Q"], (n, n))
            fetch(g, "L", sf["L"], (n, 1))
    hand_over(g, "A")  # the reduction waited for this copy before overwriting A
    if _cached(sf["B"], forceSys):
        log(f"IMCOM: Reading B matrix from {sf['B']}")
        g.set_B(_sized(fits.readfits(sf["B"]), (n, m), sf["B"]))
    else:
        g.lookup_B(t.npoly, t.npad)
        if sf["B"] != "None":
            fetch(g, "B", sf["B"], (n, m))
    if _cached(sf["P"], forceSys):
        log(f"IMCOM: Reading P matrix from {sf['P']}")
        g.set_P(_sized(fits.readfits(sf["P"]), (n, m), sf["P"]))
    else:
        g.build_P()
        if sf["P"] != "None":
            fetch(g, "P", sf["P"], (n, m))


class _InDir:
    """Work in a configuration file's directory, as a run of the reference
    from that directory would."""

    def __init__(self, path):
        self.path = os.path.dirname(os.path.abspath(path))

    def __enter__(self):
        self.old = os.getcwd()
        os.chdir(self.path)

    def __exit__(self, *exc):
        os.chdir(self.old)


def run_batch(configs, mode: str, US_max: float, US_tol: float, device: int = 0, log=print):
    """Several tiles (one configuration file each) in concurrent batched
    solves (imcom_amd.batch).  Each configuration's file names are taken
    relative to its own directory."""
    from .batch import StreamedBatch
    USB = mode in ("U", "u")
    cfgs, tiles = [], []
    for path in configs:
        with _InDir(path):
            cfgs.append(read_config(os.path.basename(path)))
            tiles.append(build_tile(cfgs[-1], USB, US_max, US_tol, host_tables=True))
    outs = []
    with StreamedBatch(tiles, device=device) as sb:
        sb.solve()
        for i, (path, c, t) in enumerate(zip(configs, cfgs, tiles)):
            r = sb.result(i, want_T=True)
            grid = (t.n1out, t.n2out)
            U, S = r["U"], r["S"]
            with _InDir(path):
                fits.writefits(c.Kfile, r["K"].reshape(grid, order="F"))
                fits.writefits(c.Ufile, U.reshape(grid, order="F"))
                fits.writefits(c.Sfile, S.reshape(grid, order="F"))
                fits.writefits(c.Tfile, r["T"])
                fits.writefits(c.Hfile, r["H"].reshape(grid, order="F"))
            outs.append(dict(K=r["K"], U=U, S=S, T=r["T"], H=r["H"], tile=t))
            log(f"IMCOM: tile {i}: wrote {c.Hfile}, {c.Kfile}, {c.Ufile}, {c.Sfile}, {c.Tfile}")
    return outs


def _sized(a, shape, name):
    if a.shape != shape:
        raise ConfigError(f"IMCOM ERROR: Dimensions of Matrix {name} do not match "
                          f"<config_file> image dimensions: should be {shape}")
    return a


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    usage = ("usage: python -m imcom_amd.driver <config_file> <U|S> <US_max> <US_tol> "
             "[forceT [forceSys [saturation]]]\n"
             "       python -m imcom_amd.driver --batch <U|S> <US_max> <US_tol> <config_file> ...")
    try:
        if argv and argv[0] == "--batch":
            if len(argv) < 5:
                print(usage)
                return 2
            run_batch(argv[4:], argv[1], float(argv[2]), float(argv[3]))
            return 0
        if len(argv) < 4:
            print(usage)
            return 2
        forceT = int(argv[4]) if len(argv) > 4 else 0
        forceSys = int(argv[5]) if len(argv) > 5 else 0
        sat = float(argv[6]) if len(argv) > 6 else np.finfo(np.float64).max
        run(argv[0], argv[1], float(argv[2]), float(argv[3]), forceT, forceSys, sat)
        return 0
    except ConfigError as e:
        print(str(e))
        return 1


if __name__ == "__main__":
    sys.exit(main())
