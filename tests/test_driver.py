"""The driver front-end over the reference's own example inputs
(tests/golden/example: the FITS images and configuration files of the
reference's example/ directory, copied as fixtures): FITS reading and
writing, the configuration parsers, the tile set-up, and on the GPU the whole
run against the oracle on the same tile."""
import os
import shutil

import numpy as np
import pytest

from imcom_amd import driver as D
from imcom_amd import fits

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "example")
EPS = 2.0 ** -53


def _workdir(tmp_path, userxy):
    """The reference runs from the directory holding the configuration files
    and (links to) the FITS images."""
    d = tmp_path / f"userxy{userxy}"
    d.mkdir()
    for f in os.listdir(os.path.join(EX, f"userxy{userxy}")):
        shutil.copy(os.path.join(EX, f"userxy{userxy}", f), d / f)
    for f in os.listdir(EX):
        if f.endswith(".fits"):
            os.symlink(os.path.join(EX, f), d / f)
    return d


def test_fits_round_trip_and_example_shapes(tmp_path):
    rng = np.random.default_rng(1)
    a = np.asfortranarray(rng.standard_normal((7, 5)))
    p = str(tmp_path / "a.fits")
    fits.writefits(p, a)
    assert os.path.getsize(p) % 2880 == 0
    b = fits.readfits(p)
    assert b.shape == (7, 5) and np.array_equal(a, b)
    assert fits.sizefits(p) == (7, 5, -64)
    psf = fits.readfits(os.path.join(EX, "psf.example.fits"))
    im = fits.readfits(os.path.join(EX, "test1.example.fits"))
    assert psf.shape == (256, 256) and im.shape == (32, 32)
    assert np.all(np.isfinite(psf)) and psf.sum() > 0
    # NAXIS1 is the first (fastest) axis: the x map grows along it
    x = fits.readfits(os.path.join(EX, "x1.example.fits"))
    assert np.all(np.diff(x[:, 0]) > 0) and np.allclose(np.diff(x[0, :]), 0)


@pytest.mark.parametrize("userxy", [0, 1])
def test_read_config_example(tmp_path, userxy, monkeypatch):
    d = _workdir(tmp_path, userxy)
    monkeypatch.chdir(d)
    cfg = D.read_config("config_example")
    assert (cfg.psfxscale, cfg.psfyscale, cfg.nexp, cfg.userxy) == (0.1, 0.1, 4, userxy)
    assert [e.gimfile for e in cfg.exposures] == [f"test{i}.example.fits" for i in range(1, 5)]
    assert all(e.psffile == "psf.example.fits" and e.rotangdeg == 0.0 and e.noise == 1.0
               for e in cfg.exposures)
    assert cfg.gamfile == "psf.example.fits" and cfg.Hfile == "H.example.fits"
    assert cfg.stage_files == {"A": "A.example.fits", "B": "B.example.fits",
                               "Q": "Q.example.fits", "L": "L.example.fits",
                               "P": "P.example.fits"}
    if userxy == 0:
        assert cfg.exposures[0].dither == (0.0, 0.0) and cfg.exposures[0].gimxscale == 1.0
        assert (cfg.outxscale, cfg.outyscale, cfg.outpos, cfg.nout) == (0.5, 0.5, (0.0, 0.0), (65, 65))
    else:
        assert cfg.exposures[2].gimxfile == "x3.example.fits"
        assert (cfg.outxfile, cfg.outyfile) == ("xout.example.fits", "yout.example.fits")


def test_build_tile_example(tmp_path, monkeypatch):
    """The example's set-up: 4 exposures of 32^2, a 65^2 output grid at 0.5
    px, equal PSF rotations (so psfconst = 0, imcom_proc.f90:593), and the
    PSF grid padded to 2 nint(32 / 0.1) = 640."""
    monkeypatch.chdir(_workdir(tmp_path, 0))
    cfg = D.read_config("config_example")
    t = D.build_tile(cfg, True, 1e-6, 1e-9)
    assert (t.n, t.m, t.psfconst, t.nexppsf) == (4096, 4225, 0, 4)
    assert t.meta["n1pad"] == 640 and t.Alookup is None and len(t.Gt_rot) == 4
    assert t.kappa_min == t.C_a * EPS and t.kappa_max == t.C_a / EPS
    assert np.array_equal(t.exp_i, np.repeat(np.arange(1, 5, dtype=np.int32), 1024))


def test_userxy1_tile_matches_userxy0(tmp_path, monkeypatch):
    """The example's coordinate maps are the USERXY=0 grid: both forms give
    the same tile vectors."""
    monkeypatch.chdir(_workdir(tmp_path, 0))
    t0 = D.build_tile(D.read_config("config_example"), True, 1e-6, 1e-9)
    monkeypatch.chdir(_workdir(tmp_path, 1))
    t1 = D.build_tile(D.read_config("config_example"), True, 1e-6, 1e-9)
    for k in ("x_i", "y_i", "X_a", "Y_a", "I_i", "Ndiag"):
        assert np.allclose(getattr(t0, k), getattr(t1, k), rtol=0, atol=1e-12), k
    assert t0.C_a == t1.C_a


def test_driver_usage_and_bad_config(tmp_path, capsys):
    assert D.main([]) == 2
    assert D.main([str(tmp_path / "missing"), "U", "1e-6", "1e-9"]) == 1
    assert "Cannot open" in capsys.readouterr().out


@pytest.mark.gpu
def test_driver_example_matches_oracle(tmp_path, monkeypatch):
    """The reference's example (USERXY=0) run end to end through the driver
    on the GPU, against the oracle's solve of the same tile."""
    from oracle import oracle as O
    from imcom_amd import tile as TL
    print("[driver] example userxy0", flush=True)
    d = _workdir(tmp_path, 0)
    monkeypatch.chdir(d)
    out = D.run("config_example", "U", 1e-6, 1e-9, log=lambda s: None)
    for f in ("H", "K", "T", "S", "U", "A", "B", "Q", "L", "P"):
        assert os.path.exists(f"{f}.example.fits"), f
    H = fits.readfits("H.example.fits")
    assert H.shape == (65, 65) and np.array_equal(H.ravel(order="F"), out["H"])
    t = out["tile"]
    # the oracle on the same tile, with the host set-up's tables
    t.Alookup = TL.build_Alookup(t.Gt_rot, t.npad)
    t.Blookup = TL.build_Blookup(t.Gammat, t.Gt_rot, t.npad)
    ref = O.solve(t)
    # A is ill-conditioned here (cond ~ 4e12), so U(kappa) carries ~1e-10 of
    # rounding and bisection decisions whose |U - U_max| falls below that
    # flip: the pixels then stop one or more steps apart, both within the
    # bisection tolerance of U_max.  T and H are compared where kappa agrees.
    eq = out["K"] == ref["K"]
    dU = np.abs(out["U"] - ref["U"])
    tol, Umax = t.US_tol, t.US_max
    # T where kappa agrees, per pixel against the first-order change of
    # T = -1/2 Q P/(L + kappa) under eigenvalue errors of n eps lambda_max:
    # at the floor pixels kappa = kappa_min = C eps and T rests on the
    # smallest eigenvalues, which that bound covers
    Lr, Pr = ref["L"], ref["P"]
    dl = t.n * EPS * 2 * np.max(np.abs(Lr))
    g = 0.5 * np.sum(np.abs(Pr[:, eq]) * dl / (Lr[:, None] + ref["K"][None, eq]) ** 2, axis=0)
    t_tol = 1e-10 * np.abs(ref["T"]).max() + 100.0 * g
    dT = np.abs(out["T"][:, eq] - ref["T"][:, eq]).max(axis=0)
    H_eq = np.abs(out["H"][eq] - ref["H"][eq]).max() / np.abs(ref["H"]).max()
    print(f"[driver] cond={ref['cond']:.3e} max|dU|={dU.max():.3e} K equal {eq.mean():.3f} "
          f"max dT/t_tol {np.max(dT / t_tol):.3e} H rel (kappa equal) {H_eq:.3e}", flush=True)
    assert dU.max() <= tol
    assert eq.mean() >= 0.85
    floor = ref["U"] >= Umax
    assert np.array_equal(floor, out["U"] >= Umax)     # the same floor pixels
    flip = ~eq & ~floor
    assert np.all(Umax - out["U"][flip] <= tol) and np.all(Umax - ref["U"][flip] <= tol)
    assert np.all(out["U"][flip] < Umax) and np.all(ref["U"][flip] < Umax)
    assert np.all(dT <= t_tol)
    assert H_eq <= 1e-8
    # a second run with forceT = 1 solves T again, reading A, Q, L, B and P
    # back from the stage files it wrote (imcom_get_A/QL/B/P cache hits), and
    # reproduces the first run exactly
    logs = []
    again = D.run("config_example", "U", 1e-6, 1e-9, forceT=1, log=logs.append)
    assert sum("Reading" in s for s in logs) == 4
    assert not again["T_read"]
    for k in ("K", "U", "S", "H"):
        assert np.array_equal(again[k], out[k]), k
    # a third run with forceT = 0 finds TFILE: imcom_get_T reads T instead of
    # solving and the system is not built (imcom.f90:151, imcom_bisect.f90:47-64);
    # only H is formed and written, equal to the first run's bit for bit
    stamps = {f: os.stat(f"{f}.example.fits").st_mtime_ns for f in "KUST"}
    os.remove("H.example.fits")
    logs = []
    third = D.run("config_example", "U", 1e-6, 1e-9, log=logs.append)
    assert third["T_read"] and third["K"] is None
    assert any("Reading T matrix" in s for s in logs)
    assert not any("Reading A" in s or "Reading Q" in s for s in logs)
    assert np.array_equal(third["H"], out["H"])
    assert np.array_equal(fits.readfits("H.example.fits").ravel(order="F"), out["H"])
    for f, t0 in stamps.items():
        assert os.stat(f"{f}.example.fits").st_mtime_ns == t0, f


@pytest.mark.gpu
def test_driver_example_noise_mode_matches_oracle(tmp_path, monkeypatch):
    """Mode S (imcom_bisect_by_S, imcom_bisect.f90:226-267; U from
    imcom_build_U, imcom_matrix.f90:795-835) on the example, noise target
    0.05 with tolerance 1e-6: every pixel's bisection ends inside the
    target, away from both ends of the kappa range."""
    from oracle import oracle as O
    from imcom_amd import tile as TL
    print("[driver] example userxy0, mode S", flush=True)
    monkeypatch.chdir(_workdir(tmp_path, 0))
    Smax, tol = 0.05, 1e-6
    out = D.run("config_example", "S", Smax, tol, log=lambda s: None)
    assert np.array_equal(fits.readfits("S.example.fits").ravel(order="F"), out["S"])
    t = out["tile"]
    assert not t.USB and t.US_max == Smax
    t.Alookup = TL.build_Alookup(t.Gt_rot, t.npad)
    t.Blookup = TL.build_Blookup(t.Gammat, t.Gt_rot, t.npad)
    ref = O.solve(t)
    eq = out["K"] == ref["K"]
    dS = np.abs(out["S"] - ref["S"])
    print(f"[driver] mode S: K equal {eq.mean():.3f} max|dS|={dS.max():.3e} "
          f"S in [{out['S'].min():.6f}, {out['S'].max():.6f}]", flush=True)
    assert np.all(out["S"] < Smax) and np.all(Smax - out["S"] <= tol)
    assert np.all(ref["S"] < Smax) and np.all(Smax - ref["S"] <= tol)
    assert dS.max() <= tol
    assert eq.mean() >= 0.85
    assert np.all((out["K"] > t.kappa_min) & (out["K"] < t.kappa_max))
    # U recomputed at the pixel's kappa, and H, where kappa agrees
    assert np.max(np.abs(out["U"][eq] - ref["U"][eq])) <= 1e-9
    H_eq = np.abs(out["H"][eq] - ref["H"][eq]).max() / np.abs(ref["H"]).max()
    assert H_eq <= 1e-8


@pytest.mark.gpu
def test_driver_batch_matches_single_tile(tmp_path, monkeypatch):
    """The batched front-end on the example's two forms (USERXY=0 and 1, the
    same tile) gives the single-tile driver's kappa, U and H."""
    print("[driver] batch of the two example forms", flush=True)
    d0, d1 = _workdir(tmp_path, 0), _workdir(tmp_path, 1)
    outs = D.run_batch([str(d0 / "config_example"), str(d1 / "config_example")], "U", 1e-6, 1e-9,
                       log=lambda s: None)
    assert os.path.exists(d0 / "H.example.fits") and os.path.exists(d1 / "K.example.fits")
    monkeypatch.chdir(tmp_path)
    d2 = tmp_path / "single"
    d2.mkdir()
    for f in os.listdir(d0):
        if f.startswith("config_example"):
            shutil.copy(d0 / f, d2 / f)
        elif f.endswith(".example.fits") and f[0] not in "HKTSUABQLP":
            os.symlink(os.path.realpath(d0 / f), d2 / f)
    monkeypatch.chdir(d2)
    single = D.run("config_example", "U", 1e-6, 1e-9, log=lambda s: None)
    for o in outs:
        assert np.mean(o["K"] == single["K"]) >= 0.9
        assert np.max(np.abs(o["U"] - single["U"])) <= 1e-9
        assert np.max(np.abs(o["H"] - single["H"])) <= 1e-4 * np.max(np.abs(single["H"]))
