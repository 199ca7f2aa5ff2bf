"""Minimal FITS primary-image reader and writer for the reference's files
(imcom_sizefits / imcom_readfits / imcom_writefits, imcom_io.f90:339-456).

The reference reads and writes 2-D images through CFITSIO: NAXIS1 is the
first (fastest) Fortran axis, data are big-endian, and it writes BITPIX = -64.
Reading converts any BITPIX to float64 with BSCALE / BZERO applied, as
CFITSIO's FTGPVD does.  Arrays come back Fortran-ordered (naxis1, naxis2).
"""
from __future__ import annotations

import numpy as np

_BLOCK = 2880
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _cards(buf: bytes):
    for i in range(0, len(buf), 80):
        yield buf[i:i + 80].decode("ascii", errors="replace")


def _value(card: str):
    v = card[10:].split("/", 1)[0].strip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        try:
            return float(v.replace("D", "E"))
        except ValueError:
            return v.strip("'").strip()


def read_header(path: str) -> tuple[dict, int]:
    """Primary header keywords and the byte offset of the data."""
    hdr: dict = {}
    with open(path, "rb") as f:
        off = 0
        while True:
            block = f.read(_BLOCK)
            if len(block) < _BLOCK:
                raise ValueError(f"{path}: truncated FITS header")
            off += _BLOCK
            for card in _cards(block):
                key = card[:8].strip()
                if key == "END":
                    return hdr, off
                if card[8:10] == "= ":
                    hdr[key] = _value(card)


def sizefits(path: str) -> tuple[int, int, int]:
    """(naxis1, naxis2, bitpix) as imcom_sizefits returns them."""
    hdr, _ = read_header(path)
    return int(hdr.get("NAXIS1", 0)), int(hdr.get("NAXIS2", 1)), int(hdr["BITPIX"])


def readfits(path: str) -> np.ndarray:
    """The primary image as float64, Fortran-ordered (naxis1, naxis2)."""
    hdr, off = read_header(path)
    bitpix = int(hdr["BITPIX"])
    if bitpix not in _DTYPES:
        raise ValueError(f"{path}: unsupported BITPIX {bitpix}")
    naxis = int(hdr.get("NAXIS", 0))
    if naxis < 1 or naxis > 2:
        raise ValueError(f"{path}: expected a 1-D or 2-D image, NAXIS = {naxis}")
    n1 = int(hdr["NAXIS1"])
    n2 = int(hdr.get("NAXIS2", 1)) if naxis == 2 else 1
    data = np.fromfile(path, dtype=_DTYPES[bitpix], count=n1 * n2, offset=off)
    if data.size != n1 * n2:
        raise ValueError(f"{path}: truncated data")
    out = data.astype(np.float64)
    bscale, bzero = float(hdr.get("BSCALE", 1.0)), float(hdr.get("BZERO", 0.0))
    if bscale != 1.0 or bzero != 0.0:
        out = out * bscale + bzero
    return np.asfortranarray(out.reshape(n2, n1).T)


def _card(key: str, value, comment: str = "") -> bytes:
    if isinstance(value, bool):
        v = "T" if value else "F"
    elif isinstance(value, int):
        v = str(value)
    else:
        v = repr(value)
    s = f"{key:<8}= {v:>20}"
    if comment:
        s += f" /{comment}"
    return s[:80].ljust(80).encode("ascii")


def writefits(path: str, image: np.ndarray) -> None:
    """Write a 2-D float64 image (BITPIX = -64), as imcom_writefits does."""
    a = np.asarray(image, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    n1, n2 = a.shape
    hdr = b"".join([
        _card("SIMPLE", True, "image conforms to FITS standard"),
        _card("BITPIX", -64, "bits per data value"),
        _card("NAXIS", 2, "number of axes"),
        _card("NAXIS1", n1),
        _card("NAXIS2", n2),
        _card("EXTEND", True, "file may contain extensions"),
        b"END".ljust(80),
    ])
    hdr += b" " * (-len(hdr) % _BLOCK)
    data = np.ascontiguousarray(a.T).astype(">f8").tobytes()
    data += b"\0" * (-len(data) % _BLOCK)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(data)
