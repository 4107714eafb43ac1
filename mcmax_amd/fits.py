"""Minimal FITS (NASA/IAU standard 4.0) for the stage's output products:
primary and IMAGE-extension HDUs of 64-bit IEEE floats, header cards of 80
characters in 2880-byte blocks, big-endian data padded to 2880 bytes, and
gzip compression for names ending in ``.gz`` (the reference writes its
FITS products through CFITSIO, InputOutput.f:230-572, Subroutines.f:2422).

An array of numpy shape (n_k, ..., n_1) is stored with NAXIS1 = n_1 (the
fastest index), so a Fortran array a(i, j) of the reference is the numpy
array a[j, i].
"""
import gzip

import numpy as np

BLOCK = 2880


class FitsError(ValueError):
    pass


def _card(key, value=None, comment=""):
    key = key.upper()
    if len(key) > 8:
        raise FitsError(f"keyword longer than 8 characters: {key!r}")
    if value is None:
        return f"{key:<8}".ljust(80)
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, (int, np.integer)):
        v = f"{int(value):>20d}"
    elif isinstance(value, (float, np.floating)):
        v = f"{float(value):>20.12E}"
    else:
        s = str(value).replace("'", "''")
        v = f"'{s:<8}'".ljust(20)
    text = f"{key:<8}= {v}"
    if comment:
        text += f" / {comment}"
    if len(text) > 80:
        raise FitsError(f"card too long: {text!r}")
    return text.ljust(80)


def _header(cards):
    text = "".join(cards) + "END".ljust(80)
    pad = (-len(text)) % BLOCK
    return (text + " " * pad).encode("ascii")


def _data(a):
    raw = np.ascontiguousarray(a, dtype=">f8").tobytes()
    return raw + b"\0" * ((-len(raw)) % BLOCK)


def hdu_bytes(a, keys=(), primary=True):
    """One HDU: `keys` is a sequence of (keyword, value[, comment])."""
    a = np.asarray(a, dtype=np.float64)
    axes = list(reversed(a.shape))
    cards = [_card("SIMPLE", True, "file does conform to FITS standard")] if primary else \
        [_card("XTENSION", "IMAGE", "IMAGE extension")]
    cards.append(_card("BITPIX", -64, "number of bits per data pixel"))
    cards.append(_card("NAXIS", len(axes), "number of data axes"))
    for k, n in enumerate(axes, 1):
        cards.append(_card(f"NAXIS{k}", int(n), f"length of data axis {k}"))
    if primary:
        cards.append(_card("EXTEND", True, "FITS dataset may contain extensions"))
    else:
        cards.append(_card("PCOUNT", 0, "required keyword; must = 0"))
        cards.append(_card("GCOUNT", 1, "required keyword; must = 1"))
    for kv in keys:
        cards.append(_card(*kv))
    return _header(cards) + _data(a)


def write(path, arrays, primary_keys=()):
    """Write arrays[0] as the primary HDU (with primary_keys) and the rest as
    IMAGE extensions."""
    out = hdu_bytes(arrays[0], primary_keys, primary=True)
    for a in arrays[1:]:
        out += hdu_bytes(a, primary=False)
    if path.endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(out)
    else:
        with open(path, "wb") as f:
            f.write(out)


def _parse_value(v):
    v = v.strip()
    if v.startswith("'"):
        end = v.find("'", 1)
        while end != -1 and end + 1 < len(v) and v[end + 1] == "'":
            end = v.find("'", end + 2)
        return v[1:end].replace("''", "'").rstrip()
    v = v.split("/")[0].strip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        return float(v.replace("D", "E"))


def read(path):
    """Read every HDU of 64-bit floats: a list of (header dict, array)."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        raw = f.read()
    if len(raw) % BLOCK:
        raise FitsError(f"{path}: size {len(raw)} is not a multiple of {BLOCK}")
    pos, hdus = 0, []
    while pos < len(raw):
        hdr = {}
        while True:
            block = raw[pos:pos + BLOCK].decode("ascii")
            pos += BLOCK
            done = False
            for k in range(0, BLOCK, 80):
                card = block[k:k + 80]
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    hdr[key] = _parse_value(card[10:])
            if done:
                break
        if hdr.get("BITPIX") != -64:
            raise FitsError(f"{path}: BITPIX {hdr.get('BITPIX')} (this reader handles -64)")
        shape = tuple(hdr[f"NAXIS{k}"] for k in range(hdr["NAXIS"], 0, -1))
        n = int(np.prod(shape)) if shape else 0
        a = np.frombuffer(raw, dtype=">f8", count=n, offset=pos).astype(np.float64).reshape(shape)
        pos += n * 8 + ((-n * 8) % BLOCK)
        hdus.append((hdr, a))
    return hdus
