"""Read-only access to FITS files without astropy.

forcepho's pixel patches read exposures, their uncertainty images and source
catalogs from FITS files (``patches/pixel_patch.py:387-484``,
``tests/verification/verify_reference.py:64-69``).  This module reads what
those uses need, following the FITS Standard (version 4.0):

* header cards of every HDU, with long string values continued on
  ``CONTINUE`` cards;
* primary and ``IMAGE`` extension arrays of any BITPIX, with BSCALE/BZERO,
  returned as numpy arrays in C order (NAXISn reversed, as astropy does);
* ``BINTABLE`` extensions with fixed-width columns (TFORM codes L, X, B, I,
  J, K, A, E, D, C, M), returned as numpy structured arrays (TZERO/TSCAL
  applied to numeric columns).

Interface: ``open(path)`` -> list of HDUs (``.header`` mapping, ``.data``),
``getheader(path, ext)``, ``getdata(path, ext)``; negative ``ext`` counts
from the end, as in astropy.
"""

import builtins

import numpy as np

__all__ = ["Header", "HDU", "open", "getheader", "getdata", "write_table", "write_images",
           "FITSFormatError"]

BLOCK = 2880
CARD = 80

_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
_TFORM = {"L": ("u1", 1), "B": (">u1", 1), "I": (">i2", 2), "J": (">i4", 4),
          "K": (">i8", 8), "E": (">f4", 4), "D": (">f8", 8), "C": (">c8", 8),
          "M": (">c16", 16)}


class FITSFormatError(ValueError):
    """The file is not FITS, or uses a part of the standard not handled."""


class Header(dict):
    """Keyword -> value (str, bool, int or float); ``cards`` keeps the raw
    80-character cards in order and ``comments`` the value comments."""

    def __init__(self):
        super().__init__()
        self.cards = []
        self.comments = {}

    def get(self, key, default=None):
        return super().get(key.upper(), default)

    def __getitem__(self, key):
        return super().__getitem__(key.upper())

    def __contains__(self, key):
        return super().__contains__(key.upper())


def _parse_value(text):
    """Value field of a card (columns 11-80) -> (value, comment)."""
    s = text.lstrip()
    if s.startswith("'"):
        out, i = [], 1
        while i < len(s):
            c = s[i]
            if c == "'":
                if i + 1 < len(s) and s[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                i += 1
                break
            out.append(c)
            i += 1
        rest = s[i:]
        comment = rest.split("/", 1)[1].strip() if "/" in rest else ""
        return "".join(out).rstrip(), comment
    body, _, comment = s.partition("/")
    body = body.strip()
    comment = comment.strip()
    if body == "T":
        return True, comment
    if body == "F":
        return False, comment
    if body == "":
        return None, comment
    try:
        return int(body), comment
    except ValueError:
        pass
    try:
        return float(body.replace("D", "E").replace("d", "e")), comment
    except ValueError:
        return body, comment


def _read_header(buf, pos):
    """Header starting at byte pos -> (Header, position of the data)."""
    hdr = Header()
    last_string_key = None
    while True:
        if pos + BLOCK > len(buf):
            raise FITSFormatError("header runs past the end of the file")
        block = buf[pos:pos + BLOCK]
        pos += BLOCK
        for k in range(BLOCK // CARD):
            card = block[k * CARD:(k + 1) * CARD].decode("ascii", "replace")
            key = card[:8].strip()
            hdr.cards.append(card)
            if key == "END":
                return hdr, pos
            if key == "CONTINUE" and last_string_key is not None:
                more, comment = _parse_value(card[8:])
                prev = dict.__getitem__(hdr, last_string_key)
                if isinstance(more, str) and prev.endswith("&"):
                    dict.__setitem__(hdr, last_string_key, prev[:-1] + more)
                continue
            if card[8:10] != "= ":
                continue
            value, comment = _parse_value(card[10:])
            dict.__setitem__(hdr, key, value)
            hdr.comments[key] = comment
            last_string_key = key if isinstance(value, str) else None


def _data_size(hdr):
    naxis = int(hdr.get("NAXIS", 0))
    if naxis == 0:
        return 0
    n = 1
    for i in range(1, naxis + 1):
        n *= int(hdr["NAXIS%d" % i])
    bits = abs(int(hdr["BITPIX"])) * int(hdr.get("GCOUNT", 1)) * (int(hdr.get("PCOUNT", 0)) + n)
    return bits // 8


class HDU:
    """One header-data unit: ``header`` (Header) and ``data`` (numpy array,
    structured array for a binary table, or None)."""

    def __init__(self, header, raw):
        self.header = header
        self._raw = raw
        self._data = None

    @property
    def kind(self):
        if "SIMPLE" in self.header:
            return "PRIMARY"
        return str(self.header.get("XTENSION", "")).strip()

    @property
    def name(self):
        return str(self.header.get("EXTNAME", "PRIMARY" if self.kind == "PRIMARY" else "")).strip()

    @property
    def data(self):
        if self._data is None:
            self._data = self._decode()
        return self._data

    def _decode(self):
        hdr = self.header
        naxis = int(hdr.get("NAXIS", 0))
        if naxis == 0:
            return None
        if self.kind == "BINTABLE":
            return self._decode_table()
        if self.kind not in ("PRIMARY", "IMAGE"):
            raise FITSFormatError("HDU kind %r is not supported" % self.kind)
        bitpix = int(hdr["BITPIX"])
        shape = tuple(int(hdr["NAXIS%d" % i]) for i in range(naxis, 0, -1))
        n = int(np.prod(shape))
        a = np.frombuffer(self._raw, dtype=_BITPIX[bitpix], count=n).reshape(shape)
        bscale = hdr.get("BSCALE", 1)
        bzero = hdr.get("BZERO", 0)
        if bscale != 1 or bzero != 0:
            out_t = np.float32 if bitpix in (8, 16, -32) else np.float64
            return a.astype(out_t) * out_t(bscale) + out_t(bzero)
        return a.astype(a.dtype.newbyteorder("="))

    def _decode_table(self):
        hdr = self.header
        row_bytes = int(hdr["NAXIS1"])
        n_rows = int(hdr["NAXIS2"])
        names, formats, offsets = [], [], []
        scaled, logical = {}, set()
        off = 0
        for i in range(1, int(hdr["TFIELDS"]) + 1):
            form = str(hdr["TFORM%d" % i]).strip()
            name = str(hdr.get("TTYPE%d" % i, "col%d" % i)).strip()
            j = 0
            while j < len(form) and form[j].isdigit():
                j += 1
            repeat = int(form[:j]) if j else 1
            code = form[j:j + 1]
            if code == "A":
                dt, width = np.dtype("S%d" % max(repeat, 1)), repeat
            elif code == "X":
                dt, width = np.dtype((np.uint8, ((repeat + 7) // 8,))), (repeat + 7) // 8
            elif code in _TFORM:
                base, size = _TFORM[code]
                dt = np.dtype(base) if repeat == 1 else np.dtype((base, (repeat,)))
                width = size * repeat
            else:
                raise FITSFormatError("TFORM %r is not supported" % form)
            if width:
                names.append(name)
                formats.append(dt)
                offsets.append(off)
                if code == "L":
                    logical.add(name)
                elif ("TZERO%d" % i) in hdr or ("TSCAL%d" % i) in hdr:
                    scaled[name] = (hdr.get("TSCAL%d" % i, 1), hdr.get("TZERO%d" % i, 0))
            off += width
        if off != row_bytes:
            raise FITSFormatError("column widths (%d) do not match NAXIS1 (%d)" % (off, row_bytes))
        raw_dt = np.dtype({"names": names, "formats": formats, "offsets": offsets,
                           "itemsize": row_bytes})
        raw = np.frombuffer(self._raw, dtype=raw_dt, count=n_rows)
        out_fields = []
        for name, f in zip(names, formats):
            base = f.base
            if name in logical:
                out_fields.append((name, np.dtype(bool), f.shape))
            elif base.kind in "iufc":
                nb = base.newbyteorder("=")
                if name in scaled:
                    nb = np.dtype(np.float64)
                out_fields.append((name, nb, f.shape))
            else:
                out_fields.append((name, base, f.shape))
        out = np.zeros(n_rows, dtype=[(n, b, s) if s else (n, b) for n, b, s in out_fields])
        for name in names:
            col = raw[name]
            if name in scaled:
                sc, z = scaled[name]
                col = col.astype(np.float64) * sc + z
            elif name in logical:
                col = col == ord("T")       # logicals are the bytes 'T' / 'F'
            out[name] = col
        return out


def open(path):
    """Every HDU of a FITS file, in order."""
    with builtins.open(path, "rb") as fh:
        buf = fh.read()
    if not buf.startswith(b"SIMPLE  ="):
        raise FITSFormatError("%s is not a FITS file" % path)
    hdus, pos = [], 0
    while pos + BLOCK <= len(buf):
        hdr, data_pos = _read_header(buf, pos)
        size = _data_size(hdr)
        if data_pos + size > len(buf):
            raise FITSFormatError("data of HDU %d runs past the end of the file" % len(hdus))
        hdus.append(HDU(hdr, buf[data_pos:data_pos + size]))
        pos = data_pos + (size + BLOCK - 1) // BLOCK * BLOCK
        if pos < len(buf) and not buf[pos:pos + 8].startswith(b"XTENSION"):
            break
    return hdus


def getheader(path, ext=0):
    return open(path)[ext].header


def getdata(path, ext=0):
    return open(path)[ext].data


# ---------------------------------------------------------------------------
# Writing (binary tables and images), for catalog state files

_WRITE_CODES = {"?": "L", "u1": "B", "i2": "I", "i4": "J", "i8": "K", "f4": "E", "f8": "D",
                "c8": "C", "c16": "M"}


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, (int, np.integer)):
        v = "%20d" % int(value)
    elif isinstance(value, (float, np.floating)):
        v = "%20s" % repr(float(value)).upper()
    else:
        v = "'%-8s'" % str(value).replace("'", "''")
    text = "%-8s= %s" % (key[:8], v)
    if comment:
        text += " / " + comment
    if len(text) > CARD:
        raise FITSFormatError("card for %s does not fit in 80 characters" % key)
    return text.ljust(CARD)


def _header_bytes(cards):
    body = "".join(cards) + "END".ljust(CARD)
    body += " " * ((-len(body)) % BLOCK)
    return body.encode("ascii")


def _pad(data):
    return data + b"\x00" * ((-len(data)) % BLOCK)


def table_hdu_bytes(table, header=None):
    """A BINTABLE extension (header + data bytes) of a structured array."""
    table = np.asarray(table)
    cols, forms = [], []
    for name in table.dtype.names:
        f = table.dtype[name]
        base, shape = f.base, f.shape
        n = int(np.prod(shape)) if shape else 1
        if base.kind == "S":
            forms.append("%dA" % base.itemsize)
            cols.append((name, base, shape))
            continue
        key = base.str[1:] if base.kind != "b" else "?"
        if key not in _WRITE_CODES:
            raise FITSFormatError("column %s of dtype %s cannot be written" % (name, base))
        forms.append(("%d" % n if n != 1 else "") + _WRITE_CODES[key])
        cols.append((name, base.newbyteorder(">") if base.kind != "b" else base, shape))
    big = np.zeros(len(table), dtype=[(n, b if b.kind != "b" else np.dtype("u1"), s) if s
                                      else (n, b if b.kind != "b" else np.dtype("u1"))
                                      for n, b, s in cols])
    for name in table.dtype.names:
        if table.dtype[name].base.kind == "b":
            big[name] = np.where(table[name], ord("T"), ord("F"))   # logicals as 'T' / 'F'
        else:
            big[name] = table[name]
    cards = [_card("XTENSION", "BINTABLE", "binary table extension"),
             _card("BITPIX", 8), _card("NAXIS", 2), _card("NAXIS1", big.dtype.itemsize),
             _card("NAXIS2", len(big)), _card("PCOUNT", 0), _card("GCOUNT", 1),
             _card("TFIELDS", len(cols))]
    for i, (name, form) in enumerate(zip(table.dtype.names, forms), start=1):
        cards += [_card("TTYPE%d" % i, name), _card("TFORM%d" % i, form)]
    for k, v in (header or {}).items():
        cards.append(_card(k, v))
    return _header_bytes(cards) + _pad(big.tobytes())


def write_table(path, table, header=None):
    """Write a structured array as a FITS file: an empty primary HDU and
    one BINTABLE extension (extra header cards from ``header``)."""
    primary = _header_bytes([_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0),
                             _card("EXTEND", True)])
    with builtins.open(path, "wb") as fh:
        fh.write(primary + table_hdu_bytes(table, header))


def image_hdu_bytes(data, header=None, primary=False):
    """An image HDU (header + data bytes) of a 2-d float or integer array."""
    a = np.asarray(data)
    codes = {np.dtype(np.float64): -64, np.dtype(np.float32): -32, np.dtype(np.int32): 32,
             np.dtype(np.int16): 16, np.dtype(np.uint8): 8, np.dtype(np.int64): 64}
    bitpix = codes.get(a.dtype.newbyteorder("="))
    if bitpix is None:
        raise FITSFormatError("image dtype %s cannot be written" % a.dtype)
    cards = ([_card("SIMPLE", True)] if primary else [_card("XTENSION", "IMAGE")])
    cards += [_card("BITPIX", bitpix), _card("NAXIS", a.ndim)]
    cards += [_card("NAXIS%d" % (i + 1), n) for i, n in enumerate(a.shape[::-1])]
    cards += ([_card("EXTEND", True)] if primary else [_card("PCOUNT", 0), _card("GCOUNT", 1)])
    for k, v in (header or {}).items():
        cards.append(_card(k, v))
    return _header_bytes(cards) + _pad(a.astype(a.dtype.newbyteorder(">")).tobytes())


def write_images(path, images, headers=None, table=None, table_header=None):
    """Write 2-d arrays as the primary HDU and IMAGE extensions (numpy C
    order: the last axis is NAXIS1), optionally followed by a BINTABLE."""
    headers = headers or [None] * len(images)
    with builtins.open(path, "wb") as fh:
        for i, (im, hdr) in enumerate(zip(images, headers)):
            fh.write(image_hdu_bytes(im, hdr, primary=(i == 0)))
        if table is not None:
            fh.write(table_hdu_bytes(table, table_header))
