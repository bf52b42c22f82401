"""Minimal FITS headers, image HDUs and binary tables (numpy only).

The packers read exposure metadata as FITS headers: the metadata store keeps
them as header strings (forcepho/patches/storage.py:302-522, written with
``Header.tostring`` and read back with ``Header.fromstring``) and the FITS
packer reads image extensions (forcepho/patches/pixel_patch.py:316-483).
This package does not depend on astropy, so this module implements the subset of
the FITS standard (version 4.0) those paths use:

* 80-character header cards: ``KEYWORD = value / comment`` with string
  (quoted, ``''`` escapes a quote), logical (T/F), integer and real values;
  COMMENT / HISTORY / blank commentary cards; the END card;
* files of 2880-byte blocks: a primary HDU and IMAGE extensions, selected
  by index (negative counts from the end) or EXTNAME; big-endian data of
  BITPIX 8, 16, 32, 64, -32, -64, with BSCALE / BZERO applied on read;
* BINTABLE extensions with fixed-width columns (TFORM codes L, B, I, J, K,
  E, D, A, with repeat counts), read as numpy structured arrays (source
  catalogs are stored this way).

Long-string CONTINUE cards, variable-length and complex table columns,
ASCII tables and compressed HDUs are not supported.
"""

import numpy as np

__all__ = ["Header", "getheader", "getdata", "writeto", "BLOCK"]

BLOCK = 2880
CARD = 80
_COMMENTARY = ("COMMENT", "HISTORY", "")
_BITPIX = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}
_TFORM = {"L": "i1", "B": ">u1", "I": ">i2", "J": ">i4", "K": ">i8", "E": ">f4",
          "D": ">f8"}


def _parse_value(text):
    """(value, comment) of the value field of a card (columns 11-80)."""
    s = text.strip()
    if s.startswith("'"):
        out, i = [], 1
        while i < len(s):
            c = s[i]
            if c == "'":
                if i + 1 < len(s) and s[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(c)
            i += 1
        rest = s[i + 1:]
        comment = rest.split("/", 1)[1].strip() if "/" in rest else ""
        return "".join(out).rstrip(), comment
    if "/" in s:
        s, comment = s.split("/", 1)
        s, comment = s.strip(), comment.strip()
    else:
        comment = ""
    if s == "T":
        return True, comment
    if s == "F":
        return False, comment
    if s == "":
        return None, comment
    try:
        return int(s), comment
    except ValueError:
        pass
    try:
        return float(s.replace("D", "E").replace("d", "e")), comment
    except ValueError:
        return s, comment


def _format_value(v):
    if isinstance(v, (bool, np.bool_)):
        return f"{'T' if v else 'F':>20}"
    if isinstance(v, (int, np.integer)):
        return f"{int(v):>20d}"
    if isinstance(v, (float, np.floating)):
        s = repr(float(v)).upper()
        if "E" not in s and "." not in s and "INF" not in s and "NAN" not in s:
            s += ".0"
        return f"{s:>20}"
    if v is None:
        return " " * 20
    q = str(v).replace("'", "''")
    return f"'{q:<8}'"


class Header:
    """An ordered list of FITS cards with dict-style access by keyword
    (the first card of a keyword; commentary cards only through ``cards``)."""

    def __init__(self, cards=None):
        self.cards = []   # [(keyword, value, comment)]
        if isinstance(cards, dict):
            cards = cards.items()
        for c in cards or []:
            if len(c) == 2:
                self[c[0]] = c[1]
            else:
                self.append(*c)

    # -- dict-like access ---------------------------------------------------
    def _index(self, key):
        key = key.upper()
        for i, (k, _, _) in enumerate(self.cards):
            if k == key and k not in _COMMENTARY:
                return i
        return -1

    def __getitem__(self, key):
        i = self._index(key)
        if i < 0:
            raise KeyError(key)
        return self.cards[i][1]

    def __setitem__(self, key, value):
        comment = ""
        if isinstance(value, tuple):
            value, comment = value
        key = key.upper()
        i = self._index(key)
        if i < 0:
            self.cards.append((key, value, comment))
        else:
            self.cards[i] = (key, value, comment or self.cards[i][2])

    def __contains__(self, key):
        return self._index(key) >= 0

    def __delitem__(self, key):
        i = self._index(key)
        if i < 0:
            raise KeyError(key)
        del self.cards[i]

    def get(self, key, default=None):
        i = self._index(key)
        return default if i < 0 else self.cards[i][1]

    def keys(self):
        return [k for k, _, _ in self.cards if k not in _COMMENTARY]

    def items(self):
        return [(k, v) for k, v, _ in self.cards if k not in _COMMENTARY]

    def __len__(self):
        return len(self.cards)

    def copy(self):
        h = Header()
        h.cards = list(self.cards)
        return h

    def append(self, key, value=None, comment=""):
        self.cards.append((key.upper(), value, comment))

    def add_comment(self, text):
        self.append("COMMENT", str(text))

    def add_history(self, text):
        self.append("HISTORY", str(text))

    # -- serialisation ------------------------------------------------------
    @staticmethod
    def _card(key, value, comment):
        if key in _COMMENTARY:
            s = f"{key:<8}{'' if value is None else str(value)}"
        elif key == "END":
            s = "END"
        else:
            s = f"{key:<8}= {_format_value(value)}"
            if comment:
                s += f" / {comment}"
        if len(s) > CARD:
            raise ValueError(f"card too long for FITS: {s!r}")
        return f"{s:<{CARD}}"

    def tostring(self, padding=True, endcard=True):
        """The header as concatenated 80-character cards (padded to a
        multiple of 2880 bytes with blanks when ``padding``)."""
        s = "".join(self._card(*c) for c in self.cards)
        if endcard:
            s += self._card("END", None, "")
        if padding and len(s) % BLOCK:
            s += " " * (BLOCK - len(s) % BLOCK)
        return s

    @classmethod
    def fromstring(cls, s):
        """Parse concatenated 80-character cards up to the END card."""
        h = cls()
        for i in range(0, len(s), CARD):
            card = s[i:i + CARD]
            key = card[:8].strip().upper()
            if key == "END":
                break
            if key in _COMMENTARY and card[8:10] != "= ":
                text = card[8:].rstrip()
                if key == "" and text == "":
                    continue
                h.cards.append((key, text, ""))
                continue
            if card[8:10] != "= ":
                h.cards.append((key, None, card[8:].strip()))
                continue
            value, comment = _parse_value(card[10:])
            h.cards.append((key, value, comment))
        return h

    def __repr__(self):
        return "\n".join(self._card(*c).rstrip() for c in self.cards)


# ------------------------------------------------------------------ files
def _read_hdus(path):
    """[(header, data_offset, data_nbytes)] of every HDU of a FITS file."""
    out = []
    with open(path, "rb") as f:
        blob = f.read()
    pos = 0
    while pos + BLOCK <= len(blob):
        # header: blocks until the END card
        end = -1
        start = pos
        while pos + BLOCK <= len(blob):
            block = blob[pos:pos + BLOCK].decode("ascii")
            pos += BLOCK
            for i in range(0, BLOCK, CARD):
                if block[i:i + 8] == "END     ":
                    end = pos
                    break
            if end >= 0:
                break
        if end < 0:
            break
        hdr = Header.fromstring(blob[start:end].decode("ascii"))
        naxis = int(hdr.get("NAXIS", 0))
        n = 0
        if naxis > 0:
            n = 1
            for k in range(1, naxis + 1):
                n *= int(hdr[f"NAXIS{k}"])
            n = (abs(int(hdr["BITPIX"])) // 8) * int(hdr.get("GCOUNT", 1)) * \
                (int(hdr.get("PCOUNT", 0)) + n)
        out.append((hdr, pos, n))
        pos += -(-n // BLOCK) * BLOCK
    return out, blob


def _select(hdus, ext):
    if ext is None:
        ext = 0
    if isinstance(ext, (int, np.integer)):
        return hdus[int(ext)]
    for h in hdus:
        if str(h[0].get("EXTNAME", "")).strip().upper() == str(ext).upper():
            return h
    raise KeyError(f"no HDU with EXTNAME {ext!r}")


def getheader(path, ext=0):
    """Header of HDU ``ext`` (index or EXTNAME)."""
    hdus, _ = _read_hdus(path)
    return _select(hdus, ext)[0]


def _table(hdr, blob, off):
    """Structured array of a BINTABLE HDU's rows."""
    fields = []
    for k in range(1, int(hdr["TFIELDS"]) + 1):
        form = str(hdr[f"TFORM{k}"]).strip().upper()
        i = 0
        while i < len(form) and form[i].isdigit():
            i += 1
        rep = int(form[:i]) if i else 1
        code = form[i:i + 1]
        name = str(hdr.get(f"TTYPE{k}", f"col{k}")).strip()
        if code == "A":
            fields.append((name, f"S{rep}"))
        elif code in _TFORM:
            dt = np.dtype(_TFORM[code])
            fields.append((name, dt) if rep == 1 else (name, dt, (rep,)))
        else:
            raise NotImplementedError(f"table column format {form!r} is not supported")
    dt = np.dtype(fields)
    if dt.itemsize != int(hdr["NAXIS1"]):
        raise ValueError("BINTABLE row size does not match its columns")
    rows = np.frombuffer(blob, dtype=dt, count=int(hdr["NAXIS2"]), offset=off)
    native = []
    for n in dt.names:
        f = dt.fields[n][0]
        native.append((n, f.newbyteorder("=") if f.base.kind in "iuf" else f))
    out = np.empty(rows.shape, dtype=np.dtype(native))
    for n in dt.names:
        out[n] = rows[n]
    return out


def getdata(path, ext=0):
    """Data of HDU ``ext`` (index or EXTNAME): an image as a native-endian
    array of shape (NAXISn, ..., NAXIS1) with BSCALE / BZERO applied (float
    result), a binary table as a structured array of its rows."""
    hdus, blob = _read_hdus(path)
    hdr, off, nbytes = _select(hdus, ext)
    naxis = int(hdr.get("NAXIS", 0))
    if naxis == 0:
        return None
    if str(hdr.get("XTENSION", "")).strip() == "BINTABLE":
        return _table(hdr, blob, off)
    shape = tuple(int(hdr[f"NAXIS{k}"]) for k in range(naxis, 0, -1))
    dt = np.dtype(_BITPIX[int(hdr["BITPIX"])])
    a = np.frombuffer(blob, dtype=dt, count=int(np.prod(shape)), offset=off)
    a = a.reshape(shape).astype(dt.newbyteorder("="))
    bscale, bzero = hdr.get("BSCALE", 1.0), hdr.get("BZERO", 0.0)
    if bscale != 1.0 or bzero != 0.0:
        a = a * np.float64(bscale) + np.float64(bzero)
    return a


def writeto(path, hdus, overwrite=False):
    """Write [(header or dict or None, ndarray or None), ...]: the first is
    the primary HDU, the rest IMAGE extensions.  Structural keywords
    (SIMPLE/XTENSION, BITPIX, NAXISn, EXTEND, PCOUNT, GCOUNT) are set from
    the data; the given cards follow them."""
    import os
    if os.path.exists(path) and not overwrite:
        raise OSError(f"{path} exists")
    inv = {np.dtype(v).newbyteorder("="): k for k, v in _BITPIX.items()}
    with open(path, "wb") as f:
        for i, (h, data) in enumerate(hdus):
            h = Header(h) if not isinstance(h, Header) else h
            out = Header()
            if i == 0:
                out["SIMPLE"] = True
            else:
                out["XTENSION"] = "IMAGE"
            arr = None if data is None else np.ascontiguousarray(data)
            if arr is not None and arr.dtype not in inv:
                arr = arr.astype(np.float64)
            out["BITPIX"] = 8 if arr is None else inv[arr.dtype]
            out["NAXIS"] = 0 if arr is None else arr.ndim
            if arr is not None:
                for k in range(arr.ndim):
                    out[f"NAXIS{k + 1}"] = int(arr.shape[arr.ndim - 1 - k])
            if i == 0:
                out["EXTEND"] = True
            else:
                out["PCOUNT"] = 0
                out["GCOUNT"] = 1
            skip = {"SIMPLE", "XTENSION", "BITPIX", "NAXIS", "EXTEND", "PCOUNT", "GCOUNT"}
            for key, value, comment in h.cards:
                if key in skip or (key.startswith("NAXIS") and key[5:].isdigit()):
                    continue
                out.cards.append((key, value, comment))
            f.write(out.tostring().encode("ascii"))
            if arr is not None:
                raw = arr.astype(arr.dtype.newbyteorder(">")).tobytes()
                f.write(raw)
                if len(raw) % BLOCK:
                    f.write(b"\0" * (BLOCK - len(raw) % BLOCK))
