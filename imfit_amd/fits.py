"""FITS images for the imfit workflow: read data, noise, mask and PSF images, write
model images.

Covers what imfit asks of CFITSIO (core/image_io.cpp:110-375,
core/utilities.cpp:346-489):
  - CheckForImage / CheckHDUForImage: the chosen HDU must be an image HDU with
    NAXIS = 2 (core/image_io.cpp:110-172);
  - GetImageSize, ReadImageAsVector: NAXIS1 = columns, NAXIS2 = rows, pixels read
    as doubles with BSCALE/BZERO applied and no null-value substitution
    (fits_read_pix(..., TDOUBLE, ..., nulval = NULL), core/image_io.cpp:239-299);
  - SaveVectorAsImage: a single-precision (BITPIX -32) primary image with COMMENT
    cards and a DATE card (core/image_io.cpp:306-360);
  - the extended file names imfit accepts: "name.fits[n]" (HDU n+1, n = 0 is the
    primary), "name.fits[x1:x2,y1:y2]" (1-based inclusive section, "*" for a whole
    axis) and "name.fits[n][x1:x2,y1:y2]", with the image offsets
    DetermineImageOffset derives from a section (core/utilities.cpp:346-489).

Pixels come back as a float64 array of shape (rows, columns): FITS stores the first
axis (x) fastest, so a C-order reshape gives imfit's row-major pixel vector.
"""
import datetime
import os
import re
import sys

import numpy as np

BLOCK = 2880
CARD = 80

_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


class FitsError(IOError):
    pass


# ---------------------------------------------------------------- file names
_NAME_RE = re.compile(r"^(?P<path>[^\[]+)(?P<brackets>(\[[^\]]*\])*)$")


def split_filename(name):
    """(path, hdu_number or None, section string or None) of an extended file name."""
    m = _NAME_RE.match(name.strip())
    if m is None:
        raise FitsError("bad FITS file name %r" % name)
    path = m.group("path")
    parts = re.findall(r"\[([^\]]*)\]", m.group("brackets"))
    hdu, section = None, None
    for p in parts:
        p = p.strip()
        if re.fullmatch(r"\d+", p):
            if hdu is not None or section is not None:
                raise FitsError("bad FITS file name %r" % name)
            hdu = int(p)
        elif "," in p:
            if section is not None:
                raise FitsError("bad FITS file name %r" % name)
            section = p
        else:
            raise FitsError("unsupported FITS file-name specifier [%s] in %r" % (p, name))
    return path, hdu, section


def _axis_range(spec, n, name):
    spec = spec.strip()
    if spec == "*":
        return 1, n
    parts = spec.split(":")
    if len(parts) != 2:
        raise FitsError("incorrect image section format in %r" % name)
    try:
        a, b = int(parts[0]), int(parts[1])
    except ValueError:
        raise FitsError("incorrect image section format in %r" % name)
    if not (1 <= a <= b <= n):
        raise FitsError("image section %s lies outside the image (1..%d) in %r" % (spec, n, name))
    return a, b


def GetPixelStartCoords(name):
    """(x1, y1) of the image section in a file name; (1, 1) without a section
    (core/utilities.cpp:346-430)."""
    _, _, section = split_filename(name)
    if section is None:
        return 1, 1
    xs, ys = section.split(",")
    x1 = 1 if xs.strip() == "*" else int(xs.split(":")[0])
    y1 = 1 if ys.strip() == "*" else int(ys.split(":")[0])
    return x1, y1


def DetermineImageOffset(name):
    """X0/Y0 offsets implied by an image section: (x1 - 1, y1 - 1)
    (core/utilities.cpp:483-489)."""
    x1, y1 = GetPixelStartCoords(name)
    return x1 - 1, y1 - 1


# ---------------------------------------------------------------- headers and HDUs
def _card_value(card):
    """Value text of a "KEYWORD = value / comment" card (quoted strings keep their '/')."""
    v = card[10:].strip()
    if v.startswith("'"):
        out, i = [], 1
        while i < len(v):
            if v[i] == "'":
                if i + 1 < len(v) and v[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(v[i])
            i += 1
        return "".join(out).rstrip()
    return v.split("/")[0].strip()


class HDU:
    def __init__(self, header, data_offset, index):
        self.header = header          # keyword -> value text (first occurrence)
        self.data_offset = data_offset
        self.index = index            # 0 = primary

    def _int(self, key, default=None):
        v = self.header.get(key)
        if v is None:
            if default is None:
                raise FitsError("missing %s keyword in HDU %d" % (key, self.index + 1))
            return default
        return int(v)

    def _float(self, key, default):
        v = self.header.get(key)
        return default if v is None else float(v.replace("D", "E").replace("d", "e"))

    @property
    def is_image(self):
        if self.index == 0:
            return self.header.get("SIMPLE") == "T"
        return self.header.get("XTENSION", "").strip() == "IMAGE"

    @property
    def dims(self):
        return [self._int("NAXIS%d" % (k + 1)) for k in range(self._int("NAXIS"))]

    def data_bytes(self):
        naxis = self._int("NAXIS")
        if naxis == 0:
            return 0
        n = 1
        for d in self.dims:
            n *= d
        bitpix = self._int("BITPIX")
        pcount = self._int("PCOUNT", 0) if self.index > 0 else 0
        gcount = self._int("GCOUNT", 1) if self.index > 0 else 1
        return abs(bitpix) // 8 * gcount * (pcount + n)


def _read_hdus(raw, path):
    hdus, pos, index = [], 0, 0
    while pos + BLOCK <= len(raw):
        header, done = {}, False
        while not done:
            if pos + BLOCK > len(raw):
                raise FitsError("%s: truncated FITS header" % path)
            block = raw[pos:pos + BLOCK]
            pos += BLOCK
            for k in range(BLOCK // CARD):
                card = block[CARD * k:CARD * (k + 1)].decode("ascii", "replace")
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= " and key not in header:
                    header[key] = _card_value(card)
        if index == 0 and header.get("SIMPLE") is None:
            raise FitsError("%s: not a FITS file (no SIMPLE card)" % path)
        hdu = HDU(header, pos, index)
        hdus.append(hdu)
        pos += (hdu.data_bytes() + BLOCK - 1) // BLOCK * BLOCK
        index += 1
    if not hdus:
        raise FitsError("%s: not a FITS file" % path)
    return hdus


def _open(name):
    path, hdu, section = split_filename(name)
    try:
        with open(path, "rb") as f:
            raw = f.read()
    except OSError as e:
        raise FitsError("problems opening FITS file %r: %s" % (name, e))
    hdus = _read_hdus(raw, path)
    k = 0 if hdu is None else hdu
    if k >= len(hdus):
        raise FitsError("%r: HDU %d does not exist" % (name, k + 1))
    return raw, hdus[k], section


def _check_image(hdu, name):
    if not hdu.is_image or hdu._int("NAXIS") != 2:
        raise FitsError("unable to find valid 2D image in extension %d (HDU %d) of %s"
                        % (hdu.index, hdu.index + 1, name))


def read_image(name):
    """Pixels of the image HDU (and section) a file name selects, float64 (rows, cols).
    Raises FitsError where ReadImageAsVector returns nullptr."""
    raw, hdu, section = _open(name)
    _check_image(hdu, name)
    ncols, nrows = hdu.dims
    bitpix = hdu._int("BITPIX")
    if bitpix not in _DTYPES:
        raise FitsError("%r: unsupported BITPIX %d" % (name, bitpix))
    n = ncols * nrows
    if hdu.data_offset + n * abs(bitpix) // 8 > len(raw):
        raise FitsError("%r: truncated pixel data" % name)
    pix = np.frombuffer(raw, dtype=_DTYPES[bitpix], count=n, offset=hdu.data_offset)
    img = pix.astype(np.float64).reshape(nrows, ncols)
    bscale, bzero = hdu._float("BSCALE", 1.0), hdu._float("BZERO", 0.0)
    if bscale != 1.0 or bzero != 0.0:
        img = img * bscale + bzero
    if section is not None:
        xs, ys = section.split(",")
        x1, x2 = _axis_range(xs, ncols, name)
        y1, y2 = _axis_range(ys, nrows, name)
        img = img[y1 - 1:y2, x1 - 1:x2]
    return np.ascontiguousarray(img)


# ---------------------------------------------------------------- reference-named entry points
def CheckForImage(name):
    """1 if the file name selects a 2-D image HDU, else -1 (core/image_io.cpp:143-172)."""
    try:
        _, hdu, _ = _open(name)
        _check_image(hdu, name)
    except FitsError as e:
        print("*** WARNING: %s" % e, file=sys.stderr)
        return -1
    return 1


def GetImageSize(name):
    """(nColumns, nRows, status): status 0, or -1 with sizes 0 (core/image_io.cpp:180-228).
    A section in the name gives the section's size, as CFITSIO reports it."""
    try:
        img = read_image(name)
    except FitsError as e:
        print("*** WARNING: %s" % e, file=sys.stderr)
        return 0, 0, -1
    return img.shape[1], img.shape[0], 0


def ReadImageAsVector(name):
    """(pixels as a flat float64 vector, nColumns, nRows), or (None, 0, 0) on failure
    (core/image_io.cpp:239-299)."""
    try:
        img = read_image(name)
    except FitsError as e:
        print("*** WARNING: %s" % e, file=sys.stderr)
        return None, 0, 0
    return img.reshape(-1), img.shape[1], img.shape[0]


def _card(key, value=None, comment=None):
    if value is None:
        text = "%-8s%s" % (key, "" if comment is None else comment)
    else:
        text = "%-8s= %20s" % (key, value)
        if comment:
            text += " / " + comment
    return text[:CARD].ljust(CARD)


def write_image(pixels, name, comments=(), bitpix=-32):
    """Write a primary-HDU image, replacing any existing file.  BITPIX -32 is what
    SaveVectorAsImage writes (single precision); -64 keeps doubles exact."""
    img = np.asarray(pixels, dtype=np.float64)
    if img.ndim != 2:
        raise FitsError("write_image needs a 2-D array")
    if bitpix not in (-32, -64):
        raise FitsError("write_image supports BITPIX -32 and -64")
    nrows, ncols = img.shape
    cards = [_card("SIMPLE", "T", "file does conform to FITS standard"),
             _card("BITPIX", str(bitpix), "number of bits per data pixel"),
             _card("NAXIS", "2", "number of data axes"),
             _card("NAXIS1", str(ncols), "length of data axis 1"),
             _card("NAXIS2", str(nrows), "length of data axis 2"),
             _card("EXTEND", "T", "FITS dataset may contain extensions")]
    for c in comments:
        c = str(c)
        for k in range(0, max(len(c), 1), 72):
            cards.append(_card("COMMENT", comment=c[k:k + 72]))
    now = datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%S")
    cards.append(_card("DATE", "'%s'" % now, "file creation date (YYYY-MM-DDThh:mm:ss UT)"))
    cards.append("END".ljust(CARD))
    header = "".join(cards).encode("ascii")
    header += b" " * ((-len(header)) % BLOCK)
    data = img.astype(">f4" if bitpix == -32 else ">f8").tobytes()
    data += b"\0" * ((-len(data)) % BLOCK)
    with open(name, "wb") as f:
        f.write(header)
        f.write(data)


def SaveVectorAsImage(pixels, name, nColumns, nRows, comments=()):
    """0 on success, -1 when the file cannot be written, -2 for a missing pixel vector
    (core/image_io.cpp:306-360)."""
    if pixels is None:
        print("*** WARNING: input image array to SaveVectorAsImage is None!", file=sys.stderr)
        return -2
    try:
        write_image(np.asarray(pixels, dtype=np.float64).reshape(nRows, nColumns), name, comments)
    except (OSError, ValueError) as e:
        print("*** WARNING: problems writing FITS file %r: %s" % (name, e), file=sys.stderr)
        return -1
    return 0


def image_exists(name):
    """ImageFileExists: the on-disk file of an extended name exists (core/utilities.cpp:500-520)."""
    return os.path.exists(split_filename(name)[0])
