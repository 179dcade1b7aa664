"""FITS images and their linear WCS, read and written by this package.

The reference reads and writes FITS through astropy
(src/gbkfit/utils/fitsutils.py:7-30: ``load_fits`` returns the HDU's data
and an ``astropy.wcs.WCS`` of its header; ``dump_fits`` writes a primary
HDU with the WCS's header cards and CHECKSUM/DATASUM), and the prep task
uses ``fits.getdata`` / ``fits.getheader`` / ``fits.writeto``
(src/gbkfit/tasks/prep.py:8-51).  astropy is not part of this stack, so
this module implements the part of the FITS standard (version 4.0) those
calls touch:

- 2880-byte blocks of 80-character header cards (fixed-format values:
  logical, integer, real with E or D exponent, quoted strings with ''
  escapes; COMMENT / HISTORY / blank commentary cards kept verbatim);
- image HDUs: the primary array and IMAGE extensions, BITPIX 8, 16, 32,
  64, -32, -64, big-endian, NAXIS1 varying fastest (so a numpy array has
  shape (NAXISn, ..., NAXIS1));
- BSCALE / BZERO / BLANK as astropy applies them: the unsigned-integer
  conventions (BZERO 2^15, 2^31, 2^63 with BSCALE 1, and signed bytes with
  BZERO -128) give uint16 / uint32 / uint64 / int8; any other scaling gives
  float32 (BITPIX 8, 16, -32) or float64 (BITPIX 32, 64, -64), with BLANK
  pixels of a scaled integer array set to NaN;
- the FITS checksum convention (standard appendix J): DATASUM is the
  32-bit ones'-complement sum of the data blocks, CHECKSUM the 16-character
  ASCII encoding of the complement of the HDU's sum, so that the whole
  HDU sums to negative zero.

Arrays come back in native byte order, as writable copies (``memmap``
maps the file instead of reading it first; the result is still a copy).

``WCS`` holds what ``astropy.wcs.WCS(header).wcs`` gives the reference's
Data.load: cdelt, crpix, crval (FITS axis order, CRPIX as written, 1 for
a missing CDELT, 0 for a missing CRPIX / CRVAL) and the PC matrix.
"""
import os

import numpy as np

__all__ = ['Header', 'WCS', 'getdata', 'getheader', 'writeto',
           'load_fits', 'dump_fits', 'checksum_encode', 'checksum_decode',
           'ones_complement_sum', 'hdu_checksums']

BLOCK = 2880
CARD = 80

_BITPIX = {8: np.dtype('u1'), 16: np.dtype('>i2'), 32: np.dtype('>i4'),
           64: np.dtype('>i8'), -32: np.dtype('>f4'), -64: np.dtype('>f8')}
_COMMENTARY = ('COMMENT', 'HISTORY', '')


class Header:
    """An ordered list of header cards: (keyword, value, comment).

    Keyword lookups are case-insensitive and return the first card's value;
    commentary cards (COMMENT, HISTORY, blank) may repeat.
    """

    def __init__(self, cards=()):
        self._cards = []
        for c in cards:
            self.append(*c)

    def append(self, key, value=None, comment=''):
        self._cards.append((str(key).upper(), value, comment or ''))

    def _index(self, key):
        key = key.upper()
        for i, c in enumerate(self._cards):
            if c[0] == key:
                return i
        return -1

    def __contains__(self, key):
        return self._index(key) >= 0

    def __getitem__(self, key):
        i = self._index(key)
        if i < 0:
            raise KeyError(key)
        return self._cards[i][1]

    def __setitem__(self, key, value):
        comment = ''
        if isinstance(value, tuple):
            value, comment = value
        i = self._index(key)
        if i < 0 or key.upper() in _COMMENTARY:
            self.append(key, value, comment)
        else:
            self._cards[i] = (key.upper(), value,
                              comment or self._cards[i][2])

    def __delitem__(self, key):
        i = self._index(key)
        if i < 0:
            raise KeyError(key)
        del self._cards[i]

    def get(self, key, default=None):
        i = self._index(key)
        return self._cards[i][1] if i >= 0 else default

    def pop(self, key, default=None):
        i = self._index(key)
        if i < 0:
            return default
        return self._cards.pop(i)[1]

    def keys(self):
        return [c[0] for c in self._cards]

    def cards(self):
        return list(self._cards)

    def copy(self):
        return Header(self._cards)

    def __len__(self):
        return len(self._cards)

    def __repr__(self):
        return '\n'.join(format_card(*c).rstrip() for c in self._cards)


# ---------------------------------------------------------------- cards

def _parse_value(text):
    """Value field of a card (columns 11-80) -> (value, comment)."""
    s = text.lstrip()
    if s.startswith("'"):
        # quoted string; '' is an escaped quote
        out, i = [], 1
        while i < len(s):
            ch = s[i]
            if ch == "'":
                if i + 1 < len(s) and s[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                i += 1
                break
            out.append(ch)
            i += 1
        rest = s[i:]
        comment = rest.split('/', 1)[1].strip() if '/' in rest else ''
        return ''.join(out).rstrip(), comment
    if '/' in s:
        s, comment = s.split('/', 1)
        comment = comment.strip()
    else:
        comment = ''
    s = s.strip()
    if s == '':
        return None, comment
    if s == 'T':
        return True, comment
    if s == 'F':
        return False, comment
    try:
        return int(s), comment
    except ValueError:
        pass
    try:
        return float(s.replace('D', 'E').replace('d', 'e')), comment
    except ValueError:
        raise ValueError(f"unparsable FITS card value: {text!r}")


def parse_card(card):
    """One 80-character card -> (keyword, value, comment)."""
    key = card[:8].rstrip()
    if key in _COMMENTARY or card[8:10] != '= ':
        return key, card[8:].rstrip(), ''
    value, comment = _parse_value(card[10:])
    return key, value, comment


def _format_real(v):
    if v == int(v) and abs(v) < 1e15:
        s = f'{v:.1f}'
    else:
        s = repr(float(v)).upper()
        if 'E' in s and '.' not in s.split('E')[0]:
            m, e = s.split('E')
            s = f'{m}.0E{e}'
    return s


def format_card(key, value=None, comment=''):
    """(keyword, value, comment) -> one 80-character fixed-format card."""
    key = str(key).upper()
    if len(key) > 8:
        raise ValueError(f"FITS keyword longer than 8 characters: {key}")
    if key in _COMMENTARY:
        return f'{key:<8}{str(value or "")}'[:CARD].ljust(CARD)
    if value is None:
        field = ' ' * 20
    elif isinstance(value, (bool, np.bool_)):
        field = f'{"T" if value else "F":>20}'
    elif isinstance(value, (int, np.integer)):
        field = f'{int(value):>20d}'
    elif isinstance(value, (float, np.floating)):
        if not np.isfinite(value):
            raise ValueError(f"FITS header value is not finite: {key}")
        field = f'{_format_real(float(value)):>20}'
    else:
        text = str(value).replace("'", "''")
        field = f"'{text:<8}'"
        if len(field) < 20:
            field = field.ljust(20)
    card = f'{key:<8}= {field}'
    if comment:
        card += f' / {comment}'
    if len(card) > CARD:
        if len(f'{key:<8}= {field}') > CARD:
            raise ValueError(f"FITS card too long: {key}")
        card = card[:CARD]
    return card.ljust(CARD)


def _header_bytes(header):
    text = ''.join(format_card(*c) for c in header.cards())
    text += 'END'.ljust(CARD)
    pad = (-len(text)) % BLOCK
    return (text + ' ' * pad).encode('ascii')


# ---------------------------------------------------------------- HDUs

def _read_header(f):
    """Header at the file's position -> (Header, bytes consumed)."""
    header, nbytes = Header(), 0
    while True:
        block = f.read(BLOCK)
        if len(block) < BLOCK:
            raise ValueError("truncated FITS header")
        nbytes += BLOCK
        text = block.decode('ascii')
        for k in range(0, BLOCK, CARD):
            card = text[k:k + CARD]
            if card[:8].rstrip() == 'END':
                return header, nbytes
            key, value, comment = parse_card(card)
            if key == '' and value == '':
                continue
            header.append(key, value, comment)


def _data_nbytes(h):
    naxis = int(h.get('NAXIS', 0))
    if naxis == 0:
        return 0
    n = 1
    for i in range(1, naxis + 1):
        n *= int(h[f'NAXIS{i}'])
    n = int(h.get('GCOUNT', 1)) * (int(h.get('PCOUNT', 0)) + n)
    return abs(int(h['BITPIX'])) // 8 * n


def _locate(filename, hdu):
    """(header, data offset) of HDU number `hdu`."""
    with open(filename, 'rb') as f:
        pos = 0
        for k in range(hdu + 1):
            f.seek(pos)
            h, nh = _read_header(f)
            if k == 0 and h.get('SIMPLE') is not True:
                raise ValueError(f"{filename}: not a FITS file")
            if k == hdu:
                return h, pos + nh
            nd = _data_nbytes(h)
            pos += nh + nd + (-nd) % BLOCK
            f.seek(0, os.SEEK_END)
            if pos >= f.tell():
                raise IndexError(f"{filename}: no HDU {hdu}")
    raise IndexError(hdu)


def _scaled(raw, h):
    """Stored array -> physical values, with astropy's type rules."""
    bitpix = int(h['BITPIX'])
    bscale = h.get('BSCALE', 1)
    bzero = h.get('BZERO', 0)
    native = raw.astype(raw.dtype.newbyteorder('='))
    if bscale == 1 and bzero == 0:
        return native
    if bscale == 1:
        unsigned = {16: (1 << 15, np.uint16), 32: (1 << 31, np.uint32),
                    64: (1 << 63, np.uint64)}
        if bitpix in unsigned and bzero == unsigned[bitpix][0]:
            ut = unsigned[bitpix][1]
            return (native.view(ut) ^ ut(bzero)).astype(ut)
        if bitpix == 8 and bzero == -128:
            return (native.astype(np.int16) - 128).astype(np.int8)
    ftype = np.float32 if bitpix in (8, 16, -32) else np.float64
    out = native.astype(ftype)
    if bitpix > 0 and 'BLANK' in h:
        blank = native == h['BLANK']
    else:
        blank = None
    out *= ftype(bscale)
    out += ftype(bzero)
    if blank is not None:
        out[blank] = np.nan
    return out


def getheader(filename, hdu=0):
    """The Header of HDU number `hdu` (0: primary)."""
    return _locate(filename, hdu)[0]


def getdata(filename, hdu=0, memmap=True):
    """Data of HDU number `hdu`, physical values in native byte order,
    shape (NAXISn, ..., NAXIS1); None for an HDU without data."""
    h, off = _locate(filename, hdu)
    naxis = int(h.get('NAXIS', 0))
    if naxis == 0:
        return None
    if int(h.get('GCOUNT', 1)) != 1 or int(h.get('PCOUNT', 0)) != 0:
        raise ValueError("random-groups and table HDUs are not images")
    shape = tuple(int(h[f'NAXIS{i}']) for i in range(naxis, 0, -1))
    dt = _BITPIX[int(h['BITPIX'])]
    if memmap:
        raw = np.memmap(filename, dtype=dt, mode='r', offset=off,
                        shape=shape)
    else:
        with open(filename, 'rb') as f:
            f.seek(off)
            raw = np.frombuffer(f.read(int(np.prod(shape)) * dt.itemsize),
                                dtype=dt).reshape(shape)
    out = _scaled(raw, h)
    if out is raw or isinstance(out, np.memmap):
        out = np.array(out)
    del raw
    return out


def _storage(data):
    """Array -> (stored big-endian array, BITPIX, extra scaling cards)."""
    a = np.asarray(data)
    if a.dtype == np.bool_:
        a = a.astype(np.uint8)
    k, n = a.dtype.kind, a.dtype.itemsize
    if k == 'f' and n in (4, 8):
        return a.astype(f'>f{n}'), -8 * n, []
    if k == 'u' and n == 1:
        return a.astype('u1'), 8, []
    if k == 'i' and n in (2, 4, 8):
        return a.astype(f'>i{n}'), 8 * n, []
    if k == 'i' and n == 1:
        return ((a.astype(np.int16) + 128).astype('u1'), 8,
                [('BSCALE', 1), ('BZERO', -128)])
    if k == 'u' and n in (2, 4, 8):
        st = np.dtype(f'>i{n}')
        bias = np.array(1 << (8 * n - 1), dtype=a.dtype)
        return ((a ^ bias).view(np.dtype(f'i{n}')).astype(st), 8 * n,
            [('BSCALE', 1), ('BZERO', 1 << (8 * n - 1))])
    raise TypeError(f"no FITS image type for dtype {a.dtype}")


# ---------------------------------------------------------------- checksums

def ones_complement_sum(buf, start=0):
    """32-bit ones'-complement sum of big-endian words (end-around carry),
    continuing from `start`; len(buf) must be a multiple of 4."""
    w = np.frombuffer(buf, dtype='>u4').astype(np.uint64)
    total = int(w.sum()) + int(start)
    while total >> 32:
        total = (total & 0xFFFFFFFF) + (total >> 32)
    return total


_EXCLUDE = set(range(0x3A, 0x41)) | set(range(0x5B, 0x61))


def checksum_encode(value, complement=True):
    """16-character ASCII encoding of a 32-bit sum (FITS appendix J):
    each byte is spread over four characters from '0' upwards, avoiding
    the punctuation between the digits and the letters, then the string
    is rotated right by one so that it sums correctly when it starts at
    byte 11 of its card."""
    if complement:
        value = ~value & 0xFFFFFFFF
    asc = [0] * 16
    for i in range(4):
        byte = (value >> (24 - 8 * i)) & 0xFF
        ch = [byte // 4 + 0x30] * 4
        ch[0] += byte % 4
        changed = True
        while changed:
            changed = False
            for j in (0, 2):
                while ch[j] in _EXCLUDE or ch[j + 1] in _EXCLUDE:
                    ch[j] += 1
                    ch[j + 1] -= 1
                    changed = True
        for j in range(4):
            asc[4 * j + i] = ch[j]
    return ''.join(chr(asc[(k + 15) % 16]) for k in range(16))


def checksum_decode(text, complement=True):
    """Inverse of checksum_encode."""
    b = [ord(c) - 0x30 for c in text[1:] + text[0]]
    total = 0
    for j in range(4):
        word = (b[4 * j] << 24) | (b[4 * j + 1] << 16) | \
               (b[4 * j + 2] << 8) | b[4 * j + 3]
        total += word
    while total >> 32:
        total = (total & 0xFFFFFFFF) + (total >> 32)
    return (~total & 0xFFFFFFFF) if complement else total


def hdu_checksums(filename, hdu=0):
    """(ones'-complement sum of the whole HDU, of its data blocks).  A
    correct CHECKSUM card makes the first 0xFFFFFFFF; DATASUM holds the
    second."""
    h, off = _locate(filename, hdu)
    nd = _data_nbytes(h)
    with open(filename, 'rb') as f:
        hdr_start = _hdu_start(filename, hdu)
        f.seek(hdr_start)
        head = f.read(off - hdr_start)
        data = f.read(nd + (-nd) % BLOCK)
    ds = ones_complement_sum(data)
    return ones_complement_sum(head, ds), ds


def _hdu_start(filename, hdu):
    with open(filename, 'rb') as f:
        pos = 0
        for _ in range(hdu):
            f.seek(pos)
            h, nh = _read_header(f)
            nd = _data_nbytes(h)
            pos += nh + nd + (-nd) % BLOCK
        return pos


# ---------------------------------------------------------------- writing

_STRUCTURAL = ('SIMPLE', 'XTENSION', 'BITPIX', 'NAXIS', 'EXTEND', 'PCOUNT',
               'GCOUNT', 'BSCALE', 'BZERO', 'BLANK', 'CHECKSUM', 'DATASUM',
               'END')


def writeto(filename, data, header=None, overwrite=False, checksum=False):
    """Write `data` as the primary HDU of a new file.  The cards of
    `header` are kept except those the data fix (BITPIX, NAXISn, scaling),
    which are recomputed; with `checksum` DATASUM and CHECKSUM are added."""
    if os.path.exists(filename) and not overwrite:
        raise OSError(f"file exists: {filename}")
    stored, bitpix, scale = _storage(data)
    h = Header()
    h.append('SIMPLE', True, 'conforms to FITS standard')
    h.append('BITPIX', bitpix, 'array data type')
    h.append('NAXIS', stored.ndim, 'number of array dimensions')
    for i in range(stored.ndim):
        h.append(f'NAXIS{i + 1}', int(stored.shape[stored.ndim - 1 - i]))
    h.append('EXTEND', True)
    for key, value, comment in (header.cards() if header is not None
                                else []):
        k = key.upper()
        if k in _STRUCTURAL or (k.startswith('NAXIS') and k[5:].isdigit()):
            continue
        h.append(k, value, comment)
    for k, v in scale:
        h.append(k, v)
    body = np.ascontiguousarray(stored).tobytes()
    body += b'\0' * ((-len(body)) % BLOCK)
    if checksum:
        ds = ones_complement_sum(body)
        h.append('CHECKSUM', '0' * 16, 'HDU checksum')
        h.append('DATASUM', str(ds), 'data unit checksum')
        total = ones_complement_sum(_header_bytes(h), ds)
        h['CHECKSUM'] = checksum_encode(total)
    tmp = filename + '.tmp'
    with open(tmp, 'wb') as f:
        f.write(_header_bytes(h))
        f.write(body)
    os.replace(tmp, filename)


# ---------------------------------------------------------------- WCS

class WCS:
    """Linear world coordinates of an image (the attributes the reference
    reads from astropy.wcs.WCS(header).wcs): per FITS axis (1 = x)
    cdelt, crpix (as written, 1-based in the standard), crval, and the
    PC matrix, pc[i][j] = PC{i+1}_{j+1}."""

    def __init__(self, naxis=2, header=None):
        if header is not None:
            naxis = int(header.get('WCSAXES', header.get('NAXIS', naxis)))
        self.naxis = int(naxis)
        self.cdelt = np.ones(self.naxis)
        self.crpix = np.zeros(self.naxis)
        self.crval = np.zeros(self.naxis)
        self.pc = np.identity(self.naxis)
        if header is not None:
            for i in range(self.naxis):
                self.cdelt[i] = float(header.get(f'CDELT{i + 1}', 1.0))
                self.crpix[i] = float(header.get(f'CRPIX{i + 1}', 0.0))
                self.crval[i] = float(header.get(f'CRVAL{i + 1}', 0.0))
                for j in range(self.naxis):
                    key = f'PC{i + 1}_{j + 1}'
                    if key in header:
                        self.pc[i, j] = float(header[key])

    @property
    def wcs(self):
        """astropy's WCS.wcs view: the same object."""
        return self

    def to_header(self):
        h = Header()
        h.append('WCSAXES', self.naxis, 'number of coordinate axes')
        for i in range(self.naxis):
            h.append(f'CRPIX{i + 1}', float(self.crpix[i]),
                     'pixel coordinate of reference point')
        if not np.array_equal(self.pc, np.identity(self.naxis)):
            for i in range(self.naxis):
                for j in range(self.naxis):
                    h.append(f'PC{i + 1}_{j + 1}', float(self.pc[i, j]),
                             'coordinate transformation matrix element')
        for i in range(self.naxis):
            h.append(f'CDELT{i + 1}', float(self.cdelt[i]),
                     'coordinate increment at reference point')
        for i in range(self.naxis):
            h.append(f'CRVAL{i + 1}', float(self.crval[i]),
                     'coordinate value at reference point')
        return h


# ---------------------------------------------------------------- fitsutils

def load_fits(filename, hdu=0, memmap=True):
    """(data, WCS) of one HDU (fitsutils.py:7-17)."""
    data = getdata(filename, hdu, memmap)
    return data, WCS(header=getheader(filename, hdu))


def dump_fits(filename, data, wcs=None, overwrite=False):
    """Primary HDU with the WCS cards and checksums (fitsutils.py:20-30)."""
    header = wcs.to_header() if wcs is not None else None
    writeto(filename, data, header=header, overwrite=overwrite,
            checksum=True)
