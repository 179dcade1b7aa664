"""Dataset ingest on the CPU: FITS images and checksums, WCS, Data /
Dataset load and dump, and the prep task.

The first two tests restate the reference's own expectations
(tests/unit/test_dataset.py:7-120).  FITS parsing is pinned by byte
strings assembled here card by card from the FITS standard, the checksum
by the standard's worked example (868229149 <-> 'hcHjjc9ghcEghc9g') and by
the negative-zero sum of every written HDU; sigma clipping and the
connected components by hand-computed cases.
"""
import os

import numpy as np
import pytest

from gbkfit_amd import fitsio, prep
from gbkfit_amd.dataset import (Data, DatasetImage, DatasetLSlit,
                                DatasetMMaps, DatasetSCube, check_compatible,
                                data_parser, dataset_parser, dmodel_options)


def test_data(tmp_path):
    """tests/unit/test_dataset.py:7-78 of the reference."""
    size = (5, 2)
    shape = size[::-1]
    data_d = np.full(shape, 10.0)
    data_m = np.full(shape, 1.0)
    data_e = np.full(shape, 2.0)
    data_ones = np.ones(shape)
    data01 = Data(data_d)
    assert data01.data() is not data_d
    assert data01.dtype() == data_d.dtype
    assert np.array_equal(data01.data(), data_d)
    assert np.array_equal(data01.mask(), data_ones)
    assert data01.error() is None
    assert data01.data().size == data01.npix()
    assert data01.ndim() == len(size)
    assert data01.size() == size
    assert data01.step() == (1, 1)
    assert data01.zero() == (-2.0, -0.5)
    assert data01.rpix() == (2.0, 0.5)
    assert data01.rval() == (0, 0)
    assert data01.rota() == 0
    data02 = Data(data_d, mask=data_m, error=data_e)
    assert data02.data() is not data_d
    assert data02.mask() is not data_m
    assert data02.error() is not data_e
    assert np.array_equal(data02.data(), data_d)
    assert np.array_equal(data02.mask(), data_m)
    assert np.array_equal(data02.error(), data_e)
    data03 = Data(data_d, step=2, rpix=3, rval=4, rota=5)
    assert data03.step() == (2, 2)
    assert data03.rpix() == (3, 3)
    assert data03.rval() == (4, 4)
    assert data03.rota() == 5
    data04 = Data(data_d, error=data_e, step=(1, 2), rpix=(3, 4),
                  rval=(5, 6), rota=7)
    assert (data04.step(), data04.rpix(), data04.rval(), data04.rota()) \
        == ((1, 2), (3, 4), (5, 6), 7)
    fd, fm, fe = (str(tmp_path / f'data_{k}.fits') for k in 'dme')
    dumped = data_parser.dump(data04, fd, fm, fe, overwrite=True)
    assert dumped == dict(data=fd, mask=fm, error=fe, step=(1, 2),
                          rpix=(3, 4), rval=(5, 6), rota=7)
    loaded = data_parser.load(dumped)
    assert loaded.size() == data04.size()
    assert loaded.step() == data04.step()
    assert loaded.rpix() == data04.rpix()
    assert loaded.rval() == data04.rval()
    assert loaded.rota() == data04.rota()
    assert np.array_equal(loaded.data(), data04.data())
    assert np.array_equal(loaded.error(), data04.error())


def test_dataset_image(tmp_path, monkeypatch):
    """tests/unit/test_dataset.py:81-120 of the reference."""
    monkeypatch.chdir(tmp_path)
    shape = (2, 5)
    data01 = Data(np.full(shape, 10.0), np.full(shape, 1.0),
                  np.full(shape, 2.0))
    image01 = DatasetImage(data01)
    assert image01.npix() == data01.npix()
    assert image01.size() == data01.size()
    assert image01.step() == data01.step()
    assert image01.zero() == data01.zero()
    assert image01.dtype == data01.dtype()
    dumped = dataset_parser.dump(image01, overwrite=True)
    assert dumped == dict(
        type='image',
        image=dict(data='image_d.fits', mask='image_m.fits',
                   error='image_e.fits'),
        step=(1.0, 1.0), rpix=(2.0, 0.5), rval=(0.0, 0.0), rota=0)
    loaded = dataset_parser.load(dumped)
    assert isinstance(loaded, DatasetImage)
    for m in ('size', 'step', 'rpix', 'rval', 'rota'):
        assert getattr(loaded, m)() == getattr(image01, m)()


def test_data_total_mask_and_caller_arrays():
    """NaN data, a zero mask entry and a NaN error each drop the pixel from
    the total mask (data.py:188-201); the caller's arrays stay as given."""
    d = np.arange(6, dtype=np.float32).reshape(2, 3)
    d[0, 1] = np.nan
    m = np.ones_like(d)
    m[1, 0] = 0
    e = np.full_like(d, 0.5)
    e[1, 2] = np.inf
    d0, m0, e0 = d.copy(), m.copy(), e.copy()
    x = Data(d, m, e)
    want = np.array([[1, 0, 1], [0, 1, 0]], np.float32)
    assert np.array_equal(x.mask(), want)
    assert np.array_equal(np.isnan(x.data()), want == 0)
    assert np.array_equal(np.isnan(x.error()), want == 0)
    assert np.array_equal(d, d0, equal_nan=True)
    assert np.array_equal(m, m0) and np.array_equal(e, e0)
    assert x.dtype() == np.float32 and x.mask().dtype == np.float32
    with pytest.raises(RuntimeError):
        Data(d, m[:, :2])
    with pytest.raises(RuntimeError):
        Data(d, step=(1, 2, 3))
    with pytest.raises(RuntimeError):
        Data(d, mask=np.full_like(d, np.nan))
    # integer data become float32 (data.py:28-33)
    assert Data(np.ones((2, 2), np.int16)).dtype() == np.float32


# ---------------------------------------------------------------- FITS

def _card(text):
    assert len(text) <= 80
    return text.ljust(80).encode('ascii')


def _hdu(cards, payload):
    head = b''.join(_card(c) for c in cards) + _card('END')
    head += b' ' * ((-len(head)) % 2880)
    return head + payload + b'\0' * ((-len(payload)) % 2880)


def test_fits_hand_assembled(tmp_path):
    """Files laid out by hand per the FITS standard: fixed-format cards,
    big-endian data, NAXIS1 fastest, scaled and unsigned integers, BLANK,
    and an IMAGE extension after an empty primary."""
    vals = np.array([[0, 1, 65535], [32768, 2, 3]], dtype=np.uint16)
    stored = (vals.astype(np.int32) - 32768).astype('>i2')
    raw = _hdu(["SIMPLE  =                    T / standard",
                "BITPIX  =                   16",
                "NAXIS   =                    2",
                "NAXIS1  =                    3",
                "NAXIS2  =                    2",
                "BSCALE  =                  1.0",
                "BZERO   =              32768.0",
                "CDELT1  =                 2.5D0",
                "CRPIX1  =                   4.",
                "OBJECT  = 'NGC ''X'''          / name with quotes",
                "COMMENT   anything at all / here"], stored.tobytes())
    p = tmp_path / 'u16.fits'
    p.write_bytes(raw)
    got = fitsio.getdata(str(p))
    assert got.dtype == np.uint16 and np.array_equal(got, vals)
    h = fitsio.getheader(str(p))
    assert h['OBJECT'] == "NGC 'X'" and h['CDELT1'] == 2.5
    w = fitsio.WCS(header=h)
    assert list(w.cdelt) == [2.5, 1.0] and list(w.crpix) == [4.0, 0.0]
    # scaled integers with BLANK -> float32 with NaN
    s = np.array([[1, -5], [7, 100]], dtype='>i2')
    raw = _hdu(["SIMPLE  =                    T", "BITPIX  =                   16",
                "NAXIS   =                    2", "NAXIS1  =                    2",
                "NAXIS2  =                    2", "BSCALE  =                  0.5",
                "BZERO   =                 10.0", "BLANK   =                  100"],
               s.tobytes())
    p2 = tmp_path / 'scaled.fits'
    p2.write_bytes(raw)
    got = fitsio.getdata(str(p2))
    assert got.dtype == np.float32
    assert np.array_equal(got, np.array([[10.5, 7.5], [13.5, np.nan]],
                                        np.float32), equal_nan=True)
    # empty primary + IMAGE extension of big-endian float64, 3-D
    cube = (np.arange(24).reshape(2, 3, 4) / 7).astype('>f8')
    raw = _hdu(["SIMPLE  =                    T", "BITPIX  =                    8",
                "NAXIS   =                    0", "EXTEND  =                    T"],
               b'')
    raw += _hdu(["XTENSION= 'IMAGE   '", "BITPIX  =                  -64",
                 "NAXIS   =                    3", "NAXIS1  =                    4",
                 "NAXIS2  =                    3", "NAXIS3  =                    2",
                 "PCOUNT  =                    0", "GCOUNT  =                    1"],
                cube.tobytes())
    p3 = tmp_path / 'ext.fits'
    p3.write_bytes(raw)
    assert fitsio.getdata(str(p3)) is None
    got = fitsio.getdata(str(p3), hdu=1, memmap=False)
    assert got.dtype == np.float64 and got.dtype.isnative
    assert np.array_equal(got, cube.astype(np.float64))
    with pytest.raises(IndexError):
        fitsio.getdata(str(p3), hdu=2)


def test_fits_cards_fixed_format():
    c = fitsio.format_card('NAXIS1', 300, 'length')
    assert c[:30] == 'NAXIS1  = ' + ' ' * 17 + '300'
    assert c[30:] == ' / length'.ljust(50)
    c = fitsio.format_card('SIMPLE', True)
    assert c[29] == 'T' and c[:10] == 'SIMPLE  = '
    c = fitsio.format_card('OBJECT', "a'b")
    assert c[10:20] == "'a''b    '" and c[20:] == ' ' * 60
    for v in (0.0, -2.5, 1e-30, 6.02214076e23, 0.1):
        c = fitsio.format_card('X', v)
        assert fitsio.parse_card(c)[1] == v
    assert fitsio.parse_card(fitsio.format_card('X', -7))[1] == -7


def test_checksum_known_answer():
    """The FITS standard's worked example of the ASCII encoding."""
    assert fitsio.checksum_encode(868229149) == 'hcHjjc9ghcEghc9g'
    assert fitsio.checksum_decode('hcHjjc9ghcEghc9g') == 868229149
    # negative zero encodes as the sixteen '0' characters
    assert fitsio.checksum_encode(0xFFFFFFFF) == '0' * 16
    rng = np.random.default_rng(3)
    for v in rng.integers(0, 2 ** 32, 200):
        e = fitsio.checksum_encode(int(v))
        assert all(c.isalnum() for c in e)
        assert fitsio.checksum_decode(e) == int(v)
    # ones' complement sum with end-around carry
    buf = np.array([0xFFFFFFFF, 0x00000002], '>u4').tobytes()
    assert fitsio.ones_complement_sum(buf) == 2


@pytest.mark.parametrize('dtype', ['f4', 'f8', 'u1', 'i1', 'i2', 'u2',
                                   'i4', 'u4', 'i8', 'u8'])
def test_fits_round_trip_and_checksum(tmp_path, dtype):
    rng = np.random.default_rng(7)
    if dtype[0] == 'f':
        a = rng.normal(0, 1e3, (3, 5, 7)).astype(dtype)
        a[0, 0, 0] = np.nan
    else:
        info = np.iinfo(dtype)
        a = rng.integers(info.min, info.max, (3, 5, 7), dtype=dtype,
                         endpoint=True)
        a.flat[0], a.flat[1] = info.min, info.max
    wcs = fitsio.WCS(naxis=3)
    wcs.cdelt[:] = (1.5, 2.0, 10.0)
    wcs.crpix[:] = (3, 2, 1)
    p = str(tmp_path / f'{dtype}.fits')
    fitsio.dump_fits(p, a, wcs)
    b, w2 = fitsio.load_fits(p)
    assert b.dtype == np.dtype(dtype) and b.dtype.isnative
    assert np.array_equal(a, b, equal_nan=(dtype[0] == 'f'))
    assert list(w2.cdelt) == [1.5, 2.0, 10.0]
    total, datasum = fitsio.hdu_checksums(p)
    assert total == 0xFFFFFFFF
    assert fitsio.getheader(p)['DATASUM'] == str(datasum)
    assert os.path.getsize(p) % 2880 == 0
    with pytest.raises(OSError):
        fitsio.dump_fits(p, a, wcs)


# ---------------------------------------------------------------- datasets

def test_data_load_from_header(tmp_path):
    """No step / rpix / rval in the options: the header's CDELT, CRPIX (as
    written) and CRVAL; a scalar error; options given globally."""
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    h = fitsio.Header([('CDELT1', 2.0), ('CDELT2', 3.0), ('CRPIX1', 1.5),
                       ('CRPIX2', 1.0), ('CRVAL1', 10.0), ('CRVAL2', -4.0)])
    p = str(tmp_path / 'img.fits')
    fitsio.writeto(p, a, h)
    d = Data.load(dict(data=p, error=0.25))
    assert d.step() == (2.0, 3.0) and d.rpix() == (1.5, 1.0)
    assert d.rval() == (10.0, -4.0)
    assert d.zero() == (10.0 - 1.5 * 2.0, -4.0 - 1.0 * 3.0)
    assert np.all(d.error() == 0.25)
    d = Data.load(dict(data=p, rpix=(0, 0)), step=(1, 1))
    assert d.step() == (1, 1) and d.rpix() == (0, 0)
    with pytest.raises(RuntimeError):
        Data.load(dict(mask=p))


def test_dataset_load_dump_types(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    rng = np.random.default_rng(0)
    cube = Data(rng.normal(size=(4, 3, 5)).astype(np.float32), step=(1, 1, 10))
    sc = DatasetSCube(cube)
    info = sc.dump(prefix='g1_')
    assert info['type'] == 'scube' and info['scube']['data'] == 'g1_scube_d.fits'
    sc2 = dataset_parser.load(info)
    assert isinstance(sc2, DatasetSCube) and sc2.size() == (5, 3, 4)
    assert sc2.step() == (1, 1, 10)
    assert np.array_equal(sc2['scube'].data(), cube.data())
    # a scube description read as an image keeps the first two entries of
    # its dimensional options; a scalar is repeated
    img = Data(np.ones((3, 5), np.float32))
    DatasetImage(img).dump(prefix='i_')
    loaded = dataset_parser.load(dict(
        type='image', image=dict(data='i_image_d.fits'), step=(2, 3, 10),
        rpix=1.0, rval=[0, 0]))
    assert loaded.step() == (2, 3) and loaded.rpix() == (1.0, 1.0)
    ls = DatasetLSlit(Data(np.ones((6, 5), np.float32)))
    assert dataset_parser.load(ls.dump(prefix='l_')).type() == 'lslit'
    mm = DatasetMMaps(mmap0=img, mmap2=Data(np.zeros((3, 5), np.float32)))
    assert sorted(mm.keys()) == ['mmap0', 'mmap2']
    mm2 = dataset_parser.load(mm.dump(prefix='m_'))
    assert sorted(mm2.keys()) == ['mmap0', 'mmap2']
    with pytest.raises(RuntimeError):
        DatasetMMaps(mmap0=img, mmap1=Data(np.ones((3, 4), np.float32)))
    with pytest.raises(RuntimeError):
        dataset_parser.dump_many([sc, ls], prefix=['a', 'a'])
    with pytest.raises(RuntimeError):
        dataset_parser.load(dict(type='nope'))
    with pytest.raises(RuntimeError):
        dataset_parser.load(dict(type='image', image=dict(
            data='i_image_d.fits'), step=[1]))


def test_dataset_describes_dmodel():
    """A data model built from a dataset's grid passes the Objective's
    checks; another step fails them (dmodels/_detail.py:59-75,
    objective.py:79-110)."""
    from gbkfit_amd.dmodels import DModelMMaps, DModelSCube
    cube = Data(np.ones((8, 6, 10), np.float32), step=(1, 1, 10),
                rpix=(4.5, 2.5, 3.5), rval=(0, 0, 0))
    ds = DatasetSCube(cube)
    opts = dmodel_options(ds, 3, scale=2)
    assert opts['size'] == (10, 6, 8) and opts['scale'] == (2, 2, 2)
    assert opts['dtype'] == 'float32'
    dm = DModelSCube(**opts)
    check_compatible(ds, dm)
    with pytest.raises(RuntimeError, match='step'):
        check_compatible(ds, DModelSCube(**dict(opts, step=(1, 1, 5))))
    mm = DatasetMMaps(mmap0=Data(np.ones((6, 10), np.float32)),
                      mmap1=Data(np.ones((6, 10), np.float32)))
    dmm = DModelMMaps(**dmodel_options(mm, 2), orders=(0, 1))
    check_compatible(mm, dmm)
    with pytest.raises(RuntimeError, match='incompatible'):
        check_compatible(mm, DModelMMaps(**dmodel_options(mm, 2),
                                         orders=(0, 1, 2)))


# ---------------------------------------------------------------- prep

def test_sigma_clip_known_answers():
    # [1,2,3,4,100]: median 3, std 39.01 -> bounds 3 +- 58.5 rejects 100;
    # then [1,2,3,4]: median 2.5, std 1.118, bounds [0.82, 4.18]: stable
    x = np.array([1, 2, 3, 4, 100.0])
    assert np.array_equal(prep.sigma_clip_mask(x, 1.5),
                          [False, False, False, False, True])
    # a second iteration rejects more: [0,0,0,0,0,0,0,0,10,30] sigma 1.5:
    # median 0, std 9.22 -> rejects 30 (> 13.8); then median 0, std 3.14
    # -> rejects 10 (> 4.7); then all zeros, std 0: stable
    y = np.array([0.0] * 8 + [10.0, 30.0])
    assert np.array_equal(prep.sigma_clip_mask(y, 1.5),
                          [False] * 8 + [True, True])
    # one iteration only: 10 stays
    assert np.array_equal(prep.sigma_clip_mask(y, 1.5, maxiters=1),
                          [False] * 9 + [True])
    # NaN counted as rejected, ignored in the statistics
    z = np.array([1, 2, 3, 4, 100.0, np.nan])
    assert np.array_equal(prep.sigma_clip_mask(z, 1.5),
                          [False] * 4 + [True, True])


def test_ccl_known_answers():
    v = np.array([[1, 1, 0, 0, 0, 1],
                  [0, 0, 1, 0, 0, 1],
                  [0, 0, 0, 0, 0, 0],
                  [1, 0, 0, 1, 1, 1]], bool)
    # full connectivity: the diagonal (0,1)-(1,2) joins -> components of
    # 3, 2, 1 and 3 pixels, in raster order of their first pixel
    keep = prep.ccl_mask(v, lcount=1)
    assert keep.sum() == 3 and keep[0, 0] and keep[1, 2]
    keep = prep.ccl_mask(v, lcount=2)
    assert keep.sum() == 6 and keep[3, 5] and not keep[0, 5]
    assert prep.ccl_mask(v, pcount=2).sum() == 8
    assert prep.ccl_mask(v, lratio=0.9).sum() == 6
    assert prep.ccl_mask(v, lratio=0.5).sum() == 8
    assert prep.ccl_mask(np.zeros((3, 3), bool), lratio=0.5).sum() == 0


def _write(path, a, **cards):
    fitsio.writeto(str(path), a, fitsio.Header(list(cards.items())))
    return str(path)


def test_prep_image(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.normal(1.0, 0.1, (12, 16)).astype(np.float32)
    img[0, :] = np.nan
    img[5, 7] = 50.0                       # outlier
    img[10, 14] = -3.0                     # below clip_min
    err = np.full_like(img, 0.1)
    err[3, 3] = -1.0                       # non-positive error -> masked
    msk = np.ones_like(img)
    msk[4, 4] = 0.0
    fd = _write(tmp_path / 'img.fits', img, CRPIX1=8.0, CRPIX2=6.0,
                CDELT1=1.0, CDELT2=1.0)
    fe = _write(tmp_path / 'err.fits', err)
    fm = _write(tmp_path / 'msk.fits', msk)
    out = prep.prep_image(fd, fe, fm, roi_spat=(2, 15, 0, 11), clip_min=0.0,
                          sclip_sigma=5.0, minify=True, nanpad=2,
                          outdir=str(tmp_path))
    assert [os.path.basename(p) for p in out] == [
        'prep_img.fits', 'prep_err.fits', 'prep_msk.fits']
    d = fitsio.getdata(out[0])
    e = fitsio.getdata(out[1])
    m = fitsio.getdata(out[2])
    # crop to rows 0..10, columns 2..14, row 0 is NaN -> minify to rows
    # 1..10; pad by 2
    assert d.shape == (10 + 4, 13 + 4)
    core = d[2:-2, 2:-2]
    src = img[1:11, 2:15]
    assert np.isnan(core[5 - 1, 7 - 2])          # outlier clipped
    assert np.isnan(core[10 - 1, 14 - 2])        # clip_min
    assert np.isnan(core[3 - 1, 3 - 2]) and np.isnan(e[2 + 2, 2 + 1])
    assert np.all(np.isnan(d[:2])) and np.all(m[:, :2] == 0)
    ok = ~np.isnan(core)
    assert np.array_equal(core[ok], src[ok])
    # the input mask's zero is not a NaN in the data (prep.py:65-72
    # tests finiteness only) but stays 0 in the mask
    assert not np.isnan(core[4 - 1, 4 - 2]) and m[2 + 3, 2 + 2] == 0
    # the world coordinate of every kept pixel is unchanged
    h = fitsio.getheader(out[0])
    assert h['CRPIX1'] == 8.0 - 2 + 2 and h['CRPIX2'] == 6.0 - 1 + 2


def test_prep_scube_and_mmaps(tmp_path):
    rng = np.random.default_rng(6)
    cube = rng.normal(0, 1, (1, 8, 6, 5)).astype(np.float64)  # squeezed
    fd = _write(tmp_path / 'cube.fits', cube)
    out = prep.prep_scube(fd, roi_spat=(1, 4, 0, 6), roi_spec=(2, 7),
                          dtype='float64', outdir=str(tmp_path))
    assert len(out) == 1
    got = fitsio.getdata(out[0])
    assert got.dtype == np.float64
    assert np.array_equal(got, cube[0, 2:7, 0:6, 1:4])
    # moment maps: one common mask; a NaN in map 1 masks map 0 too;
    # per-map clip ranges
    m0 = rng.uniform(1, 2, (5, 7)).astype(np.float32)
    m1 = rng.uniform(-100, 100, (5, 7)).astype(np.float32)
    m1[2, 3] = np.nan
    m1[0, 0] = 500.0
    f0 = _write(tmp_path / 'm0.fits', m0)
    f1 = _write(tmp_path / 'm1.fits', m1)
    out = prep.prep_mmaps((0, 1), [f0, f1], clip_max=[None, 200.0],
                          outdir=str(tmp_path))
    g0, g1 = (fitsio.getdata(p) for p in out)
    assert np.isnan(g0[2, 3]) and np.isnan(g0[0, 0])
    assert np.isnan(g1[0, 0])
    assert np.isfinite(g0).sum() == 35 - 2
    ls = rng.normal(0, 1, (6, 9)).astype(np.float32)
    fl = _write(tmp_path / 'ls.fits', ls)
    out = prep.prep_lslit(fl, roi_spat=(1, 5), roi_spec=(2, 8),
                          outdir=str(tmp_path))
    assert np.array_equal(fitsio.getdata(out[0]), ls[1:5, 2:8])


def test_masked_residual_is_zero():
    """A zero data-mask or model-mask entry gives a zero residual, also
    where the masked value is the NaN that marks it; an unmasked NaN still
    propagates (objective.hpp:7-33 multiplies, so NaN * 0 = NaN there)."""
    import oracle
    od = np.array([1.0, np.nan, np.nan, 2.0, 3.0])
    om = np.array([1.0, 0.0, 1.0, 1.0, 1.0])
    md = np.array([1.5, 1.0, 1.0, np.nan, 2.0])
    mm = np.array([1.0, 1.0, 1.0, 0.0, 1.0])
    r = oracle.residual(od, None, om, md, None, mm, 2.0)
    assert r[0] == 1.0 and r[1] == 0.0 and np.isnan(r[2])
    assert r[3] == 0.0 and r[4] == -2.0
