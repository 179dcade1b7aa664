"""Data preparation: crop, mask, clip and trim observed FITS data before a
fit (the reference's prep task, src/gbkfit/tasks/prep.py:1-349, and its
CLI options, src/gbkfit/apps/cli.py:211-420).

Each prep_* function reads data / error / mask files, and then:

1. squeezes length-1 axes, sets non-finite data and error values (and
   errors <= 0) to NaN, and turns the mask file into 0/1, non-finite
   entries 0 (prep.py:8-31);
2. crops to the spatial / spectral region of interest (prep.py:54-62);
3. builds the mask: finite data, error and mask-file values (prep.py:65-72),
   times the clip_min / clip_max tests (NaN pixels kept, prep.py:75-87),
   times a sigma clip of the data (prep.py:90-92), times a connected-
   component selection of the finite pixels (prep.py:95-109);
4. applies it: NaN data and error, 0 mask (prep.py:112-117);
5. optionally trims to the bounding box of the mask (minify,
   prep.py:120-128) and pads by `nanpad` pixels on every side (NaN data
   and error, 0 mask; prep.py:131-137);
6. writes prep_<name>.fits files in `outdir` (the working directory by
   default) with the input headers (prep.py:34-51).

Sigma clipping follows astropy.stats.sigma_clip with the defaults the
reference uses (centre = median, spread = population standard deviation,
iterate at most `maxiters` times or until no value is rejected, NaNs
ignored and counted as rejected).  Connected components follow
skimage.measure.label on the boolean map of finite pixels: full
connectivity (8 neighbours in 2-D, 26 in 3-D), labels in raster order,
areas as pixel counts, components sorted by area (largest first, ties in
label order), then at most `ccl_lcount` of them, those of at least
`ccl_pcount` pixels, and those at least `ccl_lratio` times the largest.

Fixed reference behaviours (the reference keeps the first two because it
discards or ignores a result):
- the padding result is dropped in every prep_* function
  (prep.py:187,243,308,346), so `nanpad` had no effect: here it pads;
- the sigma-clip mask marks the rejected values, and the reference
  multiplies the keep-mask by it (prep.py:90-92; only moment-0 maps invert
  it), keeping only the outliers: here every clip keeps the inliers;
- prep_mmaps hands the whole sigma / iteration lists to every map
  (prep.py:298-300): here map i uses entry i;
- cropping and trimming leave the WCS reference pixel where it was
  (prep.py:34-51 writes the input header): here CRPIXn moves with the
  crop and the padding, so the world coordinates of every kept pixel are
  unchanged.
"""
import os

import numpy as np
from scipy import ndimage

from . import fitsio

__all__ = ['sigma_clip_mask', 'ccl_mask', 'prep_image', 'prep_lslit',
           'prep_mmaps', 'prep_scube']


def sigma_clip_mask(data, sigma=3.0, maxiters=5):
    """True where astropy's sigma_clip(data, sigma, maxiters).mask is
    True: the values it rejects, and the non-finite ones."""
    data = np.asarray(data)
    if not np.issubdtype(data.dtype, np.floating):
        data = data.astype(np.float64)
    kept = data[np.isfinite(data)].ravel()
    lo, hi = -np.inf, np.inf
    iteration, nchanged = 0, 1
    while nchanged != 0 and (maxiters is None or iteration < maxiters):
        iteration += 1
        size = kept.size
        if size == 0:
            break
        cen = np.median(kept)
        std = np.std(kept)
        lo, hi = cen - std * sigma, cen + std * sigma
        kept = kept[(kept >= lo) & (kept <= hi)]
        nchanged = size - kept.size
    with np.errstate(invalid='ignore'):
        return ~np.isfinite(data) | (data < lo) | (data > hi)


def ccl_mask(valid, lcount=None, pcount=None, lratio=None):
    """Keep-mask of the connected components of `valid` selected by count,
    minimum area and area ratio (prep.py:95-109)."""
    valid = np.asarray(valid, dtype=bool)
    structure = ndimage.generate_binary_structure(valid.ndim, valid.ndim)
    labels, n = ndimage.label(valid, structure=structure)
    areas = np.bincount(labels.ravel(), minlength=n + 1)[1:]
    order = sorted(range(n), key=lambda i: -areas[i])
    if lcount is not None:
        order = order[:lcount]
    if pcount is not None:
        order = [i for i in order if areas[i] >= pcount]
    if lratio is not None and order:
        largest = areas[order[0]]
        order = [i for i in order if areas[i] / largest >= lratio]
    keep = np.zeros(n + 1, dtype=bool)
    keep[np.asarray(order, dtype=np.int64) + 1] = True
    return keep[labels]


def _floating(x):
    x = np.array(np.squeeze(x))
    return x if np.issubdtype(x.dtype, np.floating) else \
        x.astype(np.float64)


def _read(file_d, file_e, file_m):
    d = _floating(fitsio.getdata(file_d))
    hd = fitsio.getheader(file_d)
    e = m = he = hm = None
    if file_e:
        e = _floating(fitsio.getdata(file_e))
        he = fitsio.getheader(file_e)
    if file_m:
        m = _floating(fitsio.getdata(file_m))
        hm = fitsio.getheader(file_m)
    d[~np.isfinite(d)] = np.nan
    if e is not None:
        e[~np.isfinite(e)] = np.nan
        with np.errstate(invalid='ignore'):
            e[e <= 0] = np.nan
    if m is not None:
        m[~np.isfinite(m)] = 0
        m[np.nonzero(m)] = 1
    return [d, e, m], [hd, he, hm]


class _Frame:
    """The arrays of one input (data, error, mask), their headers and the
    offset of the kept region in the input's pixels, per array axis."""

    def __init__(self, arrays, headers):
        self.a = arrays
        self.h = headers
        self.offset = [0] * arrays[0].ndim

    def crop(self, axis, rng):
        s = [slice(None)] * self.a[0].ndim
        s[axis] = slice(rng[0], rng[1])
        self.a = [None if x is None else x[tuple(s)] for x in self.a]
        self.offset[axis] += int(rng[0])

    def apply(self, mask):
        d, e, m = self.a
        d[mask == 0] = np.nan
        if e is not None:
            e[mask == 0] = np.nan
        if m is not None:
            m[mask == 0] = 0

    def minify(self, mask):
        nz = np.nonzero(mask)
        if nz[0].size == 0:
            raise RuntimeError("minify: the mask keeps no pixel")
        for axis, idx in enumerate(nz):
            self.crop(axis, (idx.min(), idx.max() + 1))

    def pad(self, size):
        d, e, m = self.a
        d = np.pad(d, size, 'constant', constant_values=np.nan)
        if e is not None:
            e = np.pad(e, size, 'constant', constant_values=np.nan)
        if m is not None:
            m = np.pad(m, size, 'constant', constant_values=0)
        self.a = [d, e, m]
        self.offset = [o - size for o in self.offset]

    def save(self, files, dtype, outdir):
        """prep_<name>.fits per array, CRPIX shifted by the kept offset.
        A missing mask file name gives <data name>_mask (prep.py:49)."""
        base = os.path.splitext(os.path.basename(files[0]))[0]
        names = [base,
                 os.path.splitext(os.path.basename(files[1]))[0]
                 if files[1] else None,
                 os.path.splitext(os.path.basename(files[2]))[0]
                 if files[2] else base + '_mask']
        out = []
        nd = self.a[0].ndim
        for arr, hdr, name in zip(self.a, self.h, names):
            if arr is None:
                continue
            hdr = hdr.copy() if hdr is not None else fitsio.Header()
            for axis in range(nd):
                fits_axis = nd - axis
                key = f'CRPIX{fits_axis}'
                if key in hdr or self.offset[axis]:
                    hdr[key] = float(hdr.get(key, 0.0)) - self.offset[axis]
            path = os.path.join(outdir or '.', f'prep_{name}.fits')
            fitsio.writeto(path, arr.astype(dtype), hdr, overwrite=True)
            out.append(path)
        return out


def _mask(frames, clip_min, clip_max, sclip_sigma, sclip_iters,
          ccl_lcount, ccl_pcount, ccl_lratio):
    mask = np.ones_like(frames[0].a[0])
    for i, fr in enumerate(frames):
        d, e, m = fr.a
        mask *= np.isfinite(d)
        if e is not None:
            mask *= np.isfinite(e)
        if m is not None:
            mask *= np.isfinite(m)
        with np.errstate(invalid='ignore'):
            if clip_min is not None and clip_min[i] is not None:
                mask *= ~(d < clip_min[i])
            if clip_max is not None and clip_max[i] is not None:
                mask *= ~(d > clip_max[i])
        if sclip_sigma is not None and sclip_sigma[i] is not None:
            mask *= ~sigma_clip_mask(d, sclip_sigma[i], sclip_iters[i])
        if (ccl_lcount is not None or ccl_pcount is not None
                or ccl_lratio):
            mask *= ccl_mask(np.isfinite(d), ccl_lcount, ccl_pcount,
                             ccl_lratio)
    return mask


def _listify(v, n):
    if v is None:
        return None
    if isinstance(v, (list, tuple, np.ndarray)):
        v = list(v)
        if len(v) == 1 and n > 1:
            v = v * n
        if len(v) != n:
            raise RuntimeError(f"expected {n} values, got {len(v)}")
        return v
    return [v] * n


def _run(files_list, crops, clip_min, clip_max, ccl_lcount, ccl_pcount,
         ccl_lratio, sclip_sigma, sclip_iters, minify, nanpad, dtype,
         outdir):
    frames = []
    for files in files_list:
        fr = _Frame(*_read(*files))
        for axis, rng in crops:
            if rng is not None:
                fr.crop(axis, rng)
        frames.append(fr)
    shapes = {fr.a[0].shape for fr in frames}
    if len(shapes) != 1:
        raise RuntimeError(f"inputs of different shapes: {shapes}")
    n = len(frames)
    mask = _mask(frames, _listify(clip_min, n), _listify(clip_max, n),
                 _listify(sclip_sigma, n), _listify(sclip_iters, n),
                 ccl_lcount, ccl_pcount, ccl_lratio)
    written = []
    for fr, files in zip(frames, files_list):
        fr.apply(mask)
        if minify:
            fr.minify(mask)
        if nanpad:
            fr.pad(int(nanpad))
        written += fr.save(files, np.dtype(dtype), outdir)
    return written


def prep_image(file_d, file_e=None, file_m=None, roi_spat=None,
               clip_min=None, clip_max=None, ccl_lcount=None,
               ccl_pcount=None, ccl_lratio=None, sclip_sigma=None,
               sclip_iters=5, minify=False, nanpad=None, dtype='float32',
               outdir=None):
    """Image (y, x): roi_spat = (xmin, xmax, ymin, ymax) (prep.py:140-190).
    Returns the paths written."""
    crops = []
    if roi_spat is not None:
        crops = [(1, roi_spat[0:2]), (0, roi_spat[2:4])]
    return _run([(file_d, file_e, file_m)], crops, clip_min, clip_max,
                ccl_lcount, ccl_pcount, ccl_lratio, sclip_sigma,
                sclip_iters, minify, nanpad, dtype, outdir)


def prep_lslit(file_d, file_e=None, file_m=None, roi_spat=None,
               roi_spec=None, clip_min=None, clip_max=None,
               ccl_lcount=None, ccl_pcount=None, ccl_lratio=None,
               sclip_sigma=None, sclip_iters=5, minify=False, nanpad=None,
               dtype='float32', outdir=None):
    """Long slit: roi_spat crops array axis 0 and roi_spec array axis 1,
    as the reference does (prep.py:193-246)."""
    crops = [(0, roi_spat), (1, roi_spec)]
    return _run([(file_d, file_e, file_m)], crops, clip_min, clip_max,
                ccl_lcount, ccl_pcount, ccl_lratio, sclip_sigma,
                sclip_iters, minify, nanpad, dtype, outdir)


def prep_mmaps(orders, file_d, file_e=None, file_m=None, roi_spat=None,
               clip_min=None, clip_max=None, ccl_lcount=None,
               ccl_pcount=None, ccl_lratio=None, sclip_sigma=None,
               sclip_iters=5, minify=False, nanpad=None, dtype='float32',
               outdir=None):
    """Moment maps: one common mask over all the maps (prep.py:249-313);
    clip_min / clip_max / sclip_sigma / sclip_iters per map."""
    n = len(file_d)
    if len(orders) != n:
        raise RuntimeError(f"{len(orders)} orders for {n} moment maps")
    file_e = list(file_e) if file_e is not None else [None] * n
    file_m = list(file_m) if file_m is not None else [None] * n
    crops = []
    if roi_spat is not None:
        crops = [(1, roi_spat[0:2]), (0, roi_spat[2:4])]
    return _run(list(zip(file_d, file_e, file_m)), crops, clip_min,
                clip_max, ccl_lcount, ccl_pcount, ccl_lratio, sclip_sigma,
                sclip_iters, minify, nanpad, dtype, outdir)


def prep_scube(file_d, file_e=None, file_m=None, roi_spat=None,
               roi_spec=None, clip_min=None, clip_max=None,
               ccl_lcount=None, ccl_pcount=None, ccl_lratio=None,
               sclip_sigma=None, sclip_iters=5, minify=False, nanpad=None,
               dtype='float32', outdir=None):
    """Spectral cube (z, y, x): roi_spat = (xmin, xmax, ymin, ymax),
    roi_spec = (zmin, zmax) (prep.py:316-349)."""
    crops = []
    if roi_spat is not None:
        crops = [(2, roi_spat[0:2]), (1, roi_spat[2:4])]
    crops.append((0, roi_spec))
    return _run([(file_d, file_e, file_m)], crops, clip_min, clip_max,
                ccl_lcount, ccl_pcount, ccl_lratio, sclip_sigma,
                sclip_iters, minify, nanpad, dtype, outdir)
