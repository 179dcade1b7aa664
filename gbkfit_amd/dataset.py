"""Datasets: observed arrays with mask, error and world coordinates.

Mirror of the reference's dataset layer:

- ``Data`` (src/gbkfit/dataset/data.py:36-253): one data array with its
  mask and error, step / rpix / rval / rota, the total mask (finite data,
  nonzero mask, finite error) applied as NaN, and load / dump through FITS
  files (``gbkfit_amd.fitsio``) with WCS header cards;
- ``Dataset`` (src/gbkfit/dataset/core.py:23-97): named Data items that
  share size, step, rpix, rval, rota and dtype;
- ``DatasetImage``, ``DatasetLSlit``, ``DatasetMMaps``, ``DatasetSCube``
  (src/gbkfit/dataset/datasets/*.py) with the common load / dump of
  datasets/_detail.py:15-104 (dimensional options given as a scalar or a
  sequence trimmed to the dataset's dimensionality; global coordinates
  passed down to the items that define none);
- ``data_parser`` / ``dataset_parser``: load(info) / dump(obj, ...) as the
  reference's parsers call them (``dataset_parser`` dispatches on
  ``info['type']``).

``dmodel_options`` returns the grid options a data model takes from a
dataset (src/gbkfit/model/dmodels/_detail.py:51-99), and
``check_compatible`` the checks the reference's Objective makes between a
dataset and its data model (src/gbkfit/objective/objective.py:79-110).

Reference behaviours kept: CRPIX is taken from the header as written
(1-based in the FITS standard) and used as the 0-based rpix, as the
reference does through astropy; rotation comes only from the options.
Fixed: the reference's Data writes the total mask and the NaNs into the
caller's own arrays when they are already native floating point; here the
caller's arrays are left alone.
"""
import logging
import os
import typing
from numbers import Real

import numpy as np

from . import fitsio

__all__ = ['Data', 'Dataset', 'DatasetImage', 'DatasetLSlit',
           'DatasetMMaps', 'DatasetSCube', 'data_parser', 'dataset_parser',
           'dmodel_options', 'check_compatible']

_log = logging.getLogger(__name__)


def _native_float(x, label):
    x = np.array(x, copy=True)
    if not x.dtype.isnative:
        x = x.byteswap().view(x.dtype.newbyteorder('='))
    if not np.issubdtype(x.dtype, np.floating):
        _log.warning(f"{label} array is not of floating type; "
                     f"will convert to float32")
        x = x.astype(np.float32)
    return x


def _make_filename(filename, dump_path):
    return filename if dump_path else os.path.basename(filename)


class Data:
    """One observed array (data.py:36-253)."""

    @classmethod
    def load(cls, info, step=None, rpix=None, rval=None, rota=None,
             prefix=''):
        """From a dict naming FITS files (data.py:38-84): 'data' (required),
        'mask' (a file), 'error' (a file or a number); step / rpix / rval /
        rota from the dict, else the arguments, else the data's header."""
        info = dict(info)
        if info.get('data') is None:
            raise RuntimeError("option 'data' is required but not provided")
        data_d, wcs_d = fitsio.load_fits(prefix + info['data'])
        data_m = data_e = None
        if info.get('mask') is not None:
            data_m = fitsio.load_fits(prefix + info['mask'])[0]
        error = info.get('error')
        if error is not None:
            if isinstance(error, (int, float)) and not isinstance(error,
                                                                  bool):
                data_e = np.full_like(data_d, error)
            elif isinstance(error, str):
                data_e = fitsio.load_fits(prefix + error)[0]
        step = info.get('step', step)
        rpix = info.get('rpix', rpix)
        rval = info.get('rval', rval)
        rota = info.get('rota', rota)
        if step is None:
            step = wcs_d.cdelt.tolist()
        if rpix is None:
            rpix = wcs_d.crpix.tolist()
        if rval is None:
            rval = wcs_d.crval.tolist()
        unknown = set(info) - {'data', 'mask', 'error', 'step', 'rpix',
                               'rval', 'rota'}
        if unknown:
            raise RuntimeError(f"unknown data options: {sorted(unknown)}")
        return cls(data_d, data_m, data_e, step, rpix, rval, rota)

    def dump(self, filename_d, filename_m=None, filename_e=None,
             dump_wcs=True, dump_path=True, overwrite=False):
        """Write data / mask / error as FITS files with the WCS cards;
        returns the dict that load() reads back (data.py:86-131)."""
        info = {}
        dat, msk, err = self.data(), self.mask(), self.error()
        step, rpix, rval, rota = (self.step(), self.rpix(), self.rval(),
                                  self.rota())
        wcs = fitsio.WCS(naxis=dat.ndim)
        wcs.cdelt[:] = step
        wcs.crpix[:] = rpix
        wcs.crval[:] = rval
        wcs.pc[0][0] = +np.cos(rota)
        wcs.pc[0][1] = -np.sin(rota)
        wcs.pc[1][0] = +np.sin(rota)
        wcs.pc[1][1] = +np.cos(rota)
        if dump_wcs:
            info.update(dict(step=step, rpix=rpix, rval=rval, rota=rota))
        info['data'] = filename_d = _make_filename(filename_d, dump_path)
        fitsio.dump_fits(filename_d, dat, wcs, overwrite)
        if filename_m and msk is not None:
            info['mask'] = filename_m = _make_filename(filename_m, dump_path)
            fitsio.dump_fits(filename_m, msk, wcs, overwrite)
        if filename_e and err is not None:
            info['error'] = filename_e = _make_filename(filename_e,
                                                        dump_path)
            fitsio.dump_fits(filename_e, err, wcs, overwrite)
        return info

    def __init__(self, data, mask=None, error=None, step=None, rpix=None,
                 rval=None, rota=None):
        data = np.asarray(data)
        if mask is None:
            mask = np.ones_like(data)
        if step is None:
            step = (1,) * data.ndim
        if rpix is None:
            rpix = tuple((np.asarray(data.shape[::-1]) / 2 - 0.5).tolist())
        if rval is None:
            rval = (0,) * data.ndim
        if rota is None:
            rota = 0
        if isinstance(step, Real):
            step = (step,) * data.ndim
        if isinstance(rpix, Real):
            rpix = (rpix,) * data.ndim
        if isinstance(rval, Real):
            rval = (rval,) * data.ndim
        data = _native_float(data, 'data')
        mask = _native_float(mask, 'mask')
        if error is not None:
            error = _native_float(error, 'error')
        if np.any(~np.isfinite(mask)):
            raise RuntimeError("mask contains non-finite values")
        if data.shape != mask.shape:
            raise RuntimeError(
                f"data and mask have incompatible shapes "
                f"({data.shape} != {mask.shape})")
        if error is not None and data.shape != error.shape:
            raise RuntimeError(
                f"data and error have incompatible shapes "
                f"({data.shape} != {error.shape})")
        for name, v in (('step', step), ('rpix', rpix), ('rval', rval)):
            if data.ndim != len(v):
                raise RuntimeError(
                    f"data dimensionality and {name} length are "
                    f"incompatible ({data.ndim} != {len(v)})")
        total = np.isfinite(data) & (mask != 0)
        if error is not None:
            total &= np.isfinite(error)
        data[~total] = np.nan
        if error is not None:
            error[~total] = np.nan
        dtype = data.dtype
        zero = (np.array(rval) - np.array(rpix) * np.array(step)).tolist()
        self._data = data
        self._mask = total.astype(dtype)
        self._error = error.astype(dtype) if error is not None else None
        self._step = tuple(step)
        self._zero = tuple(zero)
        self._rpix = tuple(rpix)
        self._rval = tuple(rval)
        self._rota = rota

    def ndim(self):
        return self._data.ndim

    def npix(self):
        return self._data.size

    def size(self):
        return self._data.shape[::-1]

    def step(self):
        return self._step

    def zero(self):
        return self._zero

    def rpix(self):
        return self._rpix

    def rval(self):
        return self._rval

    def rota(self):
        return self._rota

    def data(self):
        return self._data

    def mask(self):
        return self._mask

    def error(self):
        return self._error

    def dtype(self):
        return self._data.dtype


class _DataParser:
    """data_parser of data.py:253 (load(info, ...), dump(obj, ...))."""

    @staticmethod
    def load(info, *args, **kwargs):
        return Data.load(info, *args, **kwargs)

    @staticmethod
    def dump(obj, *args, **kwargs):
        return obj.dump(*args, **kwargs)


data_parser = _DataParser()


def _same(data, method):
    attr = {k: getattr(v, method)() for k, v in data.items()}
    if len(set(attr.values())) > 1:
        raise RuntimeError(
            f"dataset contains data items of different {method}: "
            f"{str(attr)}")


class Dataset:
    """Named Data items with one grid (core.py:23-97)."""

    def __init__(self, data):
        if not data:
            raise RuntimeError("dataset contains no data items")
        invalid = [k for k, v in data.items() if not isinstance(v, Data)]
        if invalid:
            raise RuntimeError(f"dataset contains invalid data items: "
                               f"{invalid}")
        for method in ('size', 'step', 'rpix', 'rval', 'rota', 'dtype'):
            _same(data, method)
        self._data = {k: v for k, v in data.items()}

    def __contains__(self, item):
        return item in self._data

    def __getitem__(self, item):
        return self._data[item]

    def __iter__(self):
        return iter(self._data)

    def items(self):
        return self._data.items()

    def keys(self):
        return self._data.keys()

    def values(self):
        return self._data.values()

    def get(self, item, default=None):
        return self._data.get(item, default)

    def _first(self):
        return next(iter(self._data.values()))

    def npix(self):
        return self._first().npix()

    def size(self):
        return self._first().size()

    def step(self):
        return self._first().step()

    def zero(self):
        return self._first().zero()

    def rpix(self):
        return self._first().rpix()

    def rval(self):
        return self._first().rval()

    def rota(self):
        return self._first().rota()

    @property
    def dtype(self):
        return self._first().dtype()

    def dump(self, **kwargs):
        return _dump_common(self, **kwargs)


def _sanitize(option, value, length, type_):
    """Scalar -> repeated; longer sequence -> trimmed (_detail.py:15-48)."""
    types = typing.get_args(type_) or (type_,)
    names = ' | '.join(t.__name__ for t in types)
    if isinstance(value, types) and not isinstance(value, bool):
        return [value] * length
    if (isinstance(value, (list, tuple))
            and all(isinstance(v, types) and not isinstance(v, bool)
                    for v in value)):
        if len(value) < length:
            raise RuntimeError(
                f"option '{option}' has a value with a length shorter "
                f"than expected; expected length: {length}, current "
                f"length: {len(value)}")
        if len(value) > length:
            _log.warning(f"option '{option}' trimmed from {value} to "
                         f"{value[:length]}")
            value = value[:length]
        return list(value)
    raise RuntimeError(
        f"option '{option}' should be a scalar of type {names}, or a "
        f"sequence of type {names} and length of {length}")


def _load_common(cls, info, names, ndim, prefix=''):
    """Items of a dataset from its info dict (_detail.py:51-85)."""
    info = dict(info)
    info.pop('type', None)
    for option, type_ in (('size', int), ('step', int | float),
                          ('rpix', int | float), ('rval', int | float)):
        if option in info:
            info[option] = _sanitize(option, info[option], ndim, type_)
    step = info.pop('step', None)
    rpix = info.pop('rpix', None)
    rval = info.pop('rval', None)
    rota = info.pop('rota', None)
    opts = {}
    for name in names:
        if name in info:
            opts[name] = data_parser.load(info.pop(name), step, rpix, rval,
                                          rota, prefix)
    if info:
        raise RuntimeError(f"unknown {cls.type()} dataset options: "
                           f"{sorted(info)}")
    return opts


def _dump_common(dataset, prefix='', dump_path=True, overwrite=False):
    """(_detail.py:88-104)"""
    info = dict(type=dataset.type())
    info.update(step=dataset.step(), rpix=dataset.rpix(),
                rval=dataset.rval(), rota=dataset.rota())
    for key, data in dataset.items():
        info[key] = data.dump(f'{prefix}{key}_d.fits', f'{prefix}{key}_m.fits',
                              f'{prefix}{key}_e.fits', dump_wcs=False,
                              dump_path=dump_path, overwrite=overwrite)
    return info


class DatasetImage(Dataset):
    """datasets/image.py:10-28"""

    @staticmethod
    def type():
        return 'image'

    @classmethod
    def load(cls, info, **kwargs):
        return cls(**_load_common(cls, info, ['image'], 2, **kwargs))

    def __init__(self, image):
        super().__init__(dict(image=image))


class DatasetLSlit(Dataset):
    """datasets/lslit.py:10-28"""

    @staticmethod
    def type():
        return 'lslit'

    @classmethod
    def load(cls, info, **kwargs):
        return cls(**_load_common(cls, info, ['lslit'], 2, **kwargs))

    def __init__(self, lslit):
        super().__init__(dict(lslit=lslit))


class DatasetMMaps(Dataset):
    """datasets/mmaps.py:11-38: items mmap0 ... mmap7."""

    @staticmethod
    def type():
        return 'mmaps'

    @classmethod
    def load(cls, info, **kwargs):
        names = [f'mmap{i}' for i in range(8)]
        return cls(**_load_common(cls, info, names, 2, **kwargs))

    def __init__(self, mmap0=None, mmap1=None, mmap2=None, mmap3=None,
                 mmap4=None, mmap5=None, mmap6=None, mmap7=None):
        maps = dict(mmap0=mmap0, mmap1=mmap1, mmap2=mmap2, mmap3=mmap3,
                    mmap4=mmap4, mmap5=mmap5, mmap6=mmap6, mmap7=mmap7)
        super().__init__({k: v for k, v in maps.items() if v is not None})


class DatasetSCube(Dataset):
    """datasets/scube.py:10-28"""

    @staticmethod
    def type():
        return 'scube'

    @classmethod
    def load(cls, info, **kwargs):
        return cls(**_load_common(cls, info, ['scube'], 3, **kwargs))

    def __init__(self, scube):
        super().__init__(dict(scube=scube))


class _DatasetParser:
    """dataset_parser of core.py:100-113: load(info) by info['type'];
    dump(obj, prefix=...) with unique prefixes for several datasets."""

    def __init__(self, classes):
        self._classes = {c.type(): c for c in classes}

    def register(self, cls):
        if cls.type() in self._classes:
            raise RuntimeError(f"dataset type '{cls.type()}' is already "
                               f"registered")
        self._classes[cls.type()] = cls

    def load(self, info, **kwargs):
        if 'type' not in info:
            raise RuntimeError("dataset info has no 'type'")
        cls = self._classes.get(info['type'])
        if cls is None:
            raise RuntimeError(f"unknown dataset type '{info['type']}'; "
                               f"known: {sorted(self._classes)}")
        return cls.load(info, **kwargs)

    def load_many(self, infos, **kwargs):
        return [self.load(i, **kwargs) for i in infos]

    def dump(self, obj, **kwargs):
        return obj.dump(**kwargs)

    def dump_many(self, objs, prefix=None, **kwargs):
        prefix = list(prefix) if isinstance(prefix, (list, tuple)) else \
            [prefix]
        if len(objs) != len(set(prefix)):
            raise RuntimeError(
                "when dumping multiple datasets, a unique prefix for each "
                "dataset must be provided")
        return [o.dump(prefix=p, **kwargs) for o, p in zip(objs, prefix)]


dataset_parser = _DatasetParser([DatasetImage, DatasetLSlit, DatasetMMaps,
                                 DatasetSCube])


def dmodel_options(dataset, ndim, **info):
    """Grid options of a data model described by a dataset
    (dmodels/_detail.py:59-75): size from the dataset, step / rpix / rval /
    rota / dtype from `info` if given there, else from the dataset;
    dimensional options trimmed or repeated to `ndim`."""
    out = dict(info)
    out.update(size=dataset.size(),
               step=info.get('step', dataset.step()),
               rpix=info.get('rpix', dataset.rpix()),
               rval=info.get('rval', dataset.rval()),
               rota=info.get('rota', dataset.rota()),
               dtype=info.get('dtype', np.dtype(dataset.dtype).name))
    for option, type_ in (('size', int), ('step', int | float),
                          ('rpix', int | float), ('rval', int | float),
                          ('scale', int)):
        if option in out:
            out[option] = tuple(_sanitize(option, out[option], ndim,
                                          type_))
    return out


def check_compatible(dataset, dmodel, index=0):
    """The reference Objective's checks of one (dataset, dmodel) pair
    (objective.py:79-110): same keys, dtype, size, step and zero."""
    keys_dat, keys_mdl = tuple(dataset.keys()), tuple(dmodel.keys())
    if set(keys_dat) != set(keys_mdl):
        raise RuntimeError(f"dataset and dmodel are incompatible for item "
                           f"#{index} ({keys_dat} != {keys_mdl})")
    if np.dtype(dataset.dtype) != np.dtype(dmodel.dtype()):
        raise RuntimeError(f"dataset and dmodel have incompatible dtypes "
                           f"for item #{index} ({dataset.dtype} != "
                           f"{dmodel.dtype()})")
    for name in ('size', 'step', 'zero'):
        a = tuple(getattr(dataset, name)())
        b = tuple(getattr(dmodel, name)())
        if a != b:
            raise RuntimeError(f"dataset and dmodel have incompatible "
                               f"{name}s for item #{index} ({a} != {b})")
