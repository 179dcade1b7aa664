"""Model and objective: model_d / model_h and the residual / likelihood of
one or more (data, model) pairs.

Mirror of src/gbkfit/model/model.py:21-127 (Model, ModelGroup) and
src/gbkfit/objective/objective.py:133-308 (device data/mask/error vectors,
residual per data key, residual_sum, log_likelihood = -0.5 * sum chi^2).
"""
import numpy as np


class Model:
    """A driver, a data model and a galaxy model (model.py:21-86)."""

    def __init__(self, driver, dmodel, gmodel):
        self._driver, self._dmodel, self._gmodel = driver, dmodel, gmodel

    def driver(self):
        return self._driver

    def dmodel(self):
        return self._dmodel

    def gmodel(self):
        return self._gmodel

    def pdescs(self):
        return self._gmodel.pdescs()

    def model_d(self, params, out_extra=None):
        return self._dmodel.evaluate(self._driver, self._gmodel, params,
                                     out_extra, out_extra)

    def model_h(self, params, out_extra=None):
        d = self.model_d(params, out_extra)
        return {k: {n: (None if v is None else
                        self._driver.mem_copy_d2h(v))
                    for n, v in item.items()} for k, item in d.items()}


class Data:
    """One observed array with optional mask and error (the arrays of
    src/gbkfit/dataset/data.py that the objective reads)."""

    def __init__(self, data, mask=None, error=None):
        self._d = np.asarray(data)
        self._m = None if mask is None else np.asarray(mask)
        self._e = None if error is None else np.asarray(error)

    def data(self):
        return self._d

    def mask(self):
        return self._m

    def error(self):
        return self._e


class Objective:
    """Residuals and likelihood of `models` against `datasets` (one dict of
    Data per model, keyed like the data model's keys(), or one
    gbkfit_amd.dataset.Dataset per model, checked against its data model as
    the reference does: keys, dtype, size, step and zero,
    objective.py:79-110)."""

    def __init__(self, datasets, models, weights=None):
        from .dataset import Dataset, check_compatible
        self._datasets = list(datasets)
        self._models = list(models)
        if len(self._datasets) != len(self._models):
            raise RuntimeError(
                f"the number of datasets and models are not equal "
                f"({len(self._datasets)} != {len(self._models)})")
        for i, (ds, model) in enumerate(zip(self._datasets, self._models)):
            if isinstance(ds, Dataset):
                check_compatible(ds, model.dmodel(), i)
        self._weights = weights or [dict() for _ in self._models]
        self._prepared = False

    def prepare(self):
        self._dev = []
        for ds, model in zip(self._datasets, self._models):
            drv, dm = model.driver(), model.dmodel()
            keys = dm.keys()
            shape = tuple(dm.size())[::-1]
            dtype = dm.dtype()
            npix = int(np.prod(shape))
            n = npix * len(keys)
            st = dict(d=drv.mem_alloc_d(n, dtype), m={}, e={},
                      res=drv.mem_alloc_d(n, dtype),
                      sum=drv.mem_alloc_d(1, dtype), views={}, keys=keys,
                      backend=drv.backends().objective(dtype))
            drv.mem_fill(st['d'], 0)
            for j, key in enumerate(keys):
                sl = slice(j * npix, (j + 1) * npix)
                item = ds[key]
                drv.mem_copy_h2d(item.data().ravel().astype(dtype),
                                 st['d'][sl])
                if item.mask() is not None:
                    st['m'][key] = drv.mem_copy_h2d(
                        item.mask().ravel().astype(dtype)).reshape(shape)
                if item.error() is not None:
                    st['e'][key] = drv.mem_copy_h2d(
                        item.error().ravel().astype(dtype)).reshape(shape)
                st['views'][key] = (st['d'][sl].reshape(shape),
                                    st['res'][sl].reshape(shape))
            self._dev.append(st)
        self._prepared = True

    def _update_residual_d(self, params, weighted, out_extra=None):
        if not self._prepared:
            self.prepare()
        for i, (model, st) in enumerate(zip(self._models, self._dev)):
            mdata = model.model_d(params, out_extra)
            for key in st['keys']:
                obs, res = st['views'][key]
                w = self._weights[i].get(key, 1.0) if weighted else 1.0
                st['backend'].residual(
                    obs, st['e'].get(key), st['m'].get(key), mdata[key]['d'],
                    mdata[key]['w'], mdata[key]['m'], w, res)

    def residual_vector_d(self, params, weighted=True, out_extra=None):
        self._update_residual_d(params, weighted, out_extra)
        return [st['res'] for st in self._dev]

    def residual_scalar(self, params, squared=True, out_extra=None):
        self._update_residual_d(params, True, out_extra)
        out = []
        for model, st in zip(self._models, self._dev):
            st['backend'].residual_sum(st['res'], squared, st['sum'])
            out.append(float(model.driver().mem_copy_d2h(st['sum'])[0]))
        return out

    def log_likelihood(self, params, out_extra=None):
        return [-0.5 * v for v in self.residual_scalar(params, True,
                                                        out_extra)]
