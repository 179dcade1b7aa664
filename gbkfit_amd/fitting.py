"""Vectorised log-probability for ensemble samplers.

The reference's emcee fitter calls one Python function per walker
(src/gbkfit/fitting/fitters/emcee/sampler.py:34-57,205-210): it clamps
xpos to the FP32 range, evaluates the parameter expressions, adds the
priors' log probability (prior/prior_dict.py:42-43) and the objective's
log likelihood, -0.5 chi^2 (src/gbkfit/objective/objective.py:208-226),
one model evaluation per call.  ``BatchedLogProbability`` does the same for
a whole ensemble in one batched evaluation (BatchEvaluator): theta is an
(nwalkers, ndim) matrix whose columns are the free parameters, by name.

Names follow the reference's exploded form: a scalar parameter by its
name ('incl', 'vpt_vt'), one node of a node-wise parameter as
'name[i]' (its other nodes come from `fixed`).

Priors: a dict name -> (minimum, maximum) of uniform priors
(prior/priors.py:150-162) or a callable theta -> log prior vector.  The
reference's Prior.log_prob returns 0 where the probability is zero
(prior/priors.py:77-79), so its sampler never rejects a walker outside the
bounds; here such a walker gets -inf and is not evaluated
(reject_outside=True), or, with reject_outside=False, log prior 0 and an
evaluation, as in the reference.
"""
import re

import numpy as np

_NODE = re.compile(r'^(.*)\[(\d+)\]$')


def _split(name):
    m = _NODE.match(name)
    return (m.group(1), int(m.group(2))) if m else (name, None)


class BatchedLogProbability:
    """log p(theta) = log L + log prior of every row of theta, with the
    reference's blobs (log L, log prior).

    evaluator: a BatchEvaluator (one galaxy model against its data);
    names:     the columns of theta;
    fixed:     values of every other model parameter (scalars, or arrays of
               one value per ring node for node-wise ones);
    priors:    None, a dict name -> (min, max), or a callable;
    blobs:     return (log p, structured array of (log_like, log_prior)) as
               emcee's blobs_dtype expects, else log p only.
    """

    def __init__(self, evaluator, names, fixed=None, priors=None,
                 reject_outside=True, blobs=False):
        self.ev = evaluator
        self.names = list(names)
        self.fixed = dict(fixed or {})
        self.priors = priors
        self.reject_outside = bool(reject_outside)
        self.blobs = bool(blobs)
        disk = evaluator.packer.disk
        self.pdescs = disk.pdescs()
        self.nnodes = disk.rnodes.size
        self._cols = []
        for j, n in enumerate(self.names):
            base, idx = _split(n)
            if base not in self.pdescs:
                raise KeyError(f"unknown parameter '{base}'")
            size = self.pdescs[base]
            if idx is not None and not 0 <= idx < size:
                raise IndexError(f"'{n}': node index out of range")
            if idx is None and size > 1:
                raise ValueError(f"'{n}' has {size} values: name its nodes "
                                 f"as '{base}[i]'")
            self._cols.append((j, base, idx))
        missing = []
        for base, size in self.pdescs.items():
            free = {i for _, b, i in self._cols if b == base}
            whole = None in free or (size == 1 and free)
            if not whole and len(free) < size and base not in self.fixed:
                missing.append(base)
        if missing:
            raise KeyError(f"no value for {sorted(missing)}")
        if isinstance(priors, dict):
            for n in priors:
                if n not in self.names:
                    raise KeyError(f"prior for '{n}', which is not free")

    # -- host-side pieces (no device work) ------------------------------------
    def log_prior(self, theta):
        """log prior of every row (uniform priors: -log(max - min) inside;
        outside -inf, or 0 like the reference when reject_outside is
        False)."""
        theta = np.asarray(theta, np.float64)
        n = theta.shape[0]
        if self.priors is None:
            return np.zeros(n)
        if callable(self.priors):
            return np.asarray(self.priors(theta), np.float64).reshape(n)
        out = np.zeros(n)
        for name, (lo, hi) in self.priors.items():
            x = theta[:, self.names.index(name)]
            inside = (x >= lo) & (x <= hi)
            out += np.where(inside, -np.log(hi - lo),
                            -np.inf if self.reject_outside else 0.0)
        return out

    def arrays(self, theta):
        """Parameter name -> per-row values ((n,) or (n, nnodes)) for the
        set packer: the free columns over the fixed values, xpos clamped to
        the FP32 range and rounded to FP32 (sampler.py:36-46)."""
        theta = np.asarray(theta, np.float64)
        n = theta.shape[0]
        out = {}
        for base, size in self.pdescs.items():
            if base in self.fixed:
                v = np.asarray(self.fixed[base], np.float64)
                out[base] = np.broadcast_to(
                    v.reshape(1, -1) if size > 1 else v.reshape(1),
                    (n, size) if size > 1 else (n,)).copy()
            else:
                out[base] = np.zeros((n, size) if size > 1 else n)
        for j, base, idx in self._cols:
            if idx is None:
                if out[base].ndim == 2:
                    out[base][:] = theta[:, j:j + 1]
                else:
                    out[base] = theta[:, j].copy()
            else:
                out[base][:, idx] = theta[:, j]
        if 'xpos' in out:
            f = np.finfo(np.float32)
            x = np.where(out['xpos'] >= f.max, f.max - 1, out['xpos'])
            x = np.where(x <= f.min, f.min + 1, x)
            out['xpos'] = x.astype(np.float32).astype(np.float64)
        return out

    def _dicts(self, arrays, rows):
        return [{k: (v[i] if v.ndim == 2 else float(v[i]))
                 for k, v in arrays.items()} for i in rows]

    # -- evaluation -----------------------------------------------------------
    def log_likelihood(self, theta):
        """-0.5 chi^2 of every row, in one batched evaluation."""
        arrays = self.arrays(theta)
        n = np.asarray(theta).shape[0]
        if n == 0:
            return np.zeros(0)
        ev = self.ev
        if ev.mc:                  # per-set cloud budgets from the dicts
            chi2 = ev.evaluate(self._dicts(arrays, range(n)))
        else:
            ev.upload_packed(ev.packer.pack_arrays(arrays, n))
            ev.run(n)
            chi2 = ev._chi2_dev.cpu().numpy().copy()
        return -0.5 * chi2

    def __call__(self, theta):
        if isinstance(theta, dict):
            theta = [theta]
        if len(theta) and isinstance(theta[0], dict):
            theta = np.array([[d[n] for n in self.names] for d in theta],
                             np.float64)
        theta = np.atleast_2d(np.asarray(theta, np.float64))
        lp = self.log_prior(theta)
        ll = np.full(theta.shape[0], np.nan)
        ok = np.isfinite(lp)
        if ok.any():
            ll[ok] = self.log_likelihood(theta[ok])
        logp = np.where(ok, ll + lp, -np.inf)
        if not self.blobs:
            return logp
        b = np.zeros(theta.shape[0], dtype=[('log_like', float),
                                            ('log_prior', float)])
        # a rejected walker's blobs are NaN, as the reference returns
        b['log_like'] = ll
        b['log_prior'] = np.where(ok, lp, np.nan)
        return logp, b
