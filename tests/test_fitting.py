"""Host logic of the vectorised log-probability (gbkfit_amd.fitting): the
columns of theta over the fixed values, node-wise columns, the xpos clamp
of the reference's log_probability (emcee/sampler.py:36-46), uniform
priors in both modes, and the assembly of log p and blobs; the batched
evaluation itself is replaced by a stand-in (the GPU test compares it with
per-walker Objective.log_likelihood)."""
import numpy as np
import pytest

from gbkfit_amd import gmodels
from gbkfit_amd.fitting import BatchedLogProbability
from gbkfit_amd.traits import make


class _Packer:
    def __init__(self, disk):
        self.disk = disk


class _StandIn:
    """Evaluator stand-in: records the arrays it would pack; chi^2 = sum
    of the free columns squared."""
    mc = False

    def __init__(self, disk):
        self.packer = _Packer(disk)


def _disk():
    cmp = gmodels.SpectralSMDisk2D(
        False, False, make('bpt', 'exponential'),
        make('vpt', 'nw_tan_uniform'), make('dpt', 'uniform'),
        rnodes=np.arange(0, 6, 1.0))
    return cmp.disk


FIXED = dict(xpos=0.0, ypos=0.0, vsys=0.0, bpt_a=1.0, bpt_s=3.0,
             vpt_vt=np.linspace(0, 200, 6), dpt_a=20.0)


def test_arrays_free_over_fixed_and_node_columns():
    lp = BatchedLogProbability(_StandIn(_disk()),
                               ['incl', 'posa', 'vpt_vt[2]', 'xpos'],
                               fixed=FIXED)
    theta = np.array([[30.0, 10.0, 55.0, 0.5], [60.0, 200.0, 77.0, -1.0]])
    a = lp.arrays(theta)
    assert np.array_equal(a['incl'], [30.0, 60.0])
    assert np.array_equal(a['posa'], [10.0, 200.0])
    assert a['vpt_vt'].shape == (2, 6)
    assert np.array_equal(a['vpt_vt'][:, 2], [55.0, 77.0])
    assert np.array_equal(a['vpt_vt'][0, [0, 1, 3, 4, 5]],
                          FIXED['vpt_vt'][[0, 1, 3, 4, 5]])
    assert np.array_equal(a['xpos'], [0.5, -1.0])
    assert np.array_equal(a['bpt_s'], [3.0, 3.0])


def test_xpos_clamped_to_the_fp32_range():
    lp = BatchedLogProbability(_StandIn(_disk()), ['xpos'],
                               fixed=dict(FIXED, incl=45.0, posa=0.0))
    f = np.finfo(np.float32)
    a = lp.arrays(np.array([[1e300], [-1e300], [0.1]]))
    assert a['xpos'][0] == np.float32(f.max - 1)
    assert a['xpos'][1] == np.float32(f.min + 1)
    assert a['xpos'][2] == float(np.float32(0.1))


def test_name_errors():
    ev = _StandIn(_disk())
    with pytest.raises(KeyError):
        BatchedLogProbability(ev, ['nope'], fixed=FIXED)
    with pytest.raises(ValueError):       # node-wise: name the nodes
        BatchedLogProbability(ev, ['vpt_vt'], fixed=FIXED)
    with pytest.raises(IndexError):
        BatchedLogProbability(ev, ['vpt_vt[6]'], fixed=FIXED)
    with pytest.raises(KeyError):         # incl has no value
        BatchedLogProbability(ev, ['posa'], fixed=FIXED)


def test_priors_and_assembly(monkeypatch):
    lp = BatchedLogProbability(_StandIn(_disk()), ['incl', 'posa'],
                               fixed=FIXED, priors=dict(incl=(20, 80)),
                               blobs=True)
    seen = []

    def fake_ll(theta):
        seen.append(np.array(theta))
        return -0.5 * np.sum(np.asarray(theta) ** 2, axis=1)
    monkeypatch.setattr(lp, 'log_likelihood', fake_ll)
    theta = np.array([[30.0, 1.0], [90.0, 2.0], [50.0, 3.0]])
    logp, b = lp(theta)
    # the walker outside the prior is not evaluated
    assert np.array_equal(seen[0], theta[[0, 2]])
    lpr = -np.log(60.0)
    assert logp[1] == -np.inf and np.isnan(b['log_like'][1]) and \
        np.isnan(b['log_prior'][1])
    assert logp[0] == pytest.approx(-0.5 * (900 + 1) + lpr)
    assert b['log_prior'][2] == pytest.approx(lpr)
    # the reference's behaviour: log prior 0 outside, evaluated
    ref = BatchedLogProbability(_StandIn(_disk()), ['incl', 'posa'],
                                fixed=FIXED, priors=dict(incl=(20, 80)),
                                reject_outside=False)
    monkeypatch.setattr(ref, 'log_likelihood', fake_ll)
    out = ref(theta)
    assert out[1] == pytest.approx(-0.5 * (8100 + 4))
    # emcee's named form: a list of dicts
    d = lp([dict(incl=30.0, posa=1.0)])[0]
    assert d[0] == pytest.approx(logp[0])
