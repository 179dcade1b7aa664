"""Batched moment-map objective: chi^2 of every set against data maps 0/1/2
(DModelMMaps + Objective, mmaps.py:136-201, objective.py:279-308) in one
batched evaluation, against the plain path (Objective.log_likelihood per
set) and the FP64 oracle's moments.

Tolerance: moments of the FP32 cube differ from the oracle's by the FP32
error model of test_gpu_pipeline.test_c1_moment_maps; on the spaxels kept
here (|m0| > 1e-2 max|m0|) a residual moves by < 1e-5 of its error, so chi^2
agrees within 1e-4 relative.  Masked spaxels (NaN data, mask 0) add
nothing."""
import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu


def _maps_dataset(p, seed):
    from gbkfit_amd.dataset import Data, DatasetMMaps
    ref, _, _ = P.oracle_cube(p)
    geom = P.oracle_geom(p)
    md, _, _ = oracle.moments(ref, geom.step_lo, geom.zero_lo, (0, 1, 2),
                              1e-6)
    sel = np.abs(md[0]) > 1e-2 * np.nanmax(np.abs(md[0]))
    rng = np.random.default_rng(seed)
    errs = (1e-2 * np.nanmax(np.abs(md[0])), 5.0, 5.0)
    items, data = {}, []
    for i, e in enumerate(errs):
        d = md[i] + rng.normal(0, e, md[i].shape)
        d[~sel] = np.nan
        d = d.astype(np.float32)
        data.append(d)
        items[f'mmap{i}'] = Data(d, mask=sel.astype(np.float32),
                                 error=np.full(d.shape, e, np.float32))
    return DatasetMMaps(**items), md, data, errs, sel


def _oracle_chi2(md, data, errs, sel):
    tot = 0.0
    for i in range(3):
        r = oracle.residual(data[i], np.full(data[i].shape, errs[i]),
                            sel.astype(np.float64), md[i], None, None, 1.0)
        tot += oracle.residual_sum(r, True)
    return tot


@pytest.mark.parametrize('thin', [True, False], ids=['thin', 'thick'])
def test_batched_mmaps_chi2(thin):
    from gbkfit_amd.batch import BatchEvaluator
    from gbkfit_amd.objective import Objective
    drv = H.driver()
    p = dict(P.c1_params(), thin=thin)
    if not thin:
        p.update(size=(48, 48, 48), rnodes=np.arange(0, 22, 1.5), size_z=32)
    ds, md, data, errs, sel = _maps_dataset(p, 3)
    model, params = P.build_product(drv, p, orders=(0, 1, 2))
    obj = Objective([ds], [model])
    sets = [dict(p), dict(p, incl=60.0, posa=40.0, vt=180.0),
            dict(p, disp=35.0, rt=4.0)]
    plain = [-2 * obj.log_likelihood(P.gparams(s))[0] for s in sets]
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), ds,
                        max_batch=2, template_params=params)
    got, cubes = ev.evaluate([P.gparams(s) for s in sets], want_cubes=True)
    got_nograph = ev.evaluate([P.gparams(s) for s in sets], use_graph=False)
    info = ev.info()
    ev.close()
    want = _oracle_chi2(md, data, errs, sel)
    print(info['pipeline'], 'batched', got, 'plain', plain, 'oracle', want)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, got_nograph)
    assert abs(got[0] - want) / want < 1e-4
    for g, q in zip(got, plain):
        assert abs(g - q) / q < 1e-4
    # the model cubes come back too (the moment objective reads them)
    ref, amb, _ = P.oracle_cube(p)
    assert P.excess_error(cubes[0], ref, amb) < 1e-5
