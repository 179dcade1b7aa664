"""End-to-end parity: a model evaluated through gbkfit_amd's model layer
(DModelSCube / DModelMMaps + Kinematics gmodels + Objective, the reference's
call sequence) on the GPU against the FP64 oracle pipeline.

Tolerance: |gpu - oracle| <= 1e-5 * max|oracle| + amb, amb being the
oracle's bound on FP32 near-tie decisions propagated through the
convolution and downscale (oracle.evaluate_scube)."""
import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope='module')
def drv():
    return H.driver()


def _cube(drv, p, padding='reference'):
    model, params = P.build_product(drv, p, padding=padding)
    return model.model_h(params)['scube']['d'].astype(np.float64)


def test_c1_cube(drv):
    p = P.c1_params()
    got = _cube(drv, p)
    ref, amb, _ = P.oracle_cube(p)
    err = P.excess_error(got, ref, amb)
    print('C1 excess error', err, 'normwise', H.normwise(got, ref))
    assert err < TOL


@pytest.mark.parametrize('seed', [0, 1])
def test_c3_cubes(drv, seed):
    for p in P.c3_params(np.random.default_rng(seed), 2):
        got = _cube(drv, p)
        ref, amb, _ = P.oracle_cube(p)
        err = P.excess_error(got, ref, amb)
        print('C3 excess error', err)
        assert err < TOL


def test_thick_cube(drv):
    p = P.c1_params()
    p.update(size=(32, 32, 32), thin=False, rnodes=np.arange(0, 15, 1.5),
             bpt_s=4.0, rt=3.0, size_z=48)
    got = _cube(drv, p)
    ref, amb, _ = P.oracle_cube(p)
    assert P.excess_error(got, ref, amb) < TOL


def test_smooth235_padding(drv):
    p = P.c1_params()
    got = _cube(drv, p, padding='smooth235')
    ref, amb, _ = P.oracle_cube(p, pad='smooth235')
    assert P.excess_error(got, ref, amb) < TOL


def test_c1_moment_maps(drv):
    """Moment maps 0/1/2 on spaxels with |m0| > 1e-3 max|m0|.

    Error model: an FP32 FFT convolution leaves an absolute error of a few
    FP32 ulps of the cube maximum in every voxel (the reference warns that
    moments of FFT-convolved cubes pick up this noise, mmaps.py:84-93).  With
    a per-voxel floor n = 4e-7 max|cube|, the moments move by at most
      dm1 <= n dv sum_k |v_k - m1| / m0,
      dm2 <= n dv sum_k (v_k - m1)^2 / (2 m0 m2),
    which the test adds to a 1e-5 normwise tolerance.  The mask agrees
    wherever |m0| is not within 1e-3 max|m0| of the 1e-6 cutoff."""
    p = P.c1_params()
    model, params = P.build_product(drv, p, orders=(0, 1, 2))
    out = model.model_h(params)
    ref, amb, _ = P.oracle_cube(p)
    geom = P.oracle_geom(p)
    md, mm, _ = oracle.moments(ref, geom.step_lo, geom.zero_lo, (0, 1, 2),
                               1e-6)
    m0, m1, m2 = md
    sel = np.abs(m0) > 1e-3 * np.nanmax(np.abs(m0))
    dv = geom.step_lo[2]
    v = geom.zero_lo[2] + dv * np.arange(ref.shape[0])
    floor = 4e-7 * np.abs(ref).max()
    dev = v[:, None, None] - m1[None]
    bound = {
        'mmap0': floor * dv * ref.shape[0] + 0 * m0,
        'mmap1': floor * dv * np.abs(dev).sum(0) / np.abs(m0),
        'mmap2': floor * dv * (dev ** 2).sum(0) / (2 * np.abs(m0) * m2),
    }
    for i, key in enumerate(('mmap0', 'mmap1', 'mmap2')):
        g = out[key]['d'].astype(np.float64)
        r = md[i]
        assert np.all(np.isfinite(g[sel])), key
        scale = np.nanmax(np.abs(r[sel]))
        excess = np.abs(g[sel] - r[sel]) - bound[key][sel]
        print(key, np.max(np.abs(g[sel] - r[sel])) / scale,
              np.max(excess) / scale)
        assert np.max(excess) / scale < 1e-5, key
    gm = out['mmap0']['m']
    clear = np.abs(np.abs(m0) - 1e-6) > 1e-3 * np.nanmax(np.abs(m0))
    assert np.array_equal(gm[clear & sel], mm[clear & sel])


def test_objective_chi2(drv):
    """log_likelihood through Objective vs the oracle's chi^2 of the same
    model against a seeded mock (sigma 0.01)."""
    from gbkfit_amd.objective import Data, Objective
    p = P.c1_params()
    ref, amb, _ = P.oracle_cube(p)
    rng = np.random.default_rng(42)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    err = np.full(ref.shape, 0.01, np.float32)
    model, params = P.build_product(drv, p)
    obj = Objective([dict(scube=Data(mock, error=err))], [model])
    ll = obj.log_likelihood(params)[0]
    r = oracle.residual(mock, err, None, ref, None, None, 1.0)
    chi2 = oracle.residual_sum(r, True)
    print('chi2', -2 * ll, chi2)
    assert abs(-2 * ll - chi2) / chi2 < 1e-4
