"""Batched evaluation of gmodels with several components, summed into one
model cube as the reference sums them (gmodels/_detail.py:405-422): two
thin smdisks (Kinematics2D), two thick smdisks and an mcdisk with a thick
smdisk (Kinematics3D), on the direct and padded-cube pipelines.  Each set's
cube and chi^2 against the plain path (Objective per set, which evaluates
the components one after the other into the cube), and the two-thin-disk
cube against the FP64 oracle's sum of the two disks.

Tolerance: cubes |gpu - ref| <= 1e-5 max|ref| (+ amb against the oracle),
chi^2 within 1e-5 relative (mcdisk: the same Philox clouds on both paths,
1e-5 as well)."""
import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu


def _model(drv, kind, size=(48, 48, 40)):
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    psf = psflsf.PSFGauss(2.5 * P.SIG_FWHM)
    lsf = psflsf.LSFGauss(20.0 * P.SIG_FWHM)
    rn1 = np.arange(0, 21, 1.5)
    rn2 = np.arange(0, 10.6, 1.5)
    if kind == 'thin':
        c1 = gmodels.SpectralSMDisk2D(
            False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'), rnodes=rn1)
        c2 = gmodels.SpectralSMDisk2D(
            False, False, make('bpt', 'gauss'), make('vpt', 'tan_arctan'),
            make('dpt', 'uniform'), rnodes=rn2)
        gm = gmodels.GModelKinematics2D([c1, c2])
    else:
        first = (gmodels.SpectralMCDisk3D(
            40.0, False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
            make('bht', 'sech2'), rnodes=rn1) if kind == 'mc'
            else gmodels.SpectralSMDisk3D(
                False, False, make('bpt', 'exponential'),
                make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
                make('bht', 'sech2'), rnodes=rn1))
        c2 = gmodels.SpectralSMDisk3D(
            False, False, make('bpt', 'gauss'), make('vpt', 'tan_arctan'),
            make('dpt', 'uniform'), make('bht', 'sech2'), rnodes=rn2)
        gm = gmodels.GModelKinematics3D([first, c2], size_z=32)
    dm = DModelSCube(size, (1, 1, 10), scale=(2, 2, 1), psf=psf, lsf=lsf)
    return Model(drv, dm, gm)


def _params(model, rng, thick):
    p = dict(xpos=0.0, ypos=0.0, vsys=0.0, posa=float(rng.uniform(0, 360)),
             incl=float(rng.uniform(30, 70)), bpt_a=1.0, bpt_s=5.0,
             vpt_rt=4.0, vpt_vt=float(rng.uniform(120, 250)),
             dpt_a=float(rng.uniform(10, 40)),
             cmp1_xpos=1.5, cmp1_ypos=-1.0, cmp1_vsys=20.0,
             cmp1_posa=float(rng.uniform(0, 360)),
             cmp1_incl=float(rng.uniform(30, 70)), cmp1_bpt_a=2.0,
             cmp1_bpt_s=3.0, cmp1_vpt_rt=2.0,
             cmp1_vpt_vt=float(rng.uniform(80, 150)), cmp1_dpt_a=15.0)
    if thick:
        p.update(bht_s=2.0, cmp1_bht_s=1.0)
    names = set(model.gmodel().pdescs())
    assert names <= set(p), names - set(p)
    return {k: p[k] for k in names}


@pytest.mark.parametrize('kind,path', [('thin', 'direct'), ('thin', 'padded'),
                                       ('thick', 'direct'), ('mc', 'auto')])
def test_multicomponent_batch_matches_plain(kind, path):
    from gbkfit_amd.batch import BatchEvaluator
    from gbkfit_amd.objective import Data, Objective
    drv = H.driver()
    model = _model(drv, kind)
    rng = np.random.default_rng(3)
    sets = [_params(model, rng, kind != 'thin') for _ in range(3)]
    cube0 = model.model_h(sets[0])['scube']['d'].astype(np.float32)
    mock = (cube0 + rng.normal(0, 0.01, cube0.shape)).astype(np.float32)
    obj = Objective([dict(scube=Data(mock, error=np.full(mock.shape, 0.01,
                                                         np.float32)))],
                    [model])
    plain_chi2 = [-2 * obj.log_likelihood(s)[0] for s in sets]
    plain_cubes = [model.model_h(s)['scube']['d'].astype(np.float64)
                   for s in sets]
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=2, template_params=sets[0],
                        path=path)
    chi2, cubes = ev.evaluate(sets, want_cubes=True)
    info = ev.info()
    ev.close()
    print(kind, info['pipeline'], chi2, plain_chi2)
    for i in range(len(sets)):
        ref = plain_cubes[i]
        err = np.abs(cubes[i] - ref).max() / np.abs(ref).max()
        assert err < 1e-5, (i, err)
        assert abs(chi2[i] - plain_chi2[i]) / plain_chi2[i] < 1e-5, i


def test_two_thin_disks_vs_oracle():
    """The batched cube of two thin disks against the FP64 oracle's cube of
    the sum of the two disks (each disk by the oracle's smdisk)."""
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    model = _model(drv, 'thin')
    rng = np.random.default_rng(8)
    s = _params(model, rng, False)
    T = oracle.Trait
    specs = [oracle.DiskSpec(np.arange(0, 21, 1.5),
                             dict(rp=[T(2, params=(1.0, 5.0))],
                                  vp=[T(2, params=(4.0, s['vpt_vt']))],
                                  dp=[T(1, params=(s['dpt_a'],))]),
                             posa=s['posa'], incl=s['incl']),
             oracle.DiskSpec(np.arange(0, 10.6, 1.5),
                             dict(rp=[T(3, params=(2.0, 3.0))],
                                  vp=[T(2, params=(2.0, s['cmp1_vpt_vt']))],
                                  dp=[T(1, params=(15.0,))]),
                             posa=s['cmp1_posa'], incl=s['cmp1_incl'],
                             xpos=1.5, ypos=-1.0, vsys=20.0)]
    p = dict(P.c1_params(), size=(48, 48, 40))
    geom = P.oracle_geom(p)
    grid = oracle.grid_f32(oracle.grid_for(geom, True, None))
    hi = amb = 0
    for sp in specs:
        out, _ = oracle.smdisk(sp.rounded_f32(), grid, want=('scube', 'amb'))
        hi = hi + out['scube']
        amb = amb + out['amb']
    ref, _, amb_lo = oracle.evaluate_scube(geom, hi, amb)
    mock = ref.astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=1, template_params=s)
    _, cubes = ev.evaluate([s], want_cubes=True)
    ev.close()
    err = P.excess_error(cubes[0], ref, amb_lo)
    print('two thin disks vs oracle', err)
    assert err < 1e-5
