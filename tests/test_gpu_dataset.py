"""Dataset ingest feeding the hot path: a mock cube written as FITS files
with WCS cards, read back through dataset_parser, described to the data
model by the dataset (dmodel_options) and fitted through Objective on the
GPU; chi^2 against the FP64 oracle's."""
import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu


def test_fits_dataset_to_objective(tmp_path, monkeypatch):
    from gbkfit_amd import fitsio, gmodels, psflsf
    from gbkfit_amd.dataset import dataset_parser, dmodel_options
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Model, Objective
    from gbkfit_amd.traits import make
    monkeypatch.chdir(tmp_path)
    drv = H.driver()
    p = P.c1_params()
    ref, amb, _ = P.oracle_cube(p)
    rng = np.random.default_rng(11)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    mock[3, 10, 12] = np.nan                     # a bad pixel
    err = np.full(ref.shape, 0.01, np.float32)
    # the cube's grid in the header: centre pixel at world 0, 10 km/s
    nz, ny, nx = ref.shape
    h = fitsio.Header([('CDELT1', 1.0), ('CDELT2', 1.0), ('CDELT3', 10.0),
                       ('CRPIX1', nx / 2 - 0.5), ('CRPIX2', ny / 2 - 0.5),
                       ('CRPIX3', nz / 2 - 0.5), ('CRVAL1', 0.0),
                       ('CRVAL2', 0.0), ('CRVAL3', 0.0)])
    fitsio.writeto('cube_d.fits', mock, h)
    fitsio.writeto('cube_e.fits', err, h)
    ds = dataset_parser.load(dict(
        type='scube', scube=dict(data='cube_d.fits', error='cube_e.fits')))
    assert ds.size() == p['size'] and ds.step() == (1.0, 1.0, 10.0)
    opts = dmodel_options(ds, 3, scale=p['scale'])
    dm = DModelSCube(psf=psflsf.PSFGauss(p['psf_fwhm'] * P.SIG_FWHM),
                     lsf=psflsf.LSFGauss(p['lsf_fwhm'] * P.SIG_FWHM), **opts)
    cmp = gmodels.SpectralSMDisk2D(
        False, False, make('bpt', 'exponential'), make('vpt', 'tan_arctan'),
        make('dpt', 'uniform'), rnodes=p['rnodes'])
    model = Model(drv, dm, gmodels.GModelKinematics2D([cmp]))
    obj = Objective([ds], [model])
    ll = obj.log_likelihood(P.gparams(p))[0]
    msk = ds['scube'].mask()
    assert msk[3, 10, 12] == 0 and msk.sum() == mock.size - 1
    # the masked pixel (NaN data, mask 0) contributes nothing
    r = oracle.residual(mock, err, msk, ref, None, None, 1.0)
    assert r.reshape(mock.shape)[3, 10, 12] == 0
    chi2 = oracle.residual_sum(r, True)
    print('chi2', -2 * ll, chi2)
    assert np.isfinite(ll)
    assert abs(-2 * ll - chi2) / chi2 < 1e-4
    # the batched evaluator on the same dataset arrays
    from gbkfit_amd.batch import BatchEvaluator
    item = ds['scube']
    ev = BatchEvaluator(drv, dm, model.gmodel(), item.data(),
                        error=item.error(), mask=item.mask(), max_batch=2,
                        template_params=P.gparams(p))
    got = ev.evaluate([P.gparams(p), P.gparams(p)])
    ev.close()
    print('batched chi2', got, chi2)
    assert np.all(np.abs(np.asarray(got) - chi2) / chi2 < 1e-4)
