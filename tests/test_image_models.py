"""Host logic of the image and long-slit data models and of the intensity
gmodels (no device work): geometry against the oracle's DCube geometry,
parameter names of components and opacity components against the
reference's naming rule (miscutils.py:70-95, gmodels/_detail.py:339-342),
and the constructors' trait checks."""
import numpy as np
import pytest

import oracle
from gbkfit_amd import gmodels, psflsf
from gbkfit_amd.dmodels import DModelImage, DModelLSlit
from gbkfit_amd.driver.backend import DriverBackendFFT
from gbkfit_amd.traits import make

RN = np.arange(0, 10, 1.0)


class _Shapes(DriverBackendFFT):
    """The FFT backend's shape rules without a device."""

    def __init__(self):
        super().__init__(32)

    def fft_r2c(self, *a):
        raise NotImplementedError

    def fft_c2r(self, *a):
        raise NotImplementedError

    def fft_convolve(self, *a):
        raise NotImplementedError

    def fft_convolve_cached(self, *a):
        raise NotImplementedError


def _geom(dcube):
    return dcube.geometry(_Shapes())


def test_image_dcube_is_one_channel_with_psf_only():
    """DModelImage((48, 40)): a (48, 40, 1) cube with spectral step 0,
    the PSF padding on x and y only, the spectral axis one channel at zero;
    the same geometry as the oracle's DCube of that size."""
    psf = psflsf.PSFGauss(1.2)
    dm = DModelImage((48, 40), (0.5, 0.5), scale=(2, 2), psf=psf)
    size_hi, step_hi, edge_hi, zero_hi = _geom(dm.dcube())
    g = oracle.DCubeGeom((48, 40, 1), (0.5, 0.5, 0), scale=(2, 2, 1),
                         psf_sigma=1.2)
    assert tuple(size_hi) == tuple(g.size_hi) and size_hi[2] == 1
    assert tuple(edge_hi) == tuple(g.edge_hi) and edge_hi[2] == 0
    assert np.allclose(zero_hi, g.zero_hi) and zero_hi[2] == 0
    assert np.allclose(step_hi, g.step_hi)
    assert dm.size() == (48, 40) and dm.step() == (0.5, 0.5)
    assert np.allclose(dm.zero(), (-11.75, -9.75))
    assert dm.keys() == ['image'] and dm.scale() == (2, 2)


def test_lslit_dcube_is_one_row_high():
    """DModelLSlit((30, 24)): positions x channels -> a (30, 1, 24) cube,
    default reference pixel at the centre of the slit and the spectrum;
    the geometry equals the oracle's DCube of that size."""
    dm = DModelLSlit((30, 24), (1, 2, 10), scale=(2, 1, 1),
                     psf=psflsf.PSFGauss(1.0), lsf=psflsf.LSFGauss(8.0))
    assert dm.size() == (30, 1, 24) and dm.keys() == ['lslit']
    g = oracle.DCubeGeom((30, 1, 24), (1, 2, 10), scale=(2, 1, 1),
                         psf_sigma=1.0, lsf_sigma=8.0)
    size_hi, step_hi, edge_hi, zero_hi = _geom(dm.dcube())
    assert tuple(size_hi) == tuple(g.size_hi)
    assert tuple(edge_hi) == tuple(g.edge_hi)
    assert np.allclose(zero_hi, g.zero_hi) and np.allclose(dm.zero(),
                                                           g.zero_lo)
    assert np.allclose(dm.zero(), (-14.5, 0.0, -115.0))
    with pytest.raises(RuntimeError):
        DModelLSlit((30, 1, 24))


def _bright2d(**kw):
    return gmodels.BrightnessSMDisk2D(False, False,
                                      make('bpt', 'exponential'),
                                      rnodes=RN, **kw)


def _bright3d():
    return gmodels.BrightnessSMDisk3D(False, False,
                                      make('bpt', 'exponential'),
                                      make('bht', 'sech2'), rnodes=RN)


def _opacity():
    return gmodels.OpacitySMDisk3D(False, False, make('opt', 'exponential'),
                                   make('oht', 'sech2'), rnodes=RN)


def test_brightness_components_have_no_kinematic_parameters():
    p2 = _bright2d().pdescs()
    assert set(p2) == {'xpos', 'ypos', 'posa', 'incl', 'bpt_a', 'bpt_s'}
    p3 = _bright3d().pdescs()
    assert set(p3) == {'xpos', 'ypos', 'posa', 'incl', 'bpt_a', 'bpt_s',
                       'bht_s'}
    po = _opacity().pdescs()
    assert {'opt_a', 'opt_s'} <= set(po) and 'vsys' not in po
    assert any(k.startswith('oht_') for k in po)


def test_parameter_names_of_components_and_opacity_components():
    """Components: the first unprefixed, then 'cmp1_'; opacity components
    always prefixed: 'ocmp_' for the first, then 'ocmp1_'."""
    gm = gmodels.GModelIntensity3D([_bright3d(), _bright3d()],
                                   [_opacity(), _opacity()])
    names = set(gm.pdescs())
    for n in ('incl', 'bpt_s', 'cmp1_incl', 'cmp1_bpt_s', 'ocmp_incl',
              'ocmp_opt_a', 'ocmp1_incl', 'ocmp1_opt_a'):
        assert n in names, n
    assert 'ocmp0_incl' not in names and 'cmp0_incl' not in names
    assert len(names) == 2 * len(_bright3d().pdescs()) + \
        2 * len(_opacity().pdescs())


def test_intensity3d_spatial_grid_defaults():
    """The z axis takes the larger spatial axis' size and step, zero in the
    middle (intensity_3d.py:95-107), unless given."""
    gm = gmodels.GModelIntensity3D(_bright3d())
    size, step, zero = gm.spatial_grid((40, 64), (0.5, 0.25), (-10, -8))
    assert size == (40, 64, 64) and step[2] == 0.25
    assert zero[2] == pytest.approx(-(32 - 0.5) * 0.25)
    gm = gmodels.GModelIntensity3D(_bright3d(), size_z=10, step_z=2.0,
                                   zero_z=-3.0)
    size, step, zero = gm.spatial_grid((40, 64), (0.5, 0.25), (-10, -8))
    assert size[2] == 10 and step[2] == 2.0 and zero[2] == -3.0


def test_component_trait_checks():
    with pytest.raises(RuntimeError):
        gmodels.BrightnessSMDisk3D(False, False, make('bpt', 'exponential'),
                                   None, rnodes=RN)
    with pytest.raises(RuntimeError):
        gmodels.BrightnessSMDisk3D(
            False, False, [make('bpt', 'exponential')] * 2,
            make('bht', 'sech2'), rnodes=RN)
    with pytest.raises(RuntimeError):
        gmodels.BrightnessSMDisk2D(False, False, None, rnodes=RN)
    with pytest.raises(RuntimeError):
        gmodels.GModelIntensity2D([])
    assert _bright2d(wptraits=make('wpt', 'axis_range', axis=0, angle=10,
                                   weight=2)).has_weights()


def test_batch_evaluator_rejects_opacity_and_image_gmodels():
    from gbkfit_amd.batch import BatchEvaluator
    spec = gmodels.SpectralSMDisk3D(
        False, False, make('bpt', 'exponential'), make('vpt', 'tan_arctan'),
        make('dpt', 'uniform'), make('bht', 'sech2'), rnodes=RN)
    for gm in (gmodels.GModelKinematics3D(spec, _opacity()),
               gmodels.GModelIntensity2D(_bright2d())):
        with pytest.raises(RuntimeError):
            BatchEvaluator(None, None, gm, None)
