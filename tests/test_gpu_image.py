"""Image and long-slit data models with the intensity gmodels on the GPU,
against the FP64 oracle on the same grids: brightness-only disks into an
image (thin, thick with an opacity disk in front, Monte Carlo clouds),
convolved with the PSF and downscaled (image.py, intensity_2d.py,
intensity_3d.py), and the long-slit plane of a spectral cube one row high
(lslit.py).  Tolerance: 1e-5 of the maximum plus the oracle's ambiguity
bound (FP32 near-ties of the disk edge or of a cloud's voxel)."""
import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu

TOL = 1e-5
RN = np.arange(0, 24, 1.5)
T = oracle.Trait


@pytest.fixture(scope='module')
def drv():
    return H.driver()


def _image_geom(size, step, scale, psf_sigma):
    return oracle.DCubeGeom(tuple(size) + (1,), tuple(step) + (0,),
                            scale=tuple(scale) + (1,), psf_sigma=psf_sigma)


def _oracle_image(geom, out):
    lo, _, amb = oracle.evaluate_scube(geom, out['image'][None], out['amb'])
    return lo[0], amb[0]


def _check(name, got, ref, amb):
    assert np.abs(ref).max() > 0
    nbad, err = H.within(got, ref, TOL, amb)
    print(f'{name}: normwise {err:.2e}, beyond tol + amb: {nbad}')
    assert nbad == 0


def _bright_params(**kw):
    p = dict(xpos=1.5, ypos=-2.0, posa=30.0, incl=50.0, bpt_a=1.0,
             bpt_s=5.0, bpt1_a=0.5, bpt1_s=3.0)
    p.update(kw)
    return p


def test_intensity2d_image_matches_oracle(drv):
    """Two thin brightness traits (exponential + gaussian) into a 48 x 40
    image, 2x spatial supersampling, PSF sigma 1.5 px."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelImage
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    cmp = gmodels.BrightnessSMDisk2D(
        False, False, [make('bpt', 'exponential'), make('bpt', 'gauss')],
        rnodes=RN)
    model = Model(drv, DModelImage((48, 40), (1, 1), scale=(2, 2),
                                   psf=psflsf.PSFGauss(1.5)),
                  gmodels.GModelIntensity2D(cmp))
    p = _bright_params()
    got = model.model_h(p)['image']
    assert got['d'].shape == (40, 48) and got['w'] is None
    geom = _image_geom((48, 40), (1, 1), (2, 2), 1.5)
    spec = oracle.DiskSpec(RN, dict(rp=[T(2, params=(1.0, 5.0)),
                                        T(3, params=(0.5, 3.0))]),
                           xpos=1.5, ypos=-2.0, posa=30.0,
                           incl=50.0).rounded_f32()
    grid = oracle.grid_f32(oracle.grid_for(geom, True))
    out, _ = oracle.smdisk(spec, grid, want=('image', 'amb'))
    _check('intensity2d', got['d'], *_oracle_image(geom, out))


def test_intensity2d_weight_image(drv):
    """A weight trait without PSF or supersampling: the weight image is
    the spaxel's mean / max weight (gmodels.hpp:52-82 with one channel):
    1 on the disk, 0 off it, as the oracle's wcube of its wdata."""
    from gbkfit_amd import gmodels
    from gbkfit_amd.dmodels import DModelImage
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    cmp = gmodels.BrightnessSMDisk2D(
        False, False, make('bpt', 'exponential'),
        wptraits=make('wpt', 'axis_range', axis=0, angle=20, weight=3),
        rnodes=RN)
    model = Model(drv, DModelImage((48, 40)), gmodels.GModelIntensity2D(cmp))
    p = dict(xpos=0.0, ypos=0.0, posa=30.0, incl=50.0, bpt_a=1.0, bpt_s=5.0)
    for _ in range(2):               # the weights do not accumulate
        got = model.model_h(p)['image']['w']
        geom = _image_geom((48, 40), (1, 1), (1, 1), None)
        spec = oracle.DiskSpec(RN, dict(
            rp=[T(2, params=(1.0, 5.0))], wp=[T(1, consts=(0, 20, 3))]),
            posa=30.0, incl=50.0).rounded_f32()
        grid = oracle.grid_f32(oracle.grid_for(geom, True))
        w = oracle.smdisk(spec, grid, want=('wdata',))[0]['wdata']
        ref = oracle.wcube(w, 1)[0]
        assert 0 < (ref > 0).mean() < 1
        assert np.array_equal(got, ref)


def _opacity_setup(drv, with_opacity):
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelImage
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    bright = gmodels.BrightnessSMDisk3D(
        False, False, make('bpt', 'exponential'), make('bht', 'sech2'),
        rnodes=RN)
    ocmp = gmodels.OpacitySMDisk3D(
        False, False, make('opt', 'exponential'), make('oht', 'sech2'),
        rnodes=RN) if with_opacity else None
    gm = gmodels.GModelIntensity3D(bright, ocmp)
    return Model(drv, DModelImage((40, 40), (1, 1),
                                  psf=psflsf.PSFGauss(1.0)), gm)


def test_intensity3d_image_with_opacity_disk_matches_oracle(drv):
    """A thick brightness disk behind an opacity disk (opacity exponential
    0.15 x exp(-r / 3), sech2 height of amplitude 1 and scale 1.5, tilted
    differently): the opacity disk is evaluated into the 3-D opacity cube
    first, the brightness disk dims every voxel by the opacity in front of
    it and sums the line of sight into the image."""
    model = _opacity_setup(drv, True)
    p = dict(xpos=0.0, ypos=0.0, posa=40.0, incl=60.0, bpt_a=1.0,
             bpt_s=5.0, bht_s=1.0, ocmp_xpos=0.0, ocmp_ypos=0.0,
             ocmp_posa=70.0, ocmp_incl=65.0, ocmp_opt_a=0.15,
             ocmp_opt_s=3.0, ocmp_oht_a=1.0, ocmp_oht_s=1.5)
    got = model.model_h(p)['image']['d']
    geom = _image_geom((40, 40), (1, 1), (1, 1), 1.0)
    grid = oracle.grid_f32(oracle.grid_for(geom, False))
    ospec = oracle.DiskSpec(RN, dict(
        rp=[T(1002, params=(0.15, 3.0))],
        rh=[T(1007, consts=(0, 0), params=(1.0, 1.5))]),
        posa=70.0, incl=65.0).rounded_f32()
    op = oracle.smdisk(ospec, grid, want=('rdata',))[0]['rdata']
    assert 0.02 < op.max() < 0.5
    spec = oracle.DiskSpec(RN, dict(
        rp=[T(2, params=(1.0, 5.0))],
        rh=[T(7, consts=(0, 0), params=(1.0,))]),
        posa=40.0, incl=60.0).rounded_f32()
    clear, _ = oracle.smdisk(spec, grid, want=('image',))
    spec.opacity = op
    out, _ = oracle.smdisk(spec, grid, want=('image', 'amb'))
    ref, amb = _oracle_image(geom, out)
    _check('intensity3d+opacity', got, ref, amb)
    # the opacity matters: the clear image differs by far more than tol
    clear_lo, _ = _oracle_image(geom, dict(clear, amb=out['amb']))
    assert H.normwise(clear_lo, ref) > 100 * TOL
    # a second evaluation starts from a zeroed opacity cube
    again = model.model_h(p)['image']['d']
    assert H.normwise(again, got) < 1e-6


def test_brightness_mcdisk3d_image_matches_oracle(drv):
    """Monte Carlo clouds into an image: the same Philox stream on the GPU
    and in the oracle, every cloud in the same spaxel unless its position
    is an FP32 near-tie."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelImage
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    cmp = gmodels.BrightnessMCDisk3D(
        1e-4, False, False, make('bpt', 'exponential'), make('bht', 'sech2'),
        rnodes=RN)
    model = Model(drv, DModelImage((40, 40), (1, 1), scale=(2, 2),
                                   psf=psflsf.PSFGauss(1.0)),
                  gmodels.GModelIntensity3D(cmp, size_z=32))
    p = dict(xpos=0.0, ypos=0.0, posa=120.0, incl=55.0, bpt_a=1.0,
             bpt_s=5.0, bht_s=1.0)
    got = model.model_h(p)['image']['d']
    csum, hoi = cmp.disk.cloud_budget(p)
    assert csum[-1] > 1e5
    geom = _image_geom((40, 40), (1, 1), (2, 2), 1.0)
    grid = oracle.grid_f32(oracle.grid_for(geom, False, 32))
    spec = oracle.DiskSpec(RN, dict(
        rp=[T(2, params=(1.0, 5.0))],
        rh=[T(7, consts=(0, 0), params=(1.0,))]),
        posa=120.0, incl=55.0).rounded_f32()
    seed = drv.backends().gmodel(np.float32).seed
    out, _ = oracle.mcdisk(spec, grid, cmp.disk.cflux, csum, hoi, seed=seed,
                           want=('image', 'amb'))
    _check('mcdisk image', got, *_oracle_image(geom, out))


def test_lslit_plane_matches_oracle(drv):
    """A thin spectral disk seen through a slit along x: a 48-position x
    32-channel slit one pixel high, PSF and LSF convolution (the PSF
    brings in flux from rows beside the slit), 2x along the slit."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelLSlit
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    cmp = gmodels.SpectralSMDisk2D(
        False, False, make('bpt', 'exponential'), make('vpt', 'tan_arctan'),
        make('dpt', 'uniform'), rnodes=RN)
    psf, lsf = 1.2, 8.0
    model = Model(drv, DModelLSlit((48, 32), (1, 1, 10), scale=(2, 1, 1),
                                   psf=psflsf.PSFGauss(psf),
                                   lsf=psflsf.LSFGauss(lsf)),
                  gmodels.GModelKinematics2D(cmp))
    pp = dict(P.c1_params(), rnodes=RN, posa=20.0, incl=60.0, bpt_s=5.0,
              rt=4.0, vt=150.0)
    got = model.model_h(P.gparams(pp))['lslit']
    assert got['d'].shape == (32, 48)
    geom = oracle.DCubeGeom((48, 1, 32), (1, 1, 10), scale=(2, 1, 1),
                            psf_sigma=psf, lsf_sigma=lsf)
    grid = oracle.grid_f32(oracle.grid_for(geom, True))
    out, _ = oracle.smdisk(P.oracle_spec(pp), grid, want=('scube', 'amb'))
    lo, _, amb = oracle.evaluate_scube(geom, out['scube'], out['amb'])
    _check('lslit', got['d'], lo[:, 0, :], amb[:, 0, :])
