"""Builds the same model twice -- through gbkfit_amd's model layer (what a
gbkfit user calls) and through the FP64 oracle -- for pipeline tests, the
smoke check and the benchmark's CPU baseline."""
import numpy as np

import oracle

SIG_FWHM = 1 / np.sqrt(8 * np.log(2))


def c1_params():
    """Plain smdisk, 64x64x64 (the CPU-runnable case): 20 nodes at 1.5 px,
    incl 45, PA 30, exponential (1, 8 px), arctan (vmax 200, rt 6), sigma
    20 km/s, channel 10 km/s, PSF FWHM 2.5 px, LSF FWHM 2 channels, 2x
    spatial."""
    return dict(size=(64, 64, 64), step=(1, 1, 10), scale=(2, 2, 1),
                psf_fwhm=2.5, lsf_fwhm=20.0, rnodes=np.arange(0, 30, 1.5),
                posa=30.0, incl=45.0, vt=200.0, rt=6.0, disp=20.0,
                bpt_a=1.0, bpt_s=8.0, thin=True, bht_s=2.0, size_z=None)


def c3_params(rng=None, n=1):
    """Batched smdisk, 128x128x96 (the batched workload): per set incl
    U(20,80), PA U(0,360), vmax U(100,300), rt U(2,20), sigma U(5,50)."""
    rng = rng or np.random.default_rng(0)
    base = dict(size=(128, 128, 96), step=(1, 1, 10), scale=(2, 2, 1),
                psf_fwhm=3.0, lsf_fwhm=20.0, rnodes=np.arange(0, 30, 1.5),
                bpt_a=1.0, bpt_s=8.0, thin=True, bht_s=2.0, size_z=None)
    sets = []
    for _ in range(n):
        p = dict(base)
        p.update(incl=rng.uniform(20, 80), posa=rng.uniform(0, 360),
                 vt=rng.uniform(100, 300), rt=rng.uniform(2, 20),
                 disp=rng.uniform(5, 50))
        sets.append(p)
    return sets


def gparams(p):
    """The gbkfit parameter dict of a disk description."""
    d = dict(xpos=0.0, ypos=0.0, posa=p['posa'], incl=p['incl'], vsys=0.0,
             bpt_a=p['bpt_a'], bpt_s=p['bpt_s'], vpt_rt=p['rt'],
             vpt_vt=p['vt'], dpt_a=p['disp'])
    if not p['thin']:
        d['bht_s'] = p['bht_s']
    return d


def build_product(driver, p, orders=None, padding='reference'):
    """(Model, params) through gbkfit_amd's model layer."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelMMaps, DModelSCube
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    psf = psflsf.PSFGauss(p['psf_fwhm'] * SIG_FWHM)
    lsf = psflsf.LSFGauss(p['lsf_fwhm'] * SIG_FWHM)
    if p['thin']:
        cmp = gmodels.SpectralSMDisk2D(
            False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
            rnodes=p['rnodes'])
        gm = gmodels.GModelKinematics2D([cmp])
    else:
        cmp = gmodels.SpectralSMDisk3D(
            False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
            make('bht', 'sech2'), rnodes=p['rnodes'])
        gm = gmodels.GModelKinematics3D([cmp], size_z=p['size_z'])
    if orders is None:
        dm = DModelSCube(p['size'], p['step'], scale=p['scale'], psf=psf,
                         lsf=lsf, padding=padding)
    else:
        dm = DModelMMaps(p['size'], p['step'], scale=p['scale'], psf=psf,
                         lsf=lsf, orders=orders, padding=padding)
    return Model(driver, dm, gm), gparams(p)


def build_mc_product(driver, p, cflux, padding='reference'):
    """(Model, params) of a Monte Carlo cloud disk (SpectralMCDisk3D with
    the same traits as the thick smdisk) through gbkfit_amd's model layer."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    psf = psflsf.PSFGauss(p['psf_fwhm'] * SIG_FWHM)
    lsf = psflsf.LSFGauss(p['lsf_fwhm'] * SIG_FWHM)
    cmp = gmodels.SpectralMCDisk3D(
        cflux, False, False, make('bpt', 'exponential'),
        make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
        make('bht', 'sech2'), rnodes=p['rnodes'])
    gm = gmodels.GModelKinematics3D([cmp], size_z=p['size_z'])
    dm = DModelSCube(p['size'], p['step'], scale=p['scale'], psf=psf,
                     lsf=lsf, padding=padding)
    return Model(driver, dm, gm), gparams(dict(p, thin=False))


def oracle_spec(p):
    T = oracle.Trait
    traits = dict(rp=[T(2, params=(p['bpt_a'], p['bpt_s']))],
                  vp=[T(2, params=(p['rt'], p['vt']))],
                  dp=[T(1, params=(p['disp'],))])
    if not p['thin']:
        traits['rh'] = [T(7, consts=(0, 0), params=(p['bht_s'],))]
        traits['vh'] = [T(1, consts=(0,))]
        traits['dh'] = [T(1, consts=(0,))]
    return oracle.DiskSpec(p['rnodes'], traits, posa=p['posa'],
                           incl=p['incl']).rounded_f32()


def oracle_geom(p, pad='reference'):
    return oracle.DCubeGeom(p['size'], p['step'], scale=p['scale'],
                            psf_sigma=p['psf_fwhm'] * SIG_FWHM,
                            lsf_sigma=p['lsf_fwhm'] * SIG_FWHM, pad=pad)


def oracle_cube(p, pad='reference'):
    """(low-res cube, ambiguity bound, stats) from the FP64 oracle."""
    geom = oracle_geom(p, pad)
    grid = oracle.grid_f32(oracle.grid_for(geom, p['thin'], p['size_z']))
    out, stats = oracle.smdisk(oracle_spec(p), grid, want=('scube', 'amb'))
    lo, _, amb = oracle.evaluate_scube(geom, out['scube'], out['amb'])
    return lo, amb, stats


def excess_error(got, ref, amb):
    """max(|got - ref| - amb, 0) / max|ref|: the error not explained by the
    oracle's ambiguity bound."""
    ref = np.asarray(ref, np.float64)
    d = np.abs(np.asarray(got, np.float64) - ref) - amb
    return float(max(d.max(), 0.0) / np.abs(ref).max())
