"""The mcdisk, ragged-group and serving workloads on the GPU at their own
sizes, against the FP64 oracle:

  * plain mcdisk, 256x256x128, 2^24 clouds of an exponential x sech^2 disk
    (lramp rotation curve, dispersion 15 km/s), PSF FWHM 3 px + LSF FWHM
    2 channels, moment maps 0/1/2: the convolved cube per voxel against the
    oracle's Philox cloud stream (within 1e-5 max + the propagated ambiguity
    bound), and the moment maps against the oracle's moments of its cube;
  * a ragged group spanning the ragged batch's range: a 64x64x32 thin
    smdisk at 2x, a 256x256x32 thin smdisk at 3x and a 96x96x48 mcdisk with
    2^20 clouds, all on 2/3/5-smooth padding: every member's cubes against
    the oracle, and the group equal to each member evaluated alone;
  * a serving mix of resident galaxies from 64x64x32 to 256x256x128, half
    mcdisk with 2^20 clouds: every request's chi^2 and moment maps in the
    mix equal the same request alone, bit for bit."""
import random
import threading

import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _moment_bounds(cube, amb, geom, m0, m1, m2):
    """First-order bounds on the moment maps' error from a per-voxel error
    bound b = 4e-7 max|cube| (FP32 convolution noise) + amb:
      |dm0| <= dv sum b,  |dm1| <= dv sum |v - m1| b / m0,
      |dm2| <= dv sum |(v - m1)^2 - m2^2| b / (2 m0 m2)."""
    dv = geom.step_lo[2]
    v = geom.zero_lo[2] + dv * np.arange(cube.shape[0])
    b = 4e-7 * np.abs(cube).max() + amb
    dev = v[:, None, None] - m1[None]
    return (dv * b.sum(0),
            dv * (np.abs(dev) * b).sum(0) / np.abs(m0),
            dv * (np.abs(dev ** 2 - m2[None] ** 2) * b).sum(0) /
            (2 * np.abs(m0) * m2))


def _oracle_mc_lo(p, pad, cflux, csum, seed, size_z):
    """Oracle mcdisk (Philox stream) of a pipeline_cases disk description,
    convolved and downscaled: (lo cube, lo ambiguity bound, geometry)."""
    geom = P.oracle_geom(p, pad)
    grid = oracle.grid_f32(oracle.grid_for(geom, False, size_z))
    spec = P.oracle_spec(dict(p, thin=False))
    out, _ = oracle.mcdisk(spec, grid, cflux, csum,
                           np.zeros(1, np.uint8), seed=seed,
                           want=('scube', 'amb'))
    lo, _, amb = oracle.evaluate_scube(geom, out['scube'], out['amb'])
    return lo, amb, geom


def _c2_model(drv, orders):
    """The plain mcdisk workload through the model layer, with its cloud
    flux set for 2^24 clouds."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelMMaps, DModelSCube
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    rn = np.linspace(0, 120, 41)
    truth = dict(xpos=0.0, ypos=0.0, vsys=0.0, posa=30.0, incl=50.0,
                 bpt_a=1.0, bpt_s=20.0, bht_s=2.0, vpt_rt=15.0,
                 vpt_vt=220.0, dpt_a=15.0)

    def model(cflux):
        cmp = gmodels.SpectralMCDisk3D(
            cflux, False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_lramp'), make('dpt', 'uniform'),
            make('bht', 'sech2'), rnodes=rn)
        gm = gmodels.GModelKinematics3D([cmp], size_z=32)
        kw = dict(psf=psflsf.PSFGauss(3.0 * P.SIG_FWHM),
                  lsf=psflsf.LSFGauss(20.0 * P.SIG_FWHM))
        dm = DModelMMaps((256, 256, 128), (1, 1, 10), orders=orders, **kw) \
            if orders else DModelSCube((256, 256, 128), (1, 1, 10), **kw)
        return cmp, Model(drv, dm, gm)
    cmp, _ = model(1e-3)
    total = float(cmp.disk.cloud_budget(truth)[0][-1]) * 1e-3
    cmp, m = model(total / float(1 << 24))
    return cmp, m, truth


def test_c2_plain_mcdisk_cube_and_moment_maps():
    drv = H.driver()
    cmp, model, truth = _c2_model(drv, None)
    csum, hoi = cmp.disk.cloud_budget(truth)
    assert not hoi.any()
    n = int(csum[-1])
    assert abs(n - (1 << 24)) <= 1e-4 * (1 << 24), n
    got = model.model_h(truth)['scube']['d']
    # the oracle's description of the same disk (lramp: uid 5)
    p = dict(size=(256, 256, 128), step=(1, 1, 10), scale=(1, 1, 1),
             psf_fwhm=3.0, lsf_fwhm=20.0, rnodes=np.linspace(0, 120, 41),
             posa=30.0, incl=50.0, vt=220.0, rt=15.0, disp=15.0, bpt_a=1.0,
             bpt_s=20.0, thin=False, bht_s=2.0, size_z=32)
    geom = P.oracle_geom(p)
    grid = oracle.grid_f32(oracle.grid_for(geom, False, 32))
    T = oracle.Trait
    spec = oracle.DiskSpec(p['rnodes'], dict(
        rp=[T(2, params=(1.0, 20.0))], rh=[T(7, consts=(0, 0),
                                             params=(2.0,))],
        vp=[T(5, params=(15.0, 220.0))], vh=[T(1, consts=(0,))],
        dp=[T(1, params=(15.0,))], dh=[T(1, consts=(0,))]),
        posa=30.0, incl=50.0).rounded_f32()
    seed = drv.backends().gmodel(np.float32).seed
    out, _ = oracle.mcdisk(spec, grid, cmp.disk.cflux, csum, hoi, seed=seed,
                           want=('scube', 'amb'))
    lo, _, amb = oracle.evaluate_scube(geom, out['scube'], out['amb'])
    nbad, err = H.within(got, lo, TOL, amb)
    print(f'C2 cube: {n} clouds, normwise error {err:.2e}, voxels beyond '
          f'1e-5 max + amb: {nbad}')
    assert nbad == 0
    # moment maps 0/1/2 of the same evaluation through DModelMMaps
    _, mm_model, _ = _c2_model(drv, (0, 1, 2))
    maps = mm_model.model_h(truth)
    md, mm, _ = oracle.moments(lo, geom.step_lo, geom.zero_lo, (0, 1, 2),
                               1e-6)
    sel = np.abs(md[0]) > 1e-3 * np.nanmax(np.abs(md[0]))
    bounds = _moment_bounds(lo, amb, geom, *md)
    for i, key in enumerate(('mmap0', 'mmap1', 'mmap2')):
        g = maps[key]['d'].astype(np.float64)
        assert np.all(np.isfinite(g[sel])), key
        scale = np.nanmax(np.abs(md[i][sel]))
        excess = np.abs(g[sel] - md[i][sel]) - bounds[i][sel]
        print(key, 'max rel', np.max(np.abs(g[sel] - md[i][sel])) / scale,
              'beyond bound', np.max(excess) / scale)
        assert np.max(excess) / scale < TOL, key


def _c4_members(drv):
    """(kind, description, model, params, mock, sets) of three members."""
    small = dict(P.c1_params(), size=(64, 64, 32), psf_fwhm=1.5,
                 lsf_fwhm=10.0)
    large = dict(P.c1_params(), size=(256, 256, 32), scale=(3, 3, 1),
                 psf_fwhm=4.0, lsf_fwhm=30.0, rnodes=np.arange(0, 90, 3.0),
                 bpt_s=25.0, rt=10.0)
    mc = dict(P.c1_params(), size=(96, 96, 48), rnodes=np.arange(0, 38, 2.0),
              bpt_s=10.0, rt=5.0, size_z=32)
    out = []
    for kind, p in (('sm', small), ('sm', large), ('mc', mc)):
        if kind == 'mc':
            m1, base = P.build_mc_product(drv, p, cflux=1e-3,
                                          padding='smooth235')
            total = float(m1.gmodel().components[0].disk.cloud_budget(
                base)[0][-1]) * 1e-3
            p['cflux'] = total / float(1 << 20)
            model, base = P.build_mc_product(drv, p, cflux=p['cflux'],
                                             padding='smooth235')
        else:
            model, base = P.build_product(drv, p, padding='smooth235')
        ref = model.model_h(base)['scube']['d']
        rng = np.random.default_rng(len(out) + 40)
        mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
        sets = [dict(base, incl=i, posa=a) for i, a in ((45, 25), (65, 200))]
        out.append((kind, p, model, base, mock, sets))
    return out


def test_c4_ragged_group_members_vs_oracle_and_alone():
    from gbkfit_amd.batch import BatchEvaluator, RaggedGroup
    drv = H.driver()
    members = _c4_members(drv)
    evs = [BatchEvaluator(drv, m.dmodel(), m.gmodel(), mock, error=0.01,
                          max_batch=2, template_params=base)
           for _, _, m, base, mock, _ in members]
    for ev in evs:
        assert ev.info()['pipeline'] == 'direct'
    grp = RaggedGroup(evs)
    lists = [sets for *_, sets in members]
    chi2, cubes = grp.evaluate(lists, want_cubes=True, use_graph=True)
    again = grp.evaluate(lists, use_graph=True)
    seed = drv.backends().gmodel(np.float32).seed
    for (kind, p, model, base, mock, sets), ev, c, cb, c2 in zip(
            members, evs, chi2, cubes, again):
        alone_c, alone_cb = ev.evaluate(sets, want_cubes=True,
                                        use_graph=False)
        assert np.array_equal(alone_c, c) and np.array_equal(c2, c), kind
        assert np.array_equal(alone_cb, cb), kind
        for i, s in enumerate(sets):
            q = dict(p, incl=s['incl'], posa=s['posa'])
            if kind == 'mc':
                csum, _ = model.gmodel().components[0].disk.cloud_budget(s)
                assert abs(int(csum[-1]) - (1 << 20)) <= 1e-3 * (1 << 20)
                ref, amb, _ = _oracle_mc_lo(q, 'smooth235', p['cflux'],
                                            csum, seed, p['size_z'])
            else:
                ref, amb, _ = P.oracle_cube(q, pad='smooth235')
            nbad, err = H.within(cb[i], ref, TOL, amb)
            print(kind, p['size'], p['scale'], 'set', i, 'normwise', err,
                  'beyond bound', nbad, 'hi-res', ev.size_hi)
            assert nbad == 0, (kind, p['size'], i)
    grp.close()
    for ev in evs:
        ev.close()


SERVE = [('smdisk', (64, 64, 32)), ('mcdisk', (96, 96, 48)),
         ('smdisk', (160, 160, 80)), ('mcdisk', (256, 256, 128))]


def _serve_galaxy(drv, kind, size, i, streams):
    from gbkfit_amd.serve import GpuGalaxy
    import bench
    model, truth = bench.build_galaxy_model(drv, size, kind=kind,
                                            nclouds=1 << 20)
    cube = model.model_h(truth)['scube']['d'].astype(np.float32)
    rng = np.random.default_rng(60 + i)
    mock = (cube + rng.normal(0, 0.01, cube.shape)).astype(np.float32)
    g = GpuGalaxy(drv, model.dmodel(), model.gmodel(), mock, error=0.01,
                  max_batch=8, template_params=truth, orders=(0, 1, 2),
                  streams=streams)
    rs = np.random.default_rng(200 + i)
    sets = [dict(truth, incl=float(rs.uniform(25, 75)),
                 posa=float(rs.uniform(0, 360)),
                 vpt_vt=float(rs.uniform(100, 300)),
                 dpt_a=float(rs.uniform(5, 50))) for _ in range(12)]
    return g, sets


def test_c5_serving_mix_with_mcdisk_equals_alone():
    from gbkfit_amd.serve import Server, make_streams
    drv = H.driver()
    streams = make_streams(4)
    gals, sets = {}, {}
    for i, (kind, size) in enumerate(SERVE):
        gals[i], sets[i] = _serve_galaxy(drv, kind, size, i, streams)
    assert any(g.ev.mc for g in gals.values())
    works = sorted(g.work_bytes() for g in gals.values())
    srv = Server(hbm_budget=works[-1] + works[-2], slots=4, max_batch=8)
    for k, g in gals.items():
        srv.add(k, g)
    reqs = [(k, i) for k in gals for i in range(len(sets[k]))]
    random.Random(9).shuffle(reqs)
    futs, lock = {}, threading.Lock()

    def client(part):
        rnd = random.Random(len(part))
        for k, i in part:
            f = srv.submit(k, sets[k][i])
            with lock:
                futs[(k, i)] = f
            if rnd.random() < 0.3:
                threading.Event().wait(0.0005)

    srv.start()
    th = [threading.Thread(target=client, args=(reqs[j::3],))
          for j in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.stop()
    mixed = {key: f.result(120) for key, f in futs.items()}
    print('stats', {k: v for k, v in srv.stats.items() if k != 'batch_sizes'})
    assert srv.stats['completed'] == len(reqs)
    assert srv.stats['max_batch_seen'] > 1 and srv.stats['releases'] > 0
    for (k, i), res in mixed.items():
        tok = gals[k].launch([sets[k][i]], 0)
        alone = gals[k].results(tok, 1)[0]
        assert alone['chi2'] == res['chi2'], (SERVE[k], i)
        assert np.array_equal(alone['mmaps'], res['mmaps'], equal_nan=True)
    for g in gals.values():
        g.close()
