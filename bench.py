"""Benchmark: model-cube evaluations/s of the 128x128x96 smdisk + PSF + LSF
workload (512 parameter sets per step), batched vs plain, with latencies.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nsets 512]

A step evaluates `nsets` parameter sets of a thin smooth disk (20 ring
nodes at 1.5 px, exponential 1/8 px, arctan rotation curve, uniform
dispersion; per set incl U(20,80), PA U(0,360), vmax U(100,300),
rt U(2,20), sigma U(5,50)) on a 128x128x96 cube, 2x spatial supersampling,
PSF FWHM 3 px, LSF FWHM 2 channels (10 km/s), reference DCube padding
(model grid 320x320x128), each against a seeded mock cube (noise sigma
0.01) -> one FP64 chi^2 per set back on the host.  The batched path takes
its automatic choice, the direct separable pipeline: x, y and z filters
with the block mean folded into the taps over the 296x296x110 inputs that
reach the outputs, skipping tiles no line reaches (the fused FFT pipeline,
300x300x120, gives the same result and is timed in the tests), in chunks
of 256 sets (`--chunk`).  The timed region covers packing the parameter
block, the upload, the batched native evaluation (one HIP graph replay per
chunk sequence) and the download of the chi^2 vector.

A second timed region repeats the steps with direct launches and HIP events
after every kernel (on the kernels' own stream): the per-kernel times give
the roofline of the dominant kernel, algorithmic bytes per launch over its
average launch time.

The plain path is the drop-in call sequence of the reference
(Objective.log_likelihood per set: pack + upload trait arrays, smdisk,
FFT convolution, downscale, residual, residual sum, scalar download).

A thick leg (skip with --no-thick) times the same workload with a thick
disk (sech^2 height of 2 px on a spatial z axis of 64): the deterministic
thick deposit into a hi-res cube, then the direct pipeline's filters; its
kernel times, roofline and the deposit's op count (W = 60 V + 12 sum n_ch
on the 3-D grid) go into the `thick_variant` object.

An FFT leg (skip with --no-fft) times the thin workload with an
elliptical, rotated PSF (axis ratio 0.7, PA 30 deg), which does not factor,
on the fused FFT pipeline and on the padded-cube one (256 sets per step):
the `fft_pipelines` object, with per-kernel times and each pipeline's
roofline.

A serving leg (skip with --no-serve) replays a Poisson stream of
single-set requests (default 1000/s for 60 s) over 64 resident galaxies with
cube sizes log-uniform from 64x64x32 to 256x256x128, half smdisk, half mcdisk
with 2^20 clouds; each request returns chi^2 and moment maps 0/1/2 through
gbkfit_amd.serve.Server (continuous batching over 4 streams, up to 8
galaxies per slot launch, HBM-budget admission), in steady state: before
the stream every galaxy has captured the graphs of batches of 1-16 sets and
holds its work buffers.  Its throughput and p50/p99 latencies go into the
`serving` object, with a 10 s run under a budget of half the galaxies' work
bytes (`admission`: releases and re-admissions included) and an offered-rate
sweep up to the first rate not sustained (`sweep`, `saturation_rate`).

A ragged leg (skip with --no-ragged) autotunes 32 galaxies of log-uniform
sizes from 64x64x32 to 512x512x256 (smdisk / mcdisk with 2^20..2^24 clouds,
2x or 3x supersampling, random PSF / LSF widths, 2/3/5-smooth padding) and
times one graph launch evaluating one set of every galaxy, plus the moment
maps of all of them: the `ragged` object.  A plain mcdisk leg (skip with
--no-mcdisk) times the one-model path on 256x256x128 with 2^24 clouds: the
`mcdisk_plain` object.

Prints one JSON line.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

METRIC = ("model-cube evals/sec (128²×96 smdisk+PSF+LSF), "
          "batched vs plain; p99 latency")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
SIG = 1 / math.sqrt(8 * math.log(2))


def workload(nsets, seed=1234, thick=False):
    """Parameter arrays of `nsets` sets; a thick disk adds its sech^2
    scale height (2 px, the thick variant of the 64^3 case)."""
    rng = np.random.default_rng(seed)
    p = dict(incl=rng.uniform(20, 80, nsets), posa=rng.uniform(0, 360, nsets),
             vt=rng.uniform(100, 300, nsets), rt=rng.uniform(2, 20, nsets),
             disp=rng.uniform(5, 50, nsets))
    arrays = dict(xpos=np.zeros(nsets), ypos=np.zeros(nsets),
                  vsys=np.zeros(nsets), posa=p['posa'], incl=p['incl'],
                  bpt_a=np.ones(nsets), bpt_s=np.full(nsets, 8.0),
                  vpt_rt=p['rt'], vpt_vt=p['vt'], dpt_a=p['disp'])
    if thick:
        arrays['bht_s'] = np.full(nsets, 2.0)
    return arrays


def param_dicts(arrays, n):
    return [{k: float(v[i]) for k, v in arrays.items()} for i in range(n)]


def build_model(driver, thick=False, psf=None):
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    rn = np.arange(0, 30, 1.5)
    if thick:
        cmp = gmodels.SpectralSMDisk3D(
            False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
            make('bht', 'sech2'), rnodes=rn)
        gm = gmodels.GModelKinematics3D([cmp], size_z=64)
    else:
        cmp = gmodels.SpectralSMDisk2D(
            False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'), rnodes=rn)
        gm = gmodels.GModelKinematics2D([cmp])
    dm = DModelSCube((128, 128, 96), (1, 1, 10), scale=(2, 2, 1),
                     psf=psf if psf is not None else psflsf.PSFGauss(3.0 * SIG),
                     lsf=psflsf.LSFGauss(20.0 * SIG))
    return Model(driver, dm, gm)


def build_galaxy_model(driver, size, scale=(1, 1, 1), rmax=None,
                       kind='smdisk', psf_fwhm=2.5, lsf_ch=2.0,
                       padding='reference', nclouds=None, size_z=None):
    """(Model, truth params) of a galaxy in a `size` cube: a thin smdisk
    (Kinematics2D) or a Monte Carlo cloud disk with a sech^2 height profile
    (Kinematics3D, explicit spatial z size, about `nclouds` clouds), ring
    nodes out to ~40 % of the field, Gaussian PSF / LSF (FWHM in pixels and
    channels of 10 km/s)."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    rmax = rmax or 0.4 * min(size[0], size[1])
    rn = np.linspace(0, rmax, 16)
    truth = dict(xpos=0.0, ypos=0.0, vsys=0.0, posa=30.0, incl=50.0,
                 bpt_a=1.0, bpt_s=0.15 * size[0], vpt_rt=0.05 * size[0],
                 vpt_vt=200.0, dpt_a=20.0)
    dm = DModelSCube(tuple(size), (1, 1, 10), scale=scale,
                     psf=psflsf.PSFGauss(psf_fwhm * SIG),
                     lsf=psflsf.LSFGauss(10.0 * lsf_ch * SIG),
                     padding=padding)
    if kind == 'smdisk':
        cmp = gmodels.SpectralSMDisk2D(
            False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'), rnodes=rn)
        return Model(driver, dm, gmodels.GModelKinematics2D([cmp])), truth
    truth['bht_s'] = 1.0

    def mc(cflux):
        cmp = gmodels.SpectralMCDisk3D(
            cflux, False, False, make('bpt', 'exponential'),
            make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
            make('bht', 'sech2'), rnodes=rn)
        return cmp, gmodels.GModelKinematics3D([cmp], size_z=size_z or 16)
    # flux per cloud for about `nclouds` clouds over the disk's integral
    cmp, _ = mc(1e-3)
    total = float(cmp.disk.cloud_budget(truth)[0][-1]) * 1e-3
    cmp, gm = mc(total / float(nclouds or (1 << 20)))
    return Model(driver, dm, gm), truth


def serve_benchmark(drv, ngal, seconds, rate, slots=4, max_batch=16,
                    seed=77, group=8, admission_seconds=10.0,
                    sweep_rates=(2000.0, 4000.0, 8000.0, 16000.0),
                    sweep_seconds=5.0):
    """Poisson streams of single-set requests (chi^2 + moment maps 0/1/2)
    over `ngal` resident galaxies with cube sizes log-uniform from
    64x64x32 to 256x256x128, half mcdisk with 2^20 clouds.

    Three measurements on the same galaxies:
      main:      `rate` requests/s for `seconds`, every galaxy running with
                 the work arena of the slot it is launched in (work memory
                 = slots x the largest galaxy's arena; steady state: each
                 galaxy's graphs for batches of 1-8 in every slot captured);
      admission: the same rate for `admission_seconds` with work buffers
                 of each galaxy's own under a budget of half their total,
                 so the server releases idle galaxies (least recently used)
                 to admit others (allocation and graph capture included);
      sweep:     `sweep_seconds` at each of `sweep_rates` until the first
                 rate the server does not sustain (throughput below 0.9 of
                 the offered rate, or p99 above 10 times the main p99);
                 saturation = the highest rate sustained.
    A slot launches up to `group` galaxies together."""
    import torch
    from gbkfit_amd.serve import (GpuGalaxy, Server, make_arenas,
                                  make_streams)
    rng = np.random.default_rng(seed)
    streams = make_streams(slots)
    gals, sizes = [], []
    t0 = time.perf_counter()
    kinds = []
    for g in range(ngal):
        f = np.exp(rng.uniform(np.log(1.0), np.log(4.0)))
        size = (int(64 * f) // 2 * 2, int(64 * f) // 2 * 2, int(32 * f))
        kind = 'mcdisk' if g % 2 else 'smdisk'
        model, truth = build_galaxy_model(drv, size, kind=kind,
                                          nclouds=1 << 20)
        kinds.append(kind)
        cube = model.model_h(truth)['scube']['d'].astype(np.float32)
        mock = (cube + rng.normal(0, 0.01, cube.shape)).astype(np.float32)
        gals.append(GpuGalaxy(drv, model.dmodel(), model.gmodel(), mock,
                              error=0.01, max_batch=max_batch,
                              template_params=truth, orders=(0, 1, 2),
                              streams=streams))
        sizes.append(size)
    works = [g.work_bytes() for g in gals]
    arena = max(g.arena_bytes() for g in gals)
    arenas = make_arenas(arena, slots)

    def request(t, r):
        return dict(xpos=0.0, ypos=0.0, vsys=0.0,
                    posa=float(r.uniform(0, 360)),
                    incl=float(r.uniform(20, 80)), bpt_a=1.0,
                    bpt_s=0.15 * sizes[t][0], vpt_rt=0.05 * sizes[t][0],
                    vpt_vt=float(r.uniform(100, 300)),
                    dpt_a=float(r.uniform(5, 50)),
                    **({'bht_s': 1.0} if kinds[t] == 'mcdisk' else {}))

    # steady state: every galaxy's graphs for batches of 1..8 captured in
    # every slot's arena
    def warm():
        r0 = np.random.default_rng(seed + 1)
        for i, g in enumerate(gals):
            g.use_arenas(arenas)
            p0 = request(i, r0)
            for slot in range(slots):
                for k in range(1, 9):
                    g.results(g.launch([p0] * k, slot), k)
        torch.cuda.synchronize()

    warm()
    t_setup = time.perf_counter() - t0

    def stream(rate_, seconds_, budget, sseed):
        r = np.random.default_rng(sseed)
        n = int(rate_ * seconds_)
        arrivals = np.cumsum(r.exponential(1.0 / rate_, n))
        targets = r.integers(0, ngal, n)
        pars = [request(t, r) for t in targets]
        if budget is not None:          # work buffers of the galaxies' own
            for g in gals:
                g.use_arenas(None)
        srv = Server(hbm_budget=budget, slots=slots, max_batch=max_batch,
                     group=group)
        for i, g in enumerate(gals):
            srv.add(i, g)
        futs = []
        srv.start()
        start = time.perf_counter()
        for i in range(n):
            delay = start + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(int(targets[i]), pars[i]))
        for f in futs:
            f.result(300)
        end = time.perf_counter()
        srv.stop()
        st = srv.stats
        return dict(offered_rate=rate_, seconds=seconds_, requests=n,
                    throughput=n / (end - start),
                    **srv.latency_percentiles(), launches=st['launches'],
                    galaxy_launches=st['galaxy_launches'],
                    mean_batch=float(np.mean(st['batch_sizes'])),
                    max_batch_seen=st['max_batch_seen'],
                    releases=st['releases'], hbm_budget_bytes=budget)

    main = stream(rate, seconds, None, seed + 2)
    print(f"serving main done: p99 {main['p99_ms']:.2f} ms, mean batch "
          f"{main['mean_batch']:.2f}", file=sys.stderr, flush=True)
    admission = stream(rate, admission_seconds, int(0.5 * sum(works)),
                       seed + 3)
    print(f"serving admission done: {admission['releases']} releases",
          file=sys.stderr, flush=True)
    for g in gals:    # back to the slot arenas (graphs still cached)
        g.use_arenas(arenas)
    sweep, saturation = [], rate
    for k, rr in enumerate(sweep_rates):
        res = stream(rr, sweep_seconds, None, seed + 10 + k)
        ok = (res['throughput'] >= 0.9 * rr and
              res['p99_ms'] <= 10 * main['p99_ms'])
        sweep.append(dict(res, sustained=ok))
        print(f"serving sweep {rr:.0f} req/s: {res['throughput']:.0f} "
              f"req/s, p99 {res['p99_ms']:.2f} ms", file=sys.stderr,
              flush=True)
        if not ok:
            break
        saturation = rr
    out = dict(main, galaxies=ngal, slots=slots, max_batch=max_batch,
               group=group, work_bytes_total=int(sum(works)),
               arena_bytes_per_slot=int(arena),
               work_bytes_in_use=int(arena) * slots,
               admission=admission, sweep=sweep,
               saturation_rate=saturation,
               cube_sizes=dict(min=list(min(sizes)), max=list(max(sizes))),
               models=dict(smdisk=kinds.count('smdisk'),
                           mcdisk=kinds.count('mcdisk'),
                           mcdisk_clouds=1 << 20),
               setup_s=t_setup, returns='chi2 + moment maps 0/1/2')
    for g in gals:
        g.close()
    return out


def ragged_benchmark(drv, ngal, reps, seed=2024, budget_frac=0.6):
    """Ragged batch: `ngal` galaxies with cube sizes log-uniform from
    64x64x32 to 512x512x256, smdisk or mcdisk 50/50 (2^20..2^24 clouds),
    2x or 3x spatial supersampling, PSF FWHM U(1.5, 4) px, LSF FWHM U(1, 3)
    channels, 2/3/5-smooth FFT padding, explicit spatial z size.  Each galaxy
    is autotuned (pipeline, tiles) on its own parameter set; the group then
    evaluates one set per galaxy in one graph launch, followed by the moment
    maps (orders 0, 1, 2) of every galaxy's cube."""
    import torch
    from gbkfit_amd.batch import RaggedGroup, autotune
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    free, _ = torch.cuda.mem_get_info()
    budget = budget_frac * free
    evs, metas, lists = [], [], []
    used = 0
    for g in range(ngal):
        f = np.exp(rng.uniform(0.0, np.log(8.0)))
        size = (int(64 * f) // 2 * 2, int(64 * f) // 2 * 2, int(32 * f))
        kind = 'mcdisk' if rng.uniform() < 0.5 else 'smdisk'
        ncl = int(np.exp(rng.uniform(np.log(2.0 ** 20), np.log(2.0 ** 24))))
        sc = int(rng.choice([2, 3]))
        psf = float(rng.uniform(1.5, 4.0))
        lsf = float(rng.uniform(1.0, 3.0))
        model, truth = build_galaxy_model(
            drv, size, scale=(sc, sc, 1), kind=kind, psf_fwhm=psf,
            lsf_ch=lsf, padding='smooth235', nclouds=ncl, size_z=16 * sc)
        cube = model.model_h(truth)['scube']['d'].astype(np.float32)
        mock = (cube + rng.normal(0, 0.01, cube.shape)).astype(np.float32)
        del cube
        p = dict(truth, incl=float(rng.uniform(20, 80)),
                 posa=float(rng.uniform(0, 360)))
        ev, rep = autotune(drv, model.dmodel(), model.gmodel(), mock, [p],
                           error=0.01, max_batches=(1,), reps=2,
                           budget_bytes=max(budget - used, 0))
        mem = ev.memory()
        used += mem['resident'] + mem['work']
        evs.append(ev)
        lists.append([p])
        dc = model.dmodel().dcube()
        metas.append(dict(size=size, kind=kind, scale=sc,
                          clouds=ncl if kind == 'mcdisk' else 0,
                          pipeline=rep['choice'][0], tiles=rep['choice'][2],
                          us_per_eval=1e6 * rep['s_per_set'],
                          lo=tuple(dc.size()), step=tuple(dc.step()),
                          zero=tuple(dc.zero()),
                          hi_voxels=int(np.prod(ev.size_hi))))
    t_setup = time.perf_counter() - t0
    grp = RaggedGroup(evs)
    be = drv.backends().dmodel(np.float32)
    orders = torch.tensor([0, 1, 2], dtype=torch.int32, device=drv.device)
    maps = [torch.empty((3, m['lo'][1], m['lo'][0]), dtype=torch.float32,
                        device=drv.device) for m in metas]
    mmask = [torch.empty((m['lo'][1], m['lo'][0]), dtype=torch.float32,
                         device=drv.device) for m in metas]

    def one():
        chi2, _ = grp.evaluate(lists, want_cubes=True, use_graph=True,
                               download_cubes=False)
        for ev, m, mp, mk in zip(evs, metas, maps, mmask):
            be.mmaps_moments(m['lo'], m['step'], m['zero'],
                             ev._cubes_dev[0], None, 1e-6, orders, mp, None,
                             mk)
        torch.cuda.synchronize()
        return chi2
    one()
    one()
    ts = []
    for _ in range(reps):
        t1 = time.perf_counter()
        chi2 = one()
        ts.append(time.perf_counter() - t1)
    grp.close()
    for ev in evs:
        ev.close()
    ms = 1e3 * float(np.median(ts))
    hv = sum(m['hi_voxels'] for m in metas)
    pipes = {}
    for m in metas:
        pipes[m['pipeline']] = pipes.get(m['pipeline'], 0) + 1
    return dict(galaxies=ngal, ms_per_batch=ms, galaxies_per_s=1e3 * ngal / ms,
                padded_hi_voxels_per_s=hv / (ms * 1e-3), reps=reps,
                chi2_finite=bool(all(np.isfinite(c).all() for c in chi2)),
                models=dict(smdisk=sum(m['kind'] == 'smdisk' for m in metas),
                            mcdisk=sum(m['kind'] == 'mcdisk' for m in metas)),
                pipelines=pipes, setup_s=t_setup, device_bytes=int(used),
                cube_sizes=dict(min=list(min(m['size'] for m in metas)),
                                max=list(max(m['size'] for m in metas))),
                returns='cubes + chi2 + moment maps 0/1/2 per galaxy',
                per_galaxy=[{k: m[k] for k in ('size', 'kind', 'scale',
                                               'clouds', 'pipeline',
                                               'us_per_eval')}
                            for m in metas])


def mcdisk_benchmark(drv, reps, nclouds=1 << 24):
    """Plain mcdisk (one model per call): 256x256x128 cube, 2^24 clouds of
    an exponential (h = 20 px) x sech^2 (z0 = 2 px) disk, lramp rotation
    curve (vmax 220 km/s, rt 15 px), dispersion 15 km/s, PSF FWHM 3 px, LSF
    FWHM 2 channels, moment maps 0/1/2; wall time per model evaluation and
    cloud rate."""
    import torch
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.dmodels import DModelMMaps
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
        dm = DModelMMaps((256, 256, 128), (1, 1, 10),
                         psf=psflsf.PSFGauss(3.0 * SIG),
                         lsf=psflsf.LSFGauss(20.0 * SIG), orders=(0, 1, 2))
        return cmp, Model(drv, dm, gm)
    cmp, _ = model(1e-3)
    total = float(cmp.disk.cloud_budget(truth)[0][-1]) * 1e-3
    cmp, m = model(total / nclouds)
    ncl = int(cmp.disk.cloud_budget(truth)[0][-1])
    # HIP events around the cloud deposit (bin, scan, scatter, tile
    # passes: one native call), on the stream its kernels run on (torch's
    # current stream, which the driver binds its context to)
    from gbkfit_amd.driver.hip import DriverBackendGModelHip as G
    native = G.mcdisk_evaluate
    pairs = []

    def timed(self, *a, **k):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        native(self, *a, **k)
        e1.record()
        pairs.append((e0, e1))
    G.mcdisk_evaluate = timed
    try:
        for _ in range(2):
            m.model_d(truth)
        torch.cuda.synchronize()
        pairs.clear()
        ts = []
        for _ in range(reps):
            t1 = time.perf_counter()
            m.model_d(truth)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t1)
    finally:
        G.mcdisk_evaluate = native
    dep_ms = float(np.mean([a.elapsed_time(b) for a, b in pairs]))
    ms = 1e3 * float(np.median(ts))
    # compulsory bytes of the deposit: per cloud a 16-byte record and an
    # 8-byte (tile, rank) written, both read back and the record written
    # again in tile order, then read by its tile's pass; the padded cube
    # read and written once
    n_pad = int(np.prod(m.dmodel().dcube().scratch_size()))
    dep_bytes = 80.0 * ncl + 8.0 * n_pad
    gbs = dep_bytes / (dep_ms * 1e-3) / 1e9
    return dict(workload='mcdisk 256x256x128, 2^24 clouds, PSF + LSF, '
                         'moment maps 0/1/2', clouds=ncl, reps=reps,
                ms_per_eval=ms, clouds_per_s=ncl / (ms * 1e-3),
                deposit_ms=dep_ms,
                roofline=dict(bound='hbm', achieved=gbs, peak=HBM_PEAK_GBS,
                              unit='GB/s', frac=gbs / HBM_PEAK_GBS,
                              traffic=measured_traffic('mcdisk_deposit'),
                              kernel='mcdisk deposit (mcdisk_bin_kernel, '
                                     'mc_scan_tiles, mc_scatter, '
                                     'mc_tile_deposit)',
                              bytes_per_launch=dep_bytes,
                              note='80 B per cloud + 8 B per padded-cube '
                                   'voxel over the HIP-event time of the '
                                   'deposit call'),
                **percentiles(ts))


def percentiles(xs):
    xs = np.asarray(xs, np.float64) * 1e3
    return dict(p50_ms=float(np.percentile(xs, 50)),
                p99_ms=float(np.percentile(xs, 99)),
                mean_ms=float(xs.mean()))


def oracle_errors(check, thick=False):
    """Per-set error of the batched GPU results against the FP64 oracle for
    a seeded sample of the workload's sets: the model cube's maximum error
    relative to max|oracle| (and beyond the oracle's FP32 ambiguity bound)
    and chi^2's relative error.  Runs in the CPU-baseline leg, outside the
    timed regions; `check` holds the sets, their GPU cubes and chi^2 and the
    mock."""
    import oracle
    import pipeline_cases as P
    base = P.c3_params(n=1)[0]
    base['thin'] = not thick
    base['size_z'] = 64
    mock = check['mock'].astype(np.float64)
    out = []
    for idx, s, cube, c2 in zip(check['index'], check['sets'],
                                check['cubes'], check['chi2']):
        p = dict(base, incl=s['incl'], posa=s['posa'], vt=s['vpt_vt'],
                 rt=s['vpt_rt'], disp=s['dpt_a'])
        ref, amb, _ = P.oracle_cube(p)
        r = (ref - mock) / 0.01
        c_ref = float(oracle.residual_sum(r, True))
        out.append(dict(
            set=int(idx),
            cube_max_rel_err=float(np.abs(cube - ref).max() /
                                   np.abs(ref).max()),
            cube_err_beyond_amb=P.excess_error(cube, ref, amb),
            chi2_rel_err=abs(float(c2) - c_ref) / c_ref))
    return out


def cpu_baseline(seconds, thick=False):
    """The reference's own core (compiled by oracle/Makefile into
    oracle/_ref, OpenMP over all host threads) for smdisk, downscale and the
    residual / sum, with scipy.fft (workers = all threads) for the
    convolution FFTW would do; per-set wall time on a bounded sample."""
    import scipy.fft
    from oracle import refcore
    import pipeline_cases as P
    if not refcore.available():
        return None
    R = refcore.ref(np.float32)
    nthreads = refcore.num_threads()
    # enough sets that the time limit, not the sample, ends the loop
    arrays = workload(1024, seed=99)
    sets = param_dicts(arrays, 1024)
    base = P.c3_params(n=1)[0]
    base['thin'] = not thick
    geom = P.oracle_geom(base)
    pc = geom.pcube(np.float32)
    kspec = scipy.fft.rfftn(pc, workers=nthreads)
    grid = oracle_grid = None
    import oracle
    oracle_grid = oracle.grid_f32(oracle.grid_for(geom, not thick, 64))
    mock = np.zeros(geom.size_lo[::-1], np.float32)
    err = np.full(mock.shape, 0.01, np.float32)
    done, t0 = 0, time.perf_counter()
    while done < len(sets) and time.perf_counter() - t0 < seconds:
        p = dict(base)
        s = sets[done]
        p.update(incl=s['incl'], posa=s['posa'], vt=s['vpt_vt'],
                 rt=s['vpt_rt'], disp=s['dpt_a'])
        spec = P.oracle_spec(p)
        cube = R.smdisk(spec, oracle_grid, ('scube',), parallel=True)['scube']
        f = scipy.fft.rfftn(cube, workers=nthreads)
        f *= kspec
        conv = scipy.fft.irfftn(f, s=cube.shape, axes=(0, 1, 2),
                                workers=nthreads).astype(np.float32)
        lo = R.downscale(geom.scale, geom.edge_hi, conv, geom.size_lo,
                         parallel=True)
        res = R.residual(mock, err, None, lo, None, None, 1.0, parallel=True)
        R.residual_sum(res, True, parallel=True)
        done += 1
    dt = time.perf_counter() - t0
    del grid
    return dict(value=done / dt, unit='evals/s', cores=nthreads,
                kind='reference',
                sample=f"{done} parameter sets of the same workload, "
                       f"{dt:.1f} s; reference core (smdisk, downscale, "
                       f"residual, residual_sum) compiled from its sources "
                       f"with OpenMP, FFT convolution by scipy.fft (FFTW "
                       f"is absent)")


def smdisk_op_count(drv, model, sets):
    """On-disk spaxels V and channel deposits sum n_ch of the thin smdisk
    per parameter set, counted on the plain path's hi-res cube (the spaxels
    and voxels with a nonzero deposit; the rint windows are the
    reference's, gmodels.hpp:91-116), and the op count W = 60 V + 12 sum
    n_ch with each transcendental as 4 ops (means over `sets`)."""
    import torch
    dm, g = model.dmodel(), model.gmodel()
    c = dm.dcube()
    vs, ns = [], []
    for p in sets:
        model.model_d(p)                  # prepares the scratch cube
        drv.mem_fill(c.scratch_dcube(), 0)
        g.evaluate_scube(drv, p, c.scratch_dcube(), c.scratch_wcube(),
                         c.scratch_size(), c.scratch_step(),
                         c.scratch_zero(), c.rota(), c.dtype(), None)
        nz = c.scratch_dcube() != 0
        vs.append(int(nz.any(0).sum()))
        ns.append(int(nz.sum()))
    torch.cuda.synchronize()
    v, n = float(np.mean(vs)), float(np.mean(ns))
    return dict(v_disk=v, sum_nch=n, ops=60.0 * v + 12.0 * n,
                sets=len(sets))


def thick_op_count(drv, model, sets):
    """Voxels on the disk V and channel deposits sum n_ch of the thick
    smdisk per parameter set on its 3-D grid: V counted from the per-voxel
    velocity output (written for on-disk voxels only), n_ch from each
    voxel's rint channel window (gmodels.hpp:100-105, half-width 5 sigma,
    FP32 as on the device); W = 60 V + 12 sum n_ch (means over `sets`)."""
    import torch
    dm, g = model.dmodel(), model.gmodel()
    c = dm.dcube()
    model.model_d(sets[0])                # prepares the scratch cube
    spat_size, spat_step, spat_zero = g.spatial_grid(
        c.scratch_size(), c.scratch_step(), c.scratch_zero())
    nz_s, st_s, z0_s = c.scratch_size()[2], c.scratch_step()[2], \
        c.scratch_zero()[2]
    disk = g.components[0].disk
    be = drv.backends().gmodel(np.float32)
    shape = tuple(spat_size)[::-1]
    vd = drv.mem_alloc_d(shape, np.float32)
    dd = drv.mem_alloc_d(shape, np.float32)
    vs, ns = [], []
    for p in sets:
        drv.mem_fill(vd, np.nan)
        drv.mem_fill(dd, np.nan)
        be.smdisk_evaluate(*disk.native_args(drv, p), None, spat_size,
                           spat_step, spat_zero, c.rota(), nz_s, st_s, z0_s,
                           None, None, None, None, None, None, None, None,
                           vd, dd)
        v = vd.cpu().numpy()
        sd = dd.cpu().numpy()
        on = np.isfinite(v)
        v, sd = v[on], sd[on]
        f32 = np.float32
        k0 = np.maximum(np.rint((v - sd * f32(5) - f32(z0_s)) / f32(st_s)), 0)
        k1 = np.minimum(np.rint((v + sd * f32(5) - f32(z0_s)) / f32(st_s)),
                        nz_s - 1)
        vs.append(int(on.sum()))
        ns.append(int(np.maximum(k1 - k0 + 1, 0).sum()))
    torch.cuda.synchronize()
    v, n = float(np.mean(vs)), float(np.mean(ns))
    return dict(v_disk=v, sum_nch=n, ops=60.0 * v + 12.0 * n,
                sets=len(sets), grid=list(spat_size))


def algorithmic_bytes(ev, info, live=1.0, glive=1.0, seg=1.0):
    """Compulsory HBM bytes of each kernel (or stage): (per parameter set,
    shared per launch).  The observed cube is one array read by every set
    of a launch: it counts once per launch.  Direct pipeline: `live` is the
    fraction of x-pass tiles (row x channel chunk) that some spectral line
    reaches; the others are skipped, neither written by the x pass nor read
    by the y pass; `glive` the fraction of y-pass output groups with a
    nonzero input, the only ones the y pass writes and the z pass reads;
    `seg` the fraction of hi-res spaxels in 64-spaxel row segments some
    line reaches, the only line parameters written and read (1, 1, 1: the
    dense figures)."""
    nxl, nyl, nzl = ev.size_lo
    sx, sy, _ = ev.desc.scale
    X, Y = nxl * sx, nyl * sy
    n_lo = nxl * nyl * nzl
    data = {}
    if info['pipeline'] == 'direct':
        nx, ny, nz = info['fft_size']        # inputs that reach the outputs
        p1 = 4 * nz * ny * nxl * live        # x-filtered rows (live tiles)
        p2 = 4 * nz * nyl * nxl * glive      # x- and y-filtered columns
        masks = 8 * ny                       # one 64-bit mask per row
        per = {'sep_ypass': p1 + p2, 'sep_zpass': p2}
        data['sep_zpass'] = 4 * n_lo
        if ev.cube_source:                   # deposited hi-res cube
            cube = 4 * nx * ny * nz
            per['model_deposit'] = cube
            per['sep_xpass_cube'] = cube + p1 + masks
        else:
            lines = 16 * nx * ny * seg
            per['batch_lines_thin'] = lines + 2 * masks
            per['sep_xpass'] = lines + 2 * masks + p1
        return per, data
    if info['pipeline'] == 'fused':
        nx, ny, nz = info['fft_size']
        nzh = nz // 2 + 1
        plane = 8 * nzh * ny * nx            # one complex half spectrum
        data['fz_c2r_down_chi2'] = 4 * n_lo
        return {
            'batch_lines_thin': 16 * nx * ny,
            'fz_deposit_r2c': 16 * nx * ny + plane,
            'fx_lines (y)': 2 * plane,
            'fx_rows_conv': plane + 8 * nzh * ny * X,
            'fx_lines (inverse y)': 8 * nzh * ny * X + 8 * nzh * Y * X,
            'fz_c2r_down_chi2': 8 * nzh * Y * X,
        }, data
    n_pad = int(np.prod(ev.size_hi))
    data['downscale_chi2'] = 4 * n_lo
    return {'smdisk': 4 * n_pad, 'fft_convolution': 12 * n_pad,
            'downscale_chi2': 4 * X * Y * nzl}, data


def algorithmic_flops(ev, info, live=1.0, glive=1.0):
    """FP32 flops per parameter set of the direct pipeline's filter passes:
    one multiply-add per tap an output uses (the block-mean-folded kernel,
    s + 2 edge taps along each axis, without the negligible ends the
    pipeline drops), zero-padded taps not counted; the x and y passes only
    over their live tiles / groups (see algorithmic_bytes)."""
    if info['pipeline'] != 'direct':
        return {}
    nxl, nyl, nzl = ev.size_lo
    nx, ny, nz = info['fft_size']
    (tx, ty, tz), _ = ev.taps()
    return {'sep_xpass': 2 * nz * ny * nxl * tx * live,
            'sep_ypass': 2 * nz * nyl * nxl * ty * glive,
            'sep_zpass': 2 * nzl * nyl * nxl * tz}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summaries (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel
    calibration): a line 'traffic_bytes_per_launch <kernel>: <bytes>' in
    profiles/*_traffic.txt, or None when none covers it."""
    name = kernel.split('<')[0].split(' ')[0]
    d = os.path.join(ROOT, 'profiles')
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        if not fn.endswith('_traffic.txt'):
            continue
        with open(os.path.join(d, fn)) as f:
            for line in f:
                if line.startswith(f'traffic_bytes_per_launch {name}:'):
                    return float(line.split(':', 1)[1])
    return None


def batched_leg(drv, args, thick, nsets, steps, warmup, want_check,
                psf=None, path='auto', chunk=None):
    """The batched workload (a thin or thick smdisk on the 128x128x96 cube):
    `warmup` untimed steps, `steps` timed steps of `nsets` sets (graph
    replays), then the same steps launched directly with HIP events around
    every kernel for the per-kernel averages, their algorithmic bytes and
    flops and the roofline of the dominant kernel; with `want_check`, the
    GPU cubes and chi^2 of a seeded sample of sets for the per-set oracle
    errors (computed in the CPU leg)."""
    import torch
    from gbkfit_amd.batch import BatchEvaluator
    model = build_model(drv, thick, psf)
    arrays = workload(nsets, thick=thick)
    # mock data: the model of the first set plus seeded noise
    truth = param_dicts(arrays, 1)[0]
    cube0 = model.model_h(truth)['scube']['d'].astype(np.float32)
    rng = np.random.default_rng(2024)
    mock = (cube0 + rng.normal(0, 0.01, cube0.shape)).astype(np.float32)

    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock, error=0.01,
                        max_batch=chunk or args.chunk, template_params=truth,
                        path=path)
    if args.tiles and not thick:
        ev.set_tiles(*(int(v) for v in args.tiles.split(',')))
    use_graph = not args.no_graph

    def step():
        host = ev.packer.pack_arrays(arrays, nsets)
        ev.upload_packed(host)
        ev.run(nsets, use_graph=use_graph)
        return ev._chi2_dev.cpu().numpy()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        chi2 = step()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    evals = nsets * steps
    value = evals / t_batched

    # second timed region: the same steps launched directly with HIP events
    # around each stage (events on the kernels' own stream)
    ev.set_timing(True)
    step()
    torch.cuda.synchronize()
    stage_ms = np.zeros(3)
    kern_ms = {}
    chunks = 0
    t1 = time.perf_counter()
    for _ in range(steps):
        chi2_t = step()
        ms, chunks = ev.stage_times()
        stage_ms += ms
        for k, v in ev.kernel_times().items():
            kern_ms[k] = kern_ms.get(k, 0.0) + v
    torch.cuda.synchronize()
    t_timed = time.perf_counter() - t1
    ev.set_timing(False)
    # the two regions give the same chi^2, bit for bit
    assert np.array_equal(chi2_t, chi2)
    info = ev.info()

    # roofline of the dominant kernel: its algorithmic bytes per launch (the
    # compulsory reads and writes of its inputs and outputs for the sets of
    # one chunk) over its average launch time from the HIP events above
    launches = steps * max(chunks, 1)
    per_chunk = nsets / max(chunks, 1)
    # the last timed step ran one lane: its tile statistics are the last
    # chunk's (a chunk of sets of the workload)
    live_tiles = live_groups = live_blocks = None
    if info['pipeline'] == 'direct':
        (nzt, tot), (gz, gt), (bz, bt) = ev.tile_stats()
        live_tiles = nzt / tot if tot else None
        live_groups = gz / gt if gt else None
        live_blocks = bz / bt if bt else None
    live = live_tiles if live_tiles is not None else 1.0
    glive = live_groups if live_groups is not None else 1.0
    blive = live_blocks if live_blocks is not None else 1.0
    seg = None
    if info['pipeline'] == 'direct' and not ev.cube_source:
        sl, st = ev.line_stats()
        seg = sl / st if st else None
    algo, shared = algorithmic_bytes(ev, info, live, glive,
                                     seg if seg is not None else 1.0)
    aflops = algorithmic_flops(ev, info, live, glive)
    algo_dense, _ = algorithmic_bytes(ev, info)
    flops_dense = algorithmic_flops(ev, info)
    kernels = {}
    for k, t in kern_ms.items():
        avg_ms = t / launches
        b = algo.get(k)
        kernels[k] = dict(avg_ms_per_launch=avg_ms)
        if b is not None and avg_ms > 0:
            per_launch = b * per_chunk + shared.get(k, 0)
            gbs = per_launch / (avg_ms * 1e-3) / 1e9
            kernels[k].update(bytes_per_set=b, bytes_per_launch=per_launch,
                              achieved_gbs=gbs, frac=gbs / HBM_PEAK_GBS)
            if k in shared:
                kernels[k]['shared_bytes_per_launch'] = shared[k]
        if k in algo_dense and algo_dense[k] != b:
            kernels[k]['dense_bytes_per_set'] = algo_dense[k]
        if k in flops_dense and flops_dense[k] != aflops.get(k):
            kernels[k]['dense_flops_per_set'] = flops_dense[k]
        f = aflops.get(k)
        if f is not None and avg_ms > 0:
            tfs = f * per_chunk / (avg_ms * 1e-3) / 1e12
            kernels[k].update(flops_per_set=f, achieved_tflops=tfs,
                              frac_fp32=tfs / FP32_PEAK_TFLOPS)
    dom = max(kern_ms, key=kern_ms.get)
    dk = kernels[dom]
    traffic = measured_traffic(dom)
    # op roofline of the smdisk work: per-spaxel geometry and traits in
    # batch_lines_thin, the channel deposits in the x pass (which also runs
    # the x filter, so this understates the deposit's own rate)
    smdisk = None
    if info['pipeline'] == 'direct' and thick:
        smdisk = thick_op_count(drv, model, param_dicts(arrays, 4))
        t = kernels.get('model_deposit', {}).get('avg_ms_per_launch', 0)
        if t > 0:
            tfs = smdisk['ops'] * per_chunk / (t * 1e-3) / 1e12
            smdisk.update(kernels='batch_smdisk_thick (model_deposit)',
                          achieved_tflops=tfs, frac=tfs / FP32_PEAK_TFLOPS,
                          note='W = 60 V + 12 sum n_ch FP32 ops per set (a '
                               'transcendental counts 4), V the voxels of '
                               'the 3-D grid on the disk and n_ch their '
                               'channel windows, counted on 4 sets of the '
                               'workload; the time is the thick deposit '
                               'kernel per launch')
    if info['pipeline'] == 'direct' and not thick:
        smdisk = smdisk_op_count(drv, model, param_dicts(arrays, 16))
        t_ms = sum(kernels[k]['avg_ms_per_launch']
                   for k in ('batch_lines_thin', 'sep_xpass') if k in kernels)
        if t_ms > 0:
            tfs = smdisk['ops'] * per_chunk / (t_ms * 1e-3) / 1e12
            xf = aflops.get('sep_xpass', 0.0)
            both = (smdisk['ops'] + xf) * per_chunk / (t_ms * 1e-3) / 1e12
            smdisk.update(
                kernels='batch_lines_thin + sep_xpass',
                achieved_tflops=tfs, frac=tfs / FP32_PEAK_TFLOPS,
                with_x_filter_tflops=both,
                with_x_filter_frac=both / FP32_PEAK_TFLOPS,
                note='W = 60 V + 12 sum n_ch FP32 ops per set (a '
                     'transcendental counts 4), V and n_ch counted on the '
                     'hi-res cube of 16 sets of the workload; the time is '
                     'both kernels per launch, the x pass also '
                     'filtering (with_x_filter adds its useful flops)')
    # the binding roof of the dominant kernel: FP32 compute when its flop
    # rate is the larger fraction of peak, else HBM bandwidth
    if dk.get('frac_fp32', 0) > dk.get('frac', 0):
        roofline = dict(bound='mfma', achieved=dk['achieved_tflops'],
                        peak=FP32_PEAK_TFLOPS, unit='TFLOP/s',
                        frac=dk['frac_fp32'], traffic=traffic, kernel=dom,
                        note='FP32 multiply-adds on the vector ALUs '
                             '(packed FP32 disabled: at most half of the '
                             'FP32 peak is reachable); HBM: %.0f GB/s'
                             % dk['achieved_gbs'])
    else:
        achieved = dk.get('achieved_gbs')
        roofline = dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS,
                        unit='GB/s',
                        frac=(achieved / HBM_PEAK_GBS) if achieved else None,
                        traffic=traffic, kernel=dom)
    if live_tiles is not None:
        roofline['live_tile_fraction'] = live_tiles
        roofline['live_group_fraction'] = live_groups
        roofline['live_zblock_fraction'] = live_blocks
        roofline['live_line_fraction'] = seg
        roofline['note'] = (roofline.get('note', '') + '; ' if
                            'note' in roofline else '') + (
            'achieved counts the work actually done: the x-pass tiles some '
            'line reaches (live_tile_fraction), the y-pass output groups '
            'and z-pass blocks with a nonzero input (live_group_fraction, '
            'live_zblock_fraction), the line parameters of row segments '
            'some line reaches (live_line_fraction); the rest is zero and '
            'skipped; the observed cube counts once per launch')

    check = None
    if want_check:
        pick = np.sort(np.random.default_rng(7).choice(
            nsets, size=min(4, nsets), replace=False))
        csets = [param_dicts(arrays, nsets)[i] for i in pick]
        cc, cb = ev.evaluate(csets, want_cubes=True, use_graph=False)
        check = dict(index=pick, sets=csets, cubes=cb, chi2=cc, mock=mock,
                     same_as_step=bool(np.array_equal(cc, chi2[pick])))
    size_hi = list(ev.size_hi)
    ev.close()
    return dict(model=model, arrays=arrays, mock=mock, value=value,
                t_batched=t_batched, lat=lat, chi2=chi2, evals=evals,
                t_timed=t_timed, stage_ms=stage_ms, kernels=kernels,
                smdisk=smdisk, roofline=roofline, per_chunk=per_chunk,
                info=info, use_graph=use_graph, check=check,
                size_hi=size_hi)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--nsets', type=int, default=512)
    ap.add_argument('--chunk', type=int, default=256)
    ap.add_argument('--plain-sets', type=int, default=128)
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-graph', action='store_true')
    ap.add_argument('--thick', action='store_true',
                    help='the thick disk as the headline workload')
    ap.add_argument('--no-thick', action='store_true',
                    help='skip the thick variant leg')
    ap.add_argument('--serve-galaxies', type=int, default=64)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument('--serve-rate', type=float, default=1000.0)
    ap.add_argument('--no-serve', action='store_true')
    ap.add_argument('--ragged-galaxies', type=int, default=32)
    ap.add_argument('--ragged-reps', type=int, default=5)
    ap.add_argument('--no-ragged', action='store_true')
    ap.add_argument('--no-mcdisk', action='store_true')
    ap.add_argument('--no-fft', action='store_true',
                    help='skip the fused / padded FFT pipeline legs')
    ap.add_argument('--tiles', default=None,
                    help='direct pipeline tiles "lanes,ysplit" (default: '
                         'the evaluator\'s, 2,1)')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only copy a 1 GiB device buffer 3 times (known '
                         'bytes for scaling FETCH_SIZE / WRITE_SIZE)')
    ap.add_argument('--atomic-calibration', action='store_true',
                    help='only time 2^27 int64 atomic adds at random '
                         'addresses of a 256x256x128 cube (the roof of the '
                         'mcdisk deposits)')
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 28
        x = torch.ones(n, dtype=torch.float32, device='cuda')
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_bytes_read=4 * n, copy_bytes_written=4 * n,
                              launches=3)))
        return
    if args.atomic_calibration:
        import torch
        cells, n = 256 * 256 * 128, 1 << 27
        g = torch.Generator(device='cuda').manual_seed(5)
        idx = torch.randint(0, cells, (n,), device='cuda', generator=g)
        val = torch.ones(n, dtype=torch.int64, device='cuda')
        out = torch.zeros(cells, dtype=torch.int64, device='cuda')
        for _ in range(2):
            out.index_add_(0, idx, val)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            out.index_add_(0, idx, val)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 5
        print(json.dumps(dict(atomic_adds=n, cells=cells, ms=1e3 * dt,
                              atomics_per_s=n / dt)))
        return
    if args.gpus != 1:
        print(json.dumps(dict(error="this benchmark runs on one GPU; "
                              "multi-GPU runs are not part of this build")))
        sys.exit(2)

    import torch
    from gbkfit_amd.driver.hip import DriverHip
    from gbkfit_amd.objective import Data, Objective

    drv = DriverHip()
    head = batched_leg(drv, args, args.thick, args.nsets, args.steps,
                       args.warmup, not args.no_cpu)
    model, arrays, mock = head['model'], head['arrays'], head['mock']
    value, t_batched, chi2 = head['value'], head['t_batched'], head['chi2']
    evals, info = head['evals'], head['info']

    # plain drop-in path, one parameter set per call
    obj = Objective([dict(scube=Data(mock, error=np.full(mock.shape, 0.01,
                                                          np.float32)))],
                    [model])
    plain_sets = param_dicts(arrays, args.plain_sets)
    for p in plain_sets[:4]:
        obj.log_likelihood(p)
    torch.cuda.synchronize()
    plat = []
    tp = time.perf_counter()
    for p in plain_sets:
        ts = time.perf_counter()
        obj.log_likelihood(p)
        plat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - tp
    plain_value = len(plain_sets) / t_plain


    # the thick variant of the same workload (sech^2 height, spatial z 64):
    # a secondary object with its own kernel times, roofline and op count
    thick_variant = None
    if not args.thick and not args.no_thick:
        tsteps = max(2, args.steps // 2)
        tv = batched_leg(drv, args, True, args.nsets, tsteps, 1, False)
        thick_variant = dict(
            value=tv['value'], unit='evals/s', steps=tsteps,
            ms_per_step=1e3 * tv['t_batched'] / tsteps,
            workload='smdisk thick (sech^2 height 2 px, spatial z 64) '
                     '128x128x96, 2x spatial supersampling, PSF FWHM 3 px + '
                     'LSF FWHM 2 ch, chi^2 vs seeded mock; '
                     f'{args.nsets} parameter sets per step',
            stages_ms_per_step=dict(
                model=tv['stage_ms'][0] / tsteps,
                convolution=tv['stage_ms'][1] / tsteps,
                downscale_chi2=tv['stage_ms'][2] / tsteps),
            kernels=tv['kernels'], smdisk_ops=tv['smdisk'],
            roofline=dict(tv['roofline'], sets_per_launch=tv['per_chunk']))
        print(f"thick leg done: {tv['value']:.0f} evals/s", file=sys.stderr,
              flush=True)

    # the FFT pipelines, which take a PSF that does not factor (an elliptical,
    # rotated Gaussian): the same thin workload on each, 256 sets per step
    fft_pipelines = None
    if not args.thick and not args.no_fft:
        from gbkfit_amd import psflsf
        fft_pipelines = {}
        for path in ('fused', 'padded'):
            fp = batched_leg(drv, args, False, 256, 2, 1, False,
                             psf=psflsf.PSFGauss(3.0 * SIG, 0.7, 30.0),
                             path=path)
            assert fp['info']['pipeline'] == path
            fft_pipelines[path] = dict(
                value=fp['value'], unit='evals/s', steps=2,
                ms_per_step=1e3 * fp['t_batched'] / 2,
                fft_size=list(fp['info']['fft_size']),
                kernels=fp['kernels'],
                roofline=dict(fp['roofline'],
                              sets_per_launch=fp['per_chunk']))
            print(f"{path} FFT leg done: {fp['value']:.0f} evals/s",
                  file=sys.stderr, flush=True)
        fft_pipelines['workload'] = (
            'smdisk thin 128x128x96, 2x spatial supersampling, elliptical '
            'rotated PSF (FWHM 3 px, axis ratio 0.7, PA 30 deg: does not '
            'factor) + LSF FWHM 2 ch, chi^2 vs seeded mock; 256 parameter '
            'sets per step, one launch')

    serving = None
    if not args.no_serve:
        serving = serve_benchmark(drv, args.serve_galaxies,
                                  args.serve_seconds, args.serve_rate)
        print(f"serving leg done: {serving['throughput']:.0f} req/s",
              file=sys.stderr, flush=True)
    mcdisk = None
    if not args.no_mcdisk:
        mcdisk = mcdisk_benchmark(drv, 10)
        print(f"mcdisk leg done: {mcdisk['ms_per_eval']:.1f} ms per eval",
              file=sys.stderr, flush=True)
    ragged = None
    if not args.no_ragged:
        ragged = ragged_benchmark(drv, args.ragged_galaxies,
                                  args.ragged_reps)
        print(f"ragged leg done: {ragged['ms_per_batch']:.1f} ms per batch",
              file=sys.stderr, flush=True)

    cpu = None
    oracle_err = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_seconds, args.thick)
        except Exception as e:  # the harness may be absent
            cpu = dict(error=str(e))
        try:
            check = head['check']
            oracle_err = dict(sets=oracle_errors(check, args.thick),
                              chi2_equal_to_the_step=check['same_as_step'])
        except Exception as e:  # the oracle library may be absent
            oracle_err = dict(error=str(e))

    out = dict(
        metric=METRIC, value=value, unit='evals/s', n_gpus=1,
        steps=args.steps, warmup=args.warmup,
        ms_per_step=1e3 * t_batched / args.steps, higher_is_better=True,
        scaling='weak', vs_baseline=None, dtype='fp32', data='synthetic',
        config=dict(workload=(
            f"smdisk {'thick' if args.thick else 'thin'} 128x128x96, 2x "
            f"spatial supersampling, PSF FWHM 3 px + LSF FWHM 2 ch, chi^2 vs "
            f"seeded mock; {args.nsets} parameter sets per step"),
            nsets=args.nsets, chunk=args.chunk, graph=head['use_graph'],
            padding='reference', dcube_size_hi=head['size_hi'],
            pipeline=dict(direct='direct separable convolution',
                          fused='spectral-first fused FFT',
                          padded='padded-cube FFT')[info['pipeline']],
            fft_size=list(info['fft_size'])),
        plain=dict(value=plain_value, unit='evals/s',
                   sets=len(plain_sets), **percentiles(plat)),
        batched_over_plain=value / plain_value,
        batch_latency=percentiles(head['lat']),
        per_eval_latency_batched_ms=1e3 * t_batched / evals,
        direct_launch_evals_per_s=evals / head['t_timed'],
        stages_ms_per_step=dict(
            model=head['stage_ms'][0] / args.steps,
            convolution=head['stage_ms'][1] / args.steps,
            downscale_chi2=head['stage_ms'][2] / args.steps),
        kernels=head['kernels'],
        smdisk_ops=head['smdisk'],
        roofline=dict(head['roofline'], sets_per_launch=head['per_chunk']),
        thick_variant=thick_variant,
        fft_pipelines=fft_pipelines,
        serving=serving,
        ragged=ragged,
        mcdisk_plain=mcdisk,
        cpu_baseline=cpu,
        oracle_error_per_set=oracle_err,
        host=dict(cpu=platform.processor() or platform.machine(),
                  nproc=os.cpu_count()),
        chi2_sample=[float(c) for c in chi2[:3]],
    )
    print(json.dumps(out))


if __name__ == '__main__':
    main()
