"""Writes tests/golden/*.npz from the reference's own header-only core.

Run in the build container only (it needs /root/reference):
    make -C oracle && python tests/golden/make_golden.py

The reference core is compiled by oracle/Makefile into oracle/_ref/libgbkref.so
(git-ignored); this script calls it on the cases of cases.py and stores the
inputs' generator metadata and the outputs.  FP32 outputs come from the core
instantiated with T=float (what the reference host driver runs), FP64 outputs
from T=double.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import oracle  # noqa: E402
from oracle import refcore  # noqa: E402
import cases  # noqa: E402

EXTRA = ('image', 'rdata', 'vdata_cmp', 'ddata_cmp')
F64_CASES = ('thin_basic', 'thick_sech2', 'thin_loose_tilted')


def smdisk_fixtures():
    out = {}
    for case in cases.SMDISK_CASES:
        spec = cases.build_spec(case)
        grid = cases.build_grid(case)
        want = ('scube',) + (EXTRA if case['name'] in ('thin_basic',
                                                       'thick_sech2') else ())
        r32 = refcore.ref(np.float32).smdisk(spec.rounded_f32(),
                                             oracle.grid_f32(grid), want)
        for k in want:
            out[f"{case['name']}/{k}/f32"] = r32[k]
        if case['name'] in F64_CASES:
            r64 = refcore.ref(np.float64).smdisk(spec, grid, ('scube',))
            out[f"{case['name']}/scube/f64"] = r64['scube']
    return out


def extra_fixtures():
    """Opacity and weight cases (FP32 core, every output they exercise) and
    the spectral weight cube of the thick weight case's spatial weights."""
    out = {}
    for case in cases.EXTRA_CASES:
        want = cases.EXTRA_WANT[case['name']]
        spec = cases.build_spec(case).rounded_f32()
        r32 = refcore.ref(np.float32).smdisk(spec, oracle.grid_f32(
            cases.build_grid(case)), want)
        for k in want:
            out[f"{case['name']}/{k}/f32"] = r32[k]
    w = out['thick_weights/wdata/f32']
    for dt, sfx in ((np.float32, 'f32'), (np.float64, 'f64')):
        out[f'wcube/spec/{sfx}'] = refcore.ref(dt).wcube(w, 24)
    return out


def mcdisk_fixtures():
    out = {}
    for case in cases.MCDISK_CASES:
        spec = cases.build_spec(case)
        grid = cases.build_grid(case)
        cflux, csum, hoi = cases.cloud_budget(spec, case['traits'],
                                              case['nclouds'])
        r = refcore.ref(np.float64).mcdisk(spec, grid, cflux, csum, hoi, 42,
                                           want=('scube', 'rdata'))
        out[f"{case['name']}/scube/f64"] = r['scube'].astype(np.float32)
        out[f"{case['name']}/rdata/f64"] = r['rdata'].astype(np.float32)
        out[f"{case['name']}/csum"] = csum
        out[f"{case['name']}/cflux"] = np.array(cflux)
    return out


def dmodel_fixtures():
    rng = np.random.default_rng(7)
    out = {}
    src = rng.standard_normal((12, 18, 20)).astype(np.float32)
    out['downscale/src'] = src
    for dt, sfx in ((np.float32, 'f32'), (np.float64, 'f64')):
        R = refcore.ref(dt)
        out[f'downscale/dst/{sfx}'] = R.downscale((2, 3, 2), (1, 2, 1), src,
                                                  (9, 5, 5))
        d, m = R.mask(0.5, True, src)
        out[f'mask/d/{sfx}'], out[f'mask/m/{sfx}'] = d, m
    # moments of a positive line cube (+ negative pixels and masked spaxels)
    nz, ny, nx = 24, 6, 7
    z = np.arange(nz)[:, None, None]
    cube = np.exp(-0.5 * ((z - 10 - rng.uniform(-3, 3, (ny, nx))) /
                          rng.uniform(1, 4, (ny, nx))) ** 2)
    cube *= rng.uniform(0.5, 2.0, (ny, nx))
    cube[:, 0, 0] = 0
    cube[:, 1, 1] = 1e-9
    cube = cube.astype(np.float32)
    wcube = rng.uniform(0, 1, cube.shape).astype(np.float32)
    out['moments/cube'] = cube
    out['moments/wcube'] = wcube
    step, zero, orders = (1.0, 1.0, 10.0), (0.0, 0.0, -120.0), (0, 1, 2, 3, 4)
    for dt, sfx in ((np.float32, 'f32'), (np.float64, 'f64')):
        R = refcore.ref(dt)
        md, mm, mw = R.moments(cube, step, zero, orders, 1e-6, wcube)
        out[f'moments/d/{sfx}'], out[f'moments/m/{sfx}'] = md, mm
        out[f'moments/w/{sfx}'] = mw
        md1, mm1, _ = R.moments(cube, step, zero, (1,), 1e-6)
        out[f'moments1/d/{sfx}'], out[f'moments1/m/{sfx}'] = md1, mm1
    # residual with every optional array present, and the plain form
    n = 1000
    od, md_ = rng.standard_normal(n), rng.standard_normal(n)
    oe = rng.uniform(0.5, 2, n)
    om = (rng.uniform(0, 1, n) > 0.2).astype(float)
    mw_ = rng.uniform(0, 1, n)
    mm_ = (rng.uniform(0, 1, n) > 0.1).astype(float)
    for k, v in dict(od=od, oe=oe, om=om, md=md_, mw=mw_, mm=mm_).items():
        out[f'residual/{k}'] = v.astype(np.float32)
    for dt, sfx in ((np.float32, 'f32'), (np.float64, 'f64')):
        R = refcore.ref(dt)
        a = [out[f'residual/{k}'] for k in ('od', 'oe', 'om', 'md', 'mw', 'mm')]
        full = R.residual(*a, 0.7)
        plain = R.residual(a[0], None, None, a[3], None, None, 1.0)
        out[f'residual/full/{sfx}'] = full
        out[f'residual/plain/{sfx}'] = plain
        out[f'residual_sum/sq/{sfx}'] = np.array(R.residual_sum(full, True))
        out[f'residual_sum/abs/{sfx}'] = np.array(R.residual_sum(full, False))
    return out


def main():
    if not refcore.available():
        raise SystemExit("reference-core harness missing: make -C oracle")
    fx = {}
    fx.update(smdisk_fixtures())
    fx.update(extra_fixtures())
    fx.update(mcdisk_fixtures())
    fx.update(dmodel_fixtures())
    meta = dict(
        generator='tests/golden/make_golden.py',
        harness='oracle/ref_harness.cpp + reference core headers '
                '(src/gbkfit/driver/native/gbkfit/include, '
                'gbkfit_host/include/gbkfit/host/random.hpp)',
        flags='g++ -O2 -std=c++17 -ffp-contract=off -fopenmp',
        mcdisk_rng='std::mt19937 seed 42 (host/include/gbkfit/host/random.hpp)')
    fx['meta'] = np.array(json.dumps(meta))
    path = os.path.join(HERE, 'golden_ref.npz')
    np.savez_compressed(path, **fx)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB, "
          f"{len(fx)} arrays)")


if __name__ == '__main__':
    main()
