"""gbkhip_mcdisk_evaluate_noalloc (the mcdisk evaluation that allocates
nothing, its deposit buffers in a fixed device array) against
gbkhip_mcdisk_evaluate through the C ABI: the same cube bit for bit on two
calls in a row, the second with fewer clouds and other parameters, and
requests beyond the fixed scratch or outside the binned deposit rejected
before anything is enqueued."""
import ctypes

import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu


def _setup(posa, incl, nclouds):
    """A thick exponential x sech^2 disk on the 64x64x64 cube at 2x
    (padded hi-res grid 192x192x96, spatial z 16) and its cloud budget."""
    p = dict(P.c1_params(), thin=False, posa=posa, incl=incl)
    spec = P.oracle_spec(p)
    grid = oracle.grid_f32(oracle.grid_for(P.oracle_geom(p), False, 16))
    sub = spec.sub[1:-1]
    rsep = sub[1] - sub[0]
    ring = p['bpt_a'] * np.exp(-sub / p['bpt_s']) * np.pi * (
        (sub + rsep / 2) ** 2 - (sub - rsep / 2) ** 2)
    cflux = float(ring.sum() / nclouds)
    csum = np.cumsum((ring / cflux).astype(np.int32)).astype(np.int32)
    return spec, grid, cflux, csum


def _evaluate(drv, name, spec, grid, cflux, csum, seed, image=False):
    from gbkfit_amd import _native as N
    from gbkfit_amd.driver.hip import (_ptr, make_disk, make_grid,
                                       make_outputs)
    a = H.disk_args(drv, spec)
    d = make_disk(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8:53],
                  a[53])
    g = make_grid(*H.grid_args(grid)[:3], *H.grid_args(grid)[4:])
    nx, ny, _ = grid.spat_size
    scube = drv.mem_alloc_d((grid.spec_size, ny, nx), np.float32)
    drv.mem_fill(scube, 0.0)
    extra = {}
    if image:
        extra['image'] = drv.mem_alloc_d((ny, nx), np.float32)
        drv.mem_fill(extra['image'], 0.0)
    o = make_outputs(scube=scube, **extra)
    csum_d = H.dev(drv, csum, np.int32)
    hoi_d = H.dev(drv, np.zeros(1, np.uint8), np.uint8)
    try:
        N.call(name, drv.context().handle, ctypes.c_float(cflux),
               int(csum[-1]), _ptr(csum_d), int(csum.size), _ptr(hoi_d),
               ctypes.byref(d), ctypes.byref(g), ctypes.byref(o),
               ctypes.c_uint64(seed))
    finally:
        out = scube.cpu().numpy()     # waits for the call's work
    del a, csum_d, hoi_d
    return out


def test_noalloc_equals_mcdisk_evaluate_on_two_calls():
    drv = H.driver()
    first = _setup(30.0, 50.0, 1_000_000)
    second = _setup(140.0, 65.0, 200_000)
    ref1 = _evaluate(drv, 'gbkhip_mcdisk_evaluate', *first, seed=11)
    ref2 = _evaluate(drv, 'gbkhip_mcdisk_evaluate', *second, seed=12)
    got1 = _evaluate(drv, 'gbkhip_mcdisk_evaluate_noalloc', *first, seed=11)
    got2 = _evaluate(drv, 'gbkhip_mcdisk_evaluate_noalloc', *second, seed=12)
    print('clouds', first[3][-1], second[3][-1], 'flux', ref1.sum(),
          ref2.sum())
    assert ref1.sum() > 0 and ref2.sum() > 0
    assert np.array_equal(got1, ref1)
    assert np.array_equal(got2, ref2)


def test_noalloc_rejects_requests_it_cannot_serve():
    from gbkfit_amd import _native as N
    drv = H.driver()
    big = _setup(30.0, 50.0, 3_000_000)
    with pytest.raises(N.GbkHipError, match='scratch'):
        _evaluate(drv, 'gbkhip_mcdisk_evaluate_noalloc', *big, seed=1)
    small = _setup(30.0, 50.0, 100_000)
    with pytest.raises(N.GbkHipError, match='binned'):
        _evaluate(drv, 'gbkhip_mcdisk_evaluate_noalloc', *small, seed=1,
                  image=True)
    # a rejected call enqueued nothing: the next one is still exact
    got = _evaluate(drv, 'gbkhip_mcdisk_evaluate_noalloc', *small, seed=3)
    ref = _evaluate(drv, 'gbkhip_mcdisk_evaluate', *small, seed=3)
    assert np.array_equal(got, ref)
