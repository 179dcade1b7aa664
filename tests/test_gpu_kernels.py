"""GPU parity of each native entry point against the FP64 oracle, called
through the hip driver's backends (the reference's Python signatures) and the
C ABI underneath.  Tolerances are stated per test; the cube comparisons use
|gpu - oracle| <= tol * max|oracle| + amb, where amb is the oracle's bound on
contributions decided by FP32 near-ties (see oracle/oracle.h)."""
import numpy as np
import pytest

import cases
import hiputil as H
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope='module')
def drv():
    return H.driver()


@pytest.mark.parametrize('case', cases.SMDISK_CASES, ids=lambda c: c['name'])
def test_smdisk_cases(drv, case):
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    ref, _ = oracle.smdisk(spec, grid, want=('scube', 'amb'))
    got = H.smdisk(drv, spec, grid)
    nbad, err = H.within(got['scube'], ref['scube'], TOL, ref['amb'])
    print(case['name'], 'normwise', err, 'violations', nbad)
    assert nbad == 0


@pytest.mark.parametrize('name', ['thick_sech2', 'thick_heights',
                                  'thick_loose_tilted'])
def test_thick_smdisk_reruns_bit_identical(drv, name):
    """Thick disks: the waves that split a line of sight add into their own
    partial spectra and image rows, summed in a fixed order, so reruns give
    the same cube and image bit for bit (and each stays within the oracle's
    bound)."""
    case = [c for c in cases.SMDISK_CASES if c['name'] == name][0]
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    a = H.smdisk(drv, spec, grid, want=('scube', 'image'))
    b = H.smdisk(drv, spec, grid, want=('scube', 'image'))
    assert a['scube'].any() and a['image'].any()
    assert np.array_equal(a['scube'], b['scube'])
    assert np.array_equal(a['image'], b['image'])
    ref, _ = oracle.smdisk(spec, grid, want=('image',))
    assert H.normwise(a['image'], ref['image']) < TOL


@pytest.mark.parametrize('case', cases.EXTRA_CASES, ids=lambda c: c['name'])
def test_opacity_and_weights_match_oracle(drv, case):
    """smdisk with an opacity cube (thick disk) or weight traits (wdata),
    every output of the case against the FP64 oracle: the spectral cube
    within 1e-5 max + amb, the other outputs 1e-5 normwise (the oracle is
    pinned to the reference core on the same cases)."""
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    want = cases.EXTRA_WANT[case['name']]
    ref, _ = oracle.smdisk(spec, grid, want=want + ('amb',))
    got = H.smdisk(drv, spec, grid, want=want)
    for k in want:
        if k == 'scube':
            nbad, err = H.within(got[k], ref[k], TOL, ref['amb'])
            print(case['name'], k, 'normwise', err, 'violations', nbad)
            assert nbad == 0
        else:
            assert H.normwise(got[k], ref[k]) < TOL, (case['name'], k)


def test_wcube_matches_oracle(drv):
    """gbkhip_wcube_evaluate (gmodels.hpp:52-82) on the thick weight
    case's spatial weights: mean / max of each column on every channel."""
    case = [c for c in cases.EXTRA_CASES if c['name'] == 'thick_weights'][0]
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    w = oracle.smdisk(spec, grid, want=('wdata',))[0]['wdata']
    be = drv.backends().gmodel(np.float32)
    nz, ny, nx = w.shape
    out = drv.mem_alloc_d((24, ny, nx), np.float32)
    be.wcube_evaluate((nx, ny, nz), 24, H.dev(drv, w), out)
    ref = oracle.wcube(w, 24)
    assert ref.max() > 0
    assert H.normwise(out.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize('name', ['thin_basic', 'thin_profiles_a'])
def test_smdisk_deposit_counts_match_oracle(drv, name):
    """The smdisk op count bench.py reports (spaxels with a line and voxels
    with a nonzero deposit on the hi-res cube, 60 V + 12 sum n_ch) equals
    the FP64 oracle's wherever no FP32 near-tie (amb) can move a deposit."""
    case = [c for c in cases.SMDISK_CASES if c['name'] == name][0]
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    ref, _ = oracle.smdisk(spec, grid, want=('scube', 'amb'))
    got = H.smdisk(drv, spec, grid)['scube']
    sure = ref['amb'] == 0
    assert sure.mean() > 0.9
    assert np.array_equal((got != 0)[sure], (ref['scube'] != 0)[sure])
    col_sure = sure.all(axis=0)
    assert np.array_equal((got != 0).any(axis=0)[col_sure],
                          (ref['scube'] != 0).any(axis=0)[col_sure])
    print(name, 'V', int((got != 0).any(axis=0).sum()), 'sum n_ch',
          int((got != 0).sum()))


@pytest.mark.parametrize('name', ['thin_basic', 'thick_sech2'])
def test_smdisk_side_outputs(drv, name):
    case = [c for c in cases.SMDISK_CASES if c['name'] == name][0]
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    want = ('image', 'rdata', 'vdata_cmp', 'ddata_cmp')
    ref, _ = oracle.smdisk(spec, grid, want=want)
    got = H.smdisk(drv, spec, grid, want=want,
                   fill=dict(vdata_cmp=np.nan, ddata_cmp=np.nan))
    for k in ('image', 'rdata'):
        assert H.normwise(got[k], ref[k]) < TOL, k
    # velocity / dispersion maps: compare where the oracle wrote a value
    for k in ('vdata_cmp', 'ddata_cmp'):
        m = ref[k] != 0
        assert H.normwise(got[k][m], ref[k][m]) < TOL, k


def _c1_like(size=(32, 32, 32), thin=True, psf=1.0617, lsf=8.493):
    geom = oracle.DCubeGeom(size, (1, 1, 10), scale=(2, 2, 1),
                            psf_sigma=psf, lsf_sigma=lsf)
    T = oracle.Trait
    rn = np.arange(0, 14.0, 1.5)
    traits = dict(rp=[T(2, params=(1.0, 4.0))],
                  vp=[T(2, params=(3.0, 200.0))],
                  dp=[T(1, params=(20.0,))])
    if not thin:
        traits['rh'] = [T(7, consts=(0, 0), params=(1.0,))]
        traits['vh'] = [T(1, consts=(0,))]
        traits['dh'] = [T(1, consts=(0,))]
    spec = oracle.DiskSpec(rn, traits, posa=30, incl=45).rounded_f32()
    return geom, spec


def test_fft_r2c_c2r(drv):
    be = drv.backends().fft(np.float32)
    rng = np.random.default_rng(3)
    for shape in [(4, 6, 8), (5, 7, 9), (1, 1, 16), (3, 1, 34), (2, 17, 64),
                  (8, 12, 40), (1, 4, 1)]:
        a = rng.standard_normal(shape).astype(np.float32)
        d = H.dev(drv, a)
        c = drv.mem_alloc_d(be.fft_complex_shape(list(shape)), np.complex64)
        be.fft_r2c(d, c)
        ref = np.fft.rfftn(a.astype(np.float64))
        got = c.cpu().numpy()
        assert H.normwise(got, ref) < 1e-5, shape
        back = drv.mem_alloc_d(shape, np.float32)
        be.fft_c2r(c, back)
        n = np.prod(shape)
        assert H.normwise(back.cpu().numpy() / n, a) < 1e-5, shape
        # c2r leaves its input untouched
        assert np.array_equal(c.cpu().numpy(), got)


@pytest.mark.parametrize('shape', [(128, 48, 320), (30, 300, 50),
                                   (9, 40, 1024), (2, 1536, 6),
                                   (3, 2048, 4), (5, 17, 34), (7, 96, 192),
                                   (7, 224, 448), (11, 26, 352),
                                   (13, 77, 416), (4, 91, 286)])
def test_fft_sizes_all_radices(drv, shape):
    """Lengths that exercise every radix (16, 8, 4, 2, 3, 5, 7, 11, 13),
    every lines-per-block and register bucket of the tiled engine, and the
    generic fallback (2048 lines, prime 17)."""
    be = drv.backends().fft(np.float32)
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape).astype(np.float32)
    k = rng.standard_normal(shape).astype(np.float32)
    d = H.dev(drv, a)
    c = drv.mem_alloc_d(be.fft_complex_shape(list(shape)), np.complex64)
    be.fft_r2c(d, c)
    assert H.normwise(c.cpu().numpy(), np.fft.rfftn(a.astype(np.float64))) \
        < 1e-5, shape
    back = drv.mem_alloc_d(shape, np.float32)
    be.fft_c2r(c, back)
    assert H.normwise(back.cpu().numpy() / a.size, a) < 1e-5, shape
    dk = H.dev(drv, k)
    be.fft_convolve_cached(d, dk)
    assert H.normwise(d.cpu().numpy(), oracle.fft_convolve(a, k)) < 1e-5, shape


def test_fft_convolve_matches_closed_form(drv):
    be = drv.backends().fft(np.float32)
    rng = np.random.default_rng(5)
    for shape in [(8, 8, 8), (16, 20, 30), (6, 1, 10), (1, 1, 12),
                  (32, 32, 34)]:
        a = rng.standard_normal(shape).astype(np.float32)
        k = rng.standard_normal(shape).astype(np.float32)
        da, dk = H.dev(drv, a), H.dev(drv, k)
        be.fft_convolve_cached(da, dk)
        ref = oracle.fft_convolve(a, k)
        assert H.normwise(da.cpu().numpy(), ref) < 1e-5, shape
        # the explicit-spectrum form
        da2 = H.dev(drv, a)
        kc = drv.mem_alloc_d(be.fft_complex_shape(list(shape)), np.complex64)
        wc = drv.mem_alloc_d(be.fft_complex_shape(list(shape)), np.complex64)
        be.fft_r2c(dk, kc)
        be.fft_convolve(da2, wc, kc)
        assert H.normwise(da2.cpu().numpy(), ref) < 1e-5, shape


@pytest.mark.parametrize('shape,sig', [
    ((24, 36, 40), (1.2, 1.5, 0.9)),       # direct passes, wrap-around
    ((16, 64, 320), (3.0, 2.5, 0.7)),      # wide x taps (bucket 64)
    ((10, 12, 9), (0.8, 0.8, 0.8)),        # support wider than x: FFT
])
def test_fft_convolve_cached_separable_kernel(drv, shape, sig):
    """A kernel cube that factors into short 1-D kernels, as DCube's PSF x
    LSF cube does, takes three direct circular passes instead of the FFTs
    (the third case, whose support does not fit the x axis, the FFTs): the
    result is the circular convolution (FP64 closed form) within 1e-6,
    across the cube's faces too."""
    be = drv.backends().fft(np.float32)
    rng = np.random.default_rng(sum(shape))
    nz, ny, nx = shape

    def g1(n, s):
        d = np.minimum(np.arange(n), n - np.arange(n)).astype(np.float64)
        v = np.exp(-0.5 * (d / s) ** 2)
        v[d > 8 * s] = 0.0
        return v
    k = (g1(nz, sig[2])[:, None, None] * g1(ny, sig[1])[None, :, None] *
         g1(nx, sig[0])[None, None, :])
    k = (k / k.sum()).astype(np.float32)
    a = rng.standard_normal(shape).astype(np.float32)
    da, dk = H.dev(drv, a), H.dev(drv, k)
    be.fft_convolve_cached(da, dk)
    ref = oracle.fft_convolve(a, k)
    assert H.normwise(da.cpu().numpy(), ref) < 1e-6, shape
    # the same kernel again (cached analysis), on another cube
    b = rng.standard_normal(shape).astype(np.float32)
    db = H.dev(drv, b)
    be.fft_convolve_cached(db, dk)
    assert H.normwise(db.cpu().numpy(), oracle.fft_convolve(b, k)) < 1e-6


def test_fft_convolve_cached_invalidation(drv):
    be = drv.backends().fft(np.float32)
    rng = np.random.default_rng(9)
    shape = (8, 10, 12)
    a = rng.standard_normal(shape).astype(np.float32)
    k1 = rng.standard_normal(shape).astype(np.float32)
    k2 = rng.standard_normal(shape).astype(np.float32)
    dk = H.dev(drv, k1)
    da = H.dev(drv, a)
    be.fft_convolve_cached(da, dk)
    dk.copy_(H.dev(drv, k2))          # rewritten in place
    da = H.dev(drv, a)
    be.fft_convolve_cached(da, dk)
    assert H.normwise(da.cpu().numpy(), oracle.fft_convolve(a, k2)) < 1e-5


def test_fft_brute_force_tiny(drv):
    be = drv.backends().fft(np.float32)
    rng = np.random.default_rng(11)
    shape = (3, 5, 6)
    a = rng.standard_normal(shape).astype(np.float32)
    k = rng.standard_normal(shape).astype(np.float32)
    ref = oracle.circular_convolve(a, k)
    da, dk = H.dev(drv, a), H.dev(drv, k)
    be.fft_convolve_cached(da, dk)
    assert H.normwise(da.cpu().numpy(), ref) < 1e-5


def test_downscale_mask(drv):
    be = drv.backends().dmodel(np.float32)
    rng = np.random.default_rng(1)
    src = rng.standard_normal((20, 30, 40)).astype(np.float32)
    dsrc = H.dev(drv, src)
    dst = drv.mem_alloc_d((8, 12, 17), np.float32)
    be.dcube_downscale((2, 2, 2), (3, 2, 1), dsrc, dst)
    ref = oracle.downscale((2, 2, 2), (3, 2, 1), src, (17, 12, 8))
    assert H.normwise(dst.cpu().numpy(), ref) < 1e-6
    m = drv.mem_alloc_d(src.shape, np.float32)
    be.dcube_mask(0.3, True, m, dsrc, None)
    rd, rm = oracle.mask(0.3, True, src)
    got = dsrc.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(rd))
    assert np.array_equal(got[~np.isnan(got)], src[~np.isnan(rd)])
    assert np.array_equal(m.cpu().numpy(), rm)


def test_moments(drv):
    fx = np.load(cases.__file__.replace('cases.py', 'golden_ref.npz'))
    cube, wcube = fx['moments/cube'], fx['moments/wcube']
    be = drv.backends().dmodel(np.float32)
    orders = np.array([0, 1, 2, 3, 4], np.int32)
    nz, ny, nx = cube.shape
    md = drv.mem_alloc_d((5, ny, nx), np.float32)
    mm = drv.mem_alloc_d((ny, nx), np.float32)
    mw = drv.mem_alloc_d((5, ny, nx), np.float32)
    drv.mem_fill(md, np.nan)
    drv.mem_fill(mw, 1)
    be.mmaps_moments((nx, ny, nz), (1, 1, 10), (0, 0, -120), H.dev(drv, cube),
                     H.dev(drv, wcube), 1e-6, H.dev(drv, orders, np.int32),
                     md, mw, mm)
    rd, rm, rw = oracle.moments(cube, (1, 1, 10), (0, 0, -120), orders, 1e-6,
                                wcube)
    got_d = md.cpu().numpy()
    assert np.array_equal(np.isnan(got_d), np.isnan(rd))
    for i in range(5):
        assert H.normwise(got_d[i], rd[i]) < 1e-5, i
    assert np.array_equal(mm.cpu().numpy(), rm)
    assert H.normwise(mw.cpu().numpy()[0], rw[0]) < 1e-5


def test_moments_batched_equal_one_by_one(drv):
    """gbkhip_mmaps_moments_batched: the maps of three cubes in one launch
    equal gbkhip_mmaps_moments' of each cube, bit for bit."""
    import torch
    from gbkfit_amd import _native as N
    fx = np.load(cases.__file__.replace('cases.py', 'golden_ref.npz'))
    base = fx['moments/cube']
    nz, ny, nx = base.shape
    rng = np.random.default_rng(4)
    cubes = np.stack([base, base[:, ::-1, :] * 0.5,
                      base + rng.normal(0, 1e-3, base.shape)]).astype(
                          np.float32)
    orders = np.array([0, 1, 2, 4], np.int32)
    be = drv.backends().dmodel(np.float32)
    dev = H.dev(drv, cubes)
    od = H.dev(drv, orders, np.int32)
    md = drv.mem_alloc_d((3, 4, ny, nx), np.float32)
    mm = drv.mem_alloc_d((3, ny, nx), np.float32)
    N.call('gbkhip_mmaps_moments_batched', drv.context().handle,
           N._i3((nx, ny, nz)), N._f3((1, 1, 10)), N._f3((0, 0, -120)),
           dev.data_ptr(), 1e-6, 4, od.data_ptr(), md.data_ptr(),
           mm.data_ptr(), 3)
    for i in range(3):
        d1 = drv.mem_alloc_d((4, ny, nx), np.float32)
        m1 = drv.mem_alloc_d((ny, nx), np.float32)
        be.mmaps_moments((nx, ny, nz), (1, 1, 10), (0, 0, -120), dev[i],
                         None, 1e-6, od, d1, None, m1)
        torch.cuda.synchronize()
        assert np.array_equal(md[i].cpu().numpy(), d1.cpu().numpy(),
                              equal_nan=True), i
        assert np.array_equal(mm[i].cpu().numpy(), m1.cpu().numpy()), i


def test_residual_and_sum(drv):
    be = drv.backends().objective(np.float32)
    rng = np.random.default_rng(2)
    n = 100003
    arrs = [rng.standard_normal(n).astype(np.float32) for _ in range(2)]
    oe = rng.uniform(0.5, 2, n).astype(np.float32)
    om = (rng.uniform(0, 1, n) > 0.2).astype(np.float32)
    mw = rng.uniform(0, 1, n).astype(np.float32)
    mm = (rng.uniform(0, 1, n) > 0.1).astype(np.float32)
    # masked pixels carry NaN data / NaN model values: zero residual there
    arrs[0][om == 0] = np.nan
    arrs[1][:50][mm[:50] == 0] = np.nan
    res = drv.mem_alloc_d(n, np.float32)
    be.residual(H.dev(drv, arrs[0]), H.dev(drv, oe), H.dev(drv, om),
                H.dev(drv, arrs[1]), H.dev(drv, mw), H.dev(drv, mm), 0.7, res)
    ref = oracle.residual(arrs[0], oe, om, arrs[1], mw, mm, 0.7)
    assert np.all(np.isfinite(ref)) and np.all(ref[om == 0] == 0)
    assert np.array_equal(res.cpu().numpy() == 0, ref == 0)
    assert H.normwise(res.cpu().numpy(), ref) < 1e-6
    s = drv.mem_alloc_d(1, np.float32)
    be.residual_sum(res, True, s)
    rs = oracle.residual_sum(res.cpu().numpy().astype(np.float64), True)
    assert abs(float(s.cpu()[0]) - rs) / rs < 1e-6
    be.residual_sum(res, False, s)
    ra = oracle.residual_sum(res.cpu().numpy().astype(np.float64), False)
    assert abs(float(s.cpu()[0]) - ra) / ra < 1e-6


@pytest.mark.parametrize('case', cases.MCDISK_CASES, ids=lambda c: c['name'])
def test_mcdisk_matches_oracle(drv, case):
    """Same Philox4x32-10 cloud stream (keyed by seed and cloud id) on GPU
    and in the FP64 oracle: every cloud lands in the same voxel unless its
    position is an FP32 near-tie (covered by amb), and the int64 fixed-point
    sum differs from the FP64 one by FP32 rounding of the deposits."""
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    cflux, csum, hoi = cases.cloud_budget(spec, case['traits'],
                                          case['nclouds'])
    ref, _ = oracle.mcdisk(spec, grid, cflux, csum, hoi, seed=1234)
    got = H.mcdisk(drv, spec, grid, cflux, csum, hoi, seed=1234)
    nbad, err = H.within(got['scube'], ref['scube'], TOL, ref['amb'])
    print(case['name'], 'normwise', err, 'violations', nbad)
    assert nbad == 0
    # total flux agrees to FP32 rounding (ties only move flux between voxels)
    assert abs(got['scube'].sum() - ref['scube'].sum()) <= \
        1e-5 * abs(ref['scube'].sum())


@pytest.mark.parametrize('nclouds', [60000, 500000])
def test_mcdisk_binned_split_tiles_equal_global_path(drv, nclouds):
    """With the spectral cube the only output, clouds are binned by 8 x 8
    spaxel tile and summed per tile in LDS; a tile of more than 8192 records
    (the central tiles here: 16 tiles share the clouds) is split over
    several blocks whose int64 sums meet in a tile accumulator.  With
    another output requested, every deposit is a global int64 atomic.  Both
    are the same integer sums: the cubes are bit-identical."""
    case = dict(cases.MCDISK_CASES[0], nclouds=nclouds)
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    cflux, csum, hoi = cases.cloud_budget(spec, case['traits'],
                                          case['nclouds'])
    binned = H.mcdisk(drv, spec, grid, cflux, csum, hoi, seed=7)['scube']
    glob = H.mcdisk(drv, spec, grid, cflux, csum, hoi, seed=7,
                    want=('scube', 'image'))['scube']
    assert binned.any()
    assert np.array_equal(binned, glob)


def test_mcdisk_reruns_bit_identical(drv):
    """int64 fixed-point accumulation: the cube does not depend on the order
    in which clouds land, so reruns are bit-identical; another seed is not."""
    case = cases.MCDISK_CASES[0]
    spec = cases.build_spec(case).rounded_f32()
    grid = oracle.grid_f32(cases.build_grid(case))
    cflux, csum, hoi = cases.cloud_budget(spec, case['traits'],
                                          case['nclouds'])
    a = H.mcdisk(drv, spec, grid, cflux, csum, hoi, seed=99)['scube']
    b = H.mcdisk(drv, spec, grid, cflux, csum, hoi, seed=99)['scube']
    c = H.mcdisk(drv, spec, grid, cflux, csum, hoi, seed=100)['scube']
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
