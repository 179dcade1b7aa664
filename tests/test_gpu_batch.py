"""Batched evaluation (gbkhip_batch_*) against the plain drop-in path and the
FP64 oracle: chi^2 per set within 1e-5 relative of the oracle's, model cubes
within 1e-5 normwise (+ ambiguity bound), results bit-identical between graph
replay and direct launch and independent of batch composition."""
import numpy as np
import pytest

import hiputil as H
import oracle
import pipeline_cases as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def setup():
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    sets = P.c3_params(np.random.default_rng(123), 6)
    truth = sets[0]
    ref0, _, _ = P.oracle_cube(truth)
    rng = np.random.default_rng(7)
    mock = (ref0 + rng.normal(0, 0.01, ref0.shape)).astype(np.float32)
    model, _ = P.build_product(drv, truth)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=4)
    return drv, ev, sets, mock, model


def test_batch_matches_oracle(setup):
    """Per set of the 128x128x96 workload: the model cube's maximum error
    relative to max|oracle| (printed, and beyond the ambiguity bound
    asserted < 1e-5) and chi^2 against the oracle's."""
    drv, ev, sets, mock, _ = setup
    chi2, cubes = ev.evaluate([P.gparams(p) for p in sets], want_cubes=True,
                              use_graph=False)
    for i, p in enumerate(sets):
        ref, amb, _ = P.oracle_cube(p)
        rel = np.abs(cubes[i] - ref).max() / np.abs(ref).max()
        exc = P.excess_error(cubes[i], ref, amb)
        assert exc < 1e-5, i
        r = oracle.residual(mock, np.full(ref.shape, 0.01), None, ref, None,
                            None, 1.0).reshape(ref.shape)
        c_ref = oracle.residual_sum(r, True)
        # ambiguous voxels can move chi^2 by at most sum 2|r| amb / err^2
        slack = np.sum(2 * np.abs(r) * amb / 0.01 + (amb / 0.01) ** 2)
        print(f"set {i}: cube max rel error {rel:.2e} (beyond amb "
              f"{exc:.2e}), chi2 {chi2[i]:.8g} vs oracle {c_ref:.8g} "
              f"(rel {abs(chi2[i] - c_ref) / c_ref:.2e}, slack {slack:.3g})")
        assert abs(chi2[i] - c_ref) <= 1e-5 * c_ref + slack


def test_batch_matches_plain_path(setup):
    from gbkfit_amd.objective import Data, Objective
    drv, ev, sets, mock, model = setup
    chi2 = ev.evaluate([P.gparams(p) for p in sets])
    obj = Objective([dict(scube=Data(mock, error=np.full(mock.shape, 0.01,
                                                          np.float32)))],
                    [model])
    for i, p in enumerate(sets):
        plain = -2 * obj.log_likelihood(P.gparams(p))[0]
        assert abs(plain - chi2[i]) <= 1e-5 * chi2[i], (i, plain, chi2[i])


def test_graph_replay_bit_identical(setup):
    drv, ev, sets, mock, _ = setup
    params = [P.gparams(p) for p in sets]
    direct = ev.evaluate(params, use_graph=False)
    g1 = ev.evaluate(params, use_graph=True)
    # eager work on the stream between replays
    x = drv.mem_alloc_d((1 << 20,), np.float32)
    drv.mem_fill(x, 3.0)
    g2 = ev.evaluate(params[::-1], use_graph=True)
    g3 = ev.evaluate(params, use_graph=True)
    assert np.array_equal(direct, g1)
    assert np.array_equal(direct, g3)
    assert np.array_equal(direct[::-1], g2)


def test_composition_independence(setup):
    drv, ev, sets, mock, _ = setup
    params = [P.gparams(p) for p in sets]
    full = ev.evaluate(params, use_graph=False)
    for i in (0, 3, 5):
        alone = ev.evaluate([params[i]], use_graph=False)
        assert alone[0] == full[i]


def test_pipeline_selection(setup):
    """Thin disk + separable PSF x LSF: automatic selection takes the direct
    separable pipeline (fewer multiply-adds than the FFTs at this kernel
    width); the fused FFT one runs on the smallest 5-smooth sizes >= S +
    2 edge (256 + 40, 256 + 40, 96 + 14)."""
    from gbkfit_amd.batch import BatchEvaluator
    drv, ev, sets, mock, model = setup
    info = ev.info()
    assert info['pipeline'] == 'direct'
    assert info['fft_size'] == (296, 296, 110)
    # rows and channel bands no line reaches are skipped (the cube beyond
    # the outermost ring node and channels beyond the line-of-sight
    # velocities); the oracle tests cover their zero outputs
    ev.evaluate([P.gparams(p) for p in sets[:2]], use_graph=False)
    (live, total), (glive, gtotal), (blive, btotal) = ev.tile_stats()
    print('live x-pass tiles', live, 'of', total, '; live y-pass groups',
          glive, 'of', gtotal, '; live z-pass blocks', blive, 'of', btotal)
    assert 0 < live < total and 0 < glive < gtotal and 0 < blive <= btotal
    fused = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                           error=0.01, max_batch=2, path='fused')
    assert fused.info()['pipeline'] == 'fused'
    assert fused.info()['fft_size'] == (300, 300, 120)
    fused.close()


@pytest.mark.parametrize('path', ['fused', 'padded'])
def test_pipelines_agree(setup, path):
    """Direct separable pipeline vs the fused FFT one and vs the FFT of the
    whole padded cube: chi^2 within 1e-5 relative, cubes 1e-5 normwise."""
    from gbkfit_amd.batch import BatchEvaluator
    drv, ev, sets, mock, model = setup
    ref = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                         error=0.01, max_batch=4, path=path)
    assert ref.info()['pipeline'] == path
    params = [P.gparams(p) for p in sets]
    c_d, cubes_d = ev.evaluate(params, want_cubes=True, use_graph=False)
    c_p, cubes_p = ref.evaluate(params, want_cubes=True, use_graph=False)
    for i in range(len(sets)):
        assert abs(c_d[i] - c_p[i]) <= 1e-5 * c_p[i], (i, c_d[i], c_p[i])
        assert H.normwise(cubes_d[i], cubes_p[i]) < 1e-5, i
    ref.close()


@pytest.mark.parametrize('scale,size,psf,lsf', [
    ((1, 1, 1), (48, 40, 32), 2.0, 25.0),
    ((3, 2, 2), (24, 30, 20), 2.5, 30.0),
    ((2, 3, 3), (20, 18, 16), 1.5, 60.0),
    ((3, 1, 1), (184, 12, 16), 1.5, 25.0),
])
def test_direct_pipeline_other_shapes(scale, size, psf, lsf):
    """Direct pipeline at other downscale factors and tap buckets: model
    cubes against the FP64 oracle (1e-5 normwise + ambiguity bound) and
    chi^2 against the fused FFT pipeline.  The last shape has hi-res rows
    wider than the 512 lines an x-pass block keeps in registers."""
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    p = dict(P.c1_params(), size=size, scale=scale, psf_fwhm=psf,
             lsf_fwhm=lsf)
    model, _ = P.build_product(drv, p)
    ref, amb, _ = P.oracle_cube(p)
    rng = np.random.default_rng(11)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=3, path='direct')
    assert ev.info()['pipeline'] == 'direct'
    sets = [dict(p, incl=i, posa=a) for i, a in ((50, 30), (65, 140),
                                                  (80, 280))]
    chi2, cubes = ev.evaluate([P.gparams(s) for s in sets], want_cubes=True,
                              use_graph=False)
    for i, s in enumerate(sets):
        r, a, _ = P.oracle_cube(s)
        assert P.excess_error(cubes[i], r, a) < 1e-5, i
    fu = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=3, path='fused')
    c_f = fu.evaluate([P.gparams(s) for s in sets], use_graph=False)
    for i in range(len(sets)):
        assert abs(chi2[i] - c_f[i]) <= 1e-5 * c_f[i], (i, chi2[i], c_f[i])
    fu.close()
    ev.close()


@pytest.mark.parametrize('path', ['auto', 'direct', 'padded'])
def test_thick_disk_batch_matches_plain_path(path):
    """A thick disk runs the direct pipeline from its deposited hi-res cube
    (automatic choice) or the padded-cube one; chi^2 equals the plain path,
    and two lanes of chunks give the chi^2 of one, bit for bit (the thick
    deposit sums every line of sight in a fixed order), as do reruns."""
    from gbkfit_amd.batch import BatchEvaluator
    from gbkfit_amd.objective import Data, Objective
    drv = H.driver()
    p = dict(P.c1_params(), thin=False, size=(32, 32, 24))
    model, _ = P.build_product(drv, p)
    ref, _, _ = P.oracle_cube(p)
    rng = np.random.default_rng(3)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=2, path=path)
    want = 'direct' if path == 'auto' else path
    assert ev.info()['pipeline'] == want
    sets = [P.gparams(dict(p, incl=i, posa=a)) for i, a in
            ((40, 10), (60, 200), (75, 300))]
    chi2 = ev.evaluate(sets, use_graph=False)
    obj = Objective([dict(scube=Data(mock, error=np.full(mock.shape, 0.01,
                                                          np.float32)))],
                    [model])
    for i, s in enumerate(sets):
        plain = -2 * obj.log_likelihood(s)[0]
        assert abs(plain - chi2[i]) <= 1e-5 * chi2[i], (i, plain, chi2[i])
    again = ev.evaluate(sets, use_graph=True)
    assert np.array_equal(again, chi2)
    if want == 'direct':
        ev.set_tiles(2, 2)
        two = ev.evaluate(sets, use_graph=True)
        assert np.array_equal(two, chi2), (two, chi2)
    ev.close()


@pytest.mark.parametrize('path', ['auto', 'direct', 'padded'])
def test_mcdisk_batch_matches_plain_path(path):
    """Batched mcdisk (direct pipeline from the deposited hi-res cube, the
    automatic choice, or padded-cube FFT; per-set cloud budgets): model
    cubes and chi^2 of every set against the plain one-model path (which the
    kernel tests pin to the oracle's Philox stream), a set alone equal to
    the same set in the batch, graph replay equal to direct launches and,
    on the direct pipeline, two lanes of chunks equal to one."""
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    p = dict(P.c1_params(), size=(32, 32, 24), size_z=24)
    model, base = P.build_mc_product(drv, p, cflux=0.01)
    ref = model.model_h(base)['scube']['d']
    rng = np.random.default_rng(5)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=2, template_params=base,
                        path=path)
    want = 'direct' if path == 'auto' else path
    assert ev.info()['pipeline'] == want
    sets = [dict(base, incl=i, posa=a, bpt_s=s) for i, a, s in
            ((45, 30, 8.0), (60, 200, 6.0), (75, 300, 10.0))]
    chi2, cubes = ev.evaluate(sets, want_cubes=True, use_graph=False)
    for i, s in enumerate(sets):
        plain = model.model_h(s)['scube']['d']
        assert H.normwise(cubes[i], plain) < 1e-5, i
        c_ref = float(np.sum(((plain.astype(np.float64) - mock) / 0.01) ** 2))
        assert abs(chi2[i] - c_ref) <= 1e-5 * c_ref, (i, chi2[i], c_ref)
    alone = ev.evaluate([sets[1]], use_graph=False)
    assert alone[0] == chi2[1]
    g = ev.evaluate(sets, use_graph=True)
    assert np.array_equal(g, chi2)
    if want == 'direct':
        ev.set_tiles(2, 2)
        two = ev.evaluate(sets, use_graph=True)
        assert np.array_equal(two, chi2)
        print('mcdisk direct tiles', ev.tile_stats())
    ev.close()


@pytest.mark.parametrize('path', ['direct', 'fused', 'padded'])
def test_error_cube_mask_weight(path):
    """chi^2 with an error cube, a mask and a weight (objective.hpp:7-33)
    against the oracle's residual on every pipeline; on the direct pipeline
    this also covers the data-only partial sums of z-pass blocks with no
    live input, which fold in the same error and mask."""
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    p = dict(P.c1_params(), size=(48, 44, 32))
    model, base = P.build_product(drv, p)
    ref, amb, _ = P.oracle_cube(p)
    rng = np.random.default_rng(21)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    err = rng.uniform(0.005, 0.02, ref.shape).astype(np.float32)
    mask = (rng.uniform(size=ref.shape) > 0.2).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=err, mask=mask, weight=0.7, max_batch=2,
                        path=path)
    assert ev.info()['pipeline'] == path
    sets = [dict(p, incl=i, posa=a) for i, a in ((35, 10), (62, 150),
                                                  (80, 300))]
    chi2 = ev.evaluate([P.gparams(s) for s in sets], use_graph=True)
    for i, s in enumerate(sets):
        r_cube, r_amb, _ = P.oracle_cube(s)
        r = oracle.residual(mock, err, mask, r_cube, None, None,
                            0.7).reshape(r_cube.shape)
        c_ref = oracle.residual_sum(r, True)
        slack = np.sum(mask * 0.7 * (2 * np.abs(r) * r_amb / err +
                                     0.7 * (r_amb / err) ** 2))
        assert abs(chi2[i] - c_ref) <= 1e-5 * c_ref + slack, \
            (path, i, chi2[i], c_ref)
    ev.close()


@pytest.mark.parametrize('path', ['direct', 'fused', 'padded'])
def test_sets_with_no_model_flux(path):
    """Edge cases in a batch: a disk entirely outside the field and a disk
    of zero surface brightness give an all-zero model, so chi^2 is the
    data's own sum of (obs / err)^2 (on the direct pipeline every tile is
    skipped and every z-pass block takes the precomputed data-only partial);
    the normal sets around them keep their chi^2 from a batch of their
    own."""
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    p = dict(P.c1_params(), size=(48, 44, 32))
    model, base = P.build_product(drv, p)
    ref, _, _ = P.oracle_cube(p)
    rng = np.random.default_rng(31)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=4, path=path)
    assert ev.info()['pipeline'] == path
    normal = [P.gparams(dict(p, incl=i, posa=a)) for i, a in ((45, 20),
                                                             (70, 250))]
    outside = dict(normal[0], xpos=1e4, ypos=-1e4)
    dark = dict(normal[1], bpt_a=0.0)
    chi2 = ev.evaluate([normal[0], outside, dark, normal[1]],
                       use_graph=True)
    data_only = float(np.sum((mock.astype(np.float64) / 0.01) ** 2))
    for c in chi2[1:3]:
        assert abs(c - data_only) <= 1e-5 * data_only, (path, c, data_only)
    own = ev.evaluate(normal, use_graph=False)
    assert np.array_equal(own, chi2[[0, 3]]), (own, chi2)
    if path == 'direct':
        ev.evaluate([outside], use_graph=False)
        (live, _), (glive, _), (blive, _) = ev.tile_stats()
        assert live == 0 and glive == 0 and blive == 0
    ev.close()


def test_mcdisk_set_with_no_clouds():
    """A zero-brightness mcdisk set has a cloud budget of 0: its model is
    zero and its chi^2 the data-only sum, beside normal sets in the same
    batch (direct pipeline, binned deposits)."""
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    p = dict(P.c1_params(), size=(32, 32, 24), size_z=24)
    model, base = P.build_mc_product(drv, p, cflux=0.01)
    ref = model.model_h(base)['scube']['d']
    rng = np.random.default_rng(41)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=3, template_params=base)
    assert ev.info()['pipeline'] == 'direct'
    dark = dict(base, bpt_a=0.0)
    csum, _ = ev.cloud_budgets([dark])
    assert csum[0][-1] == 0
    chi2 = ev.evaluate([base, dark, dict(base, incl=70.0)], use_graph=True)
    data_only = float(np.sum((mock.astype(np.float64) / 0.01) ** 2))
    assert abs(chi2[1] - data_only) <= 1e-5 * data_only, (chi2, data_only)
    alone = ev.evaluate([base], use_graph=False)
    assert alone[0] == chi2[0]
    ev.close()


@pytest.mark.parametrize('path', ['direct', 'fused', 'padded'])
def test_loose_tilted_disk_matches_plain_path(path):
    """A loose and tilted thin disk (node-wise centre, systemic velocity,
    position angle and inclination; ring_info, gmodels.hpp:217-323) on each
    batched pipeline: chi^2 of every set within 1e-5 of the plain one-model
    path, and the model cubes within 1e-5 normwise."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.batch import BatchEvaluator
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Data, Model, Objective
    from gbkfit_amd.traits import make
    drv = H.driver()
    rn = np.arange(0, 22, 1.5)
    cmp = gmodels.SpectralSMDisk2D(
        True, True, make('bpt', 'exponential'), make('vpt', 'tan_arctan'),
        make('dpt', 'uniform'), rnodes=rn)
    dm = DModelSCube((48, 44, 32), (1, 1, 10), scale=(2, 2, 1),
                     psf=psflsf.PSFGauss(1.1), lsf=psflsf.LSFGauss(8.5))
    model = Model(drv, dm, gmodels.GModelKinematics2D([cmp]))
    n = rn.size

    def params(i):
        t = np.linspace(0.0, 1.0, n)
        return dict(xpos=0.5 * np.sin(3 * t + i), ypos=-0.4 * t,
                    vsys=5.0 * t, posa=30.0 + 40.0 * t + 10 * i,
                    incl=40.0 + 25.0 * t + 5 * i, bpt_a=1.0, bpt_s=6.0,
                    vpt_rt=4.0, vpt_vt=180.0 + 20 * i, dpt_a=25.0)
    sets = [params(i) for i in range(3)]
    ref = model.model_h(sets[0])['scube']['d']
    rng = np.random.default_rng(51)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=2, template_params=sets[0],
                        path=path)
    assert ev.info()['pipeline'] == path
    chi2, cubes = ev.evaluate(sets, want_cubes=True, use_graph=True)
    obj = Objective([dict(scube=Data(mock, error=np.full(mock.shape, 0.01,
                                                          np.float32)))],
                    [model])
    for i, s in enumerate(sets):
        plain = -2 * obj.log_likelihood(s)[0]
        assert abs(plain - chi2[i]) <= 1e-5 * chi2[i], (path, i, plain,
                                                        chi2[i])
        assert H.normwise(cubes[i], model.model_h(s)['scube']['d']) < 1e-5
    ev.close()


def test_nodewise_traits_match_plain_path():
    """Node-wise traits (a surface brightness and a rotation velocity per
    ring node, interpolated onto the sub-rings) on the direct pipeline:
    chi^2 within 1e-5 of the plain path and cubes within 1e-5 normwise."""
    from gbkfit_amd import gmodels, psflsf
    from gbkfit_amd.batch import BatchEvaluator
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Data, Model, Objective
    from gbkfit_amd.traits import make
    drv = H.driver()
    rn = np.arange(0, 22, 1.5)
    cmp = gmodels.SpectralSMDisk2D(
        False, False, make('bpt', 'nw_uniform'),
        make('vpt', 'nw_tan_uniform'), make('dpt', 'uniform'), rnodes=rn)
    dm = DModelSCube((48, 44, 32), (1, 1, 10), scale=(2, 2, 1),
                     psf=psflsf.PSFGauss(1.1), lsf=psflsf.LSFGauss(8.5))
    model = Model(drv, dm, gmodels.GModelKinematics2D([cmp]))
    t = np.linspace(0.0, 1.0, rn.size)

    def params(i):
        return dict(vsys=0.0, xpos=0.3 * i, ypos=0.0, posa=25.0 + 30 * i,
                    incl=45.0 + 10 * i, bpt_a=np.exp(-2.5 * t) + 0.05,
                    vpt_vt=40.0 + (160.0 + 15 * i) * np.sqrt(t),
                    dpt_a=20.0 + 5 * i)
    sets = [params(i) for i in range(3)]
    ref = model.model_h(sets[0])['scube']['d']
    rng = np.random.default_rng(61)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=2, template_params=sets[0],
                        path='direct')
    assert ev.info()['pipeline'] == 'direct'
    chi2, cubes = ev.evaluate(sets, want_cubes=True, use_graph=True)
    obj = Objective([dict(scube=Data(mock, error=np.full(mock.shape, 0.01,
                                                          np.float32)))],
                    [model])
    for i, s in enumerate(sets):
        plain = -2 * obj.log_likelihood(s)[0]
        assert abs(plain - chi2[i]) <= 1e-5 * chi2[i], (i, plain, chi2[i])
        assert H.normwise(cubes[i], model.model_h(s)['scube']['d']) < 1e-5
    ev.close()


def _trim(k, s):
    """Block-mean-folded taps of the kernel k (centred, 2e + 1 values) at
    downscale s, and the counts of leading and trailing taps whose |sum|
    stays within 2^-31 of the total."""
    e = len(k) // 2
    w = np.zeros(s + 2 * e)
    for u in range(s + 2 * e):
        for d in range(s):
            t = d + e - u
            if -e <= t <= e:
                w[u] += k[t + e] / s
    tot = np.abs(w).sum()
    a = int(np.sum(np.cumsum(np.abs(w)) <= 2.0 ** -31 * tot))
    b = int(np.sum(np.cumsum(np.abs(w[::-1])) <= 2.0 ** -31 * tot))
    return len(w), a, b


def test_direct_pipeline_taps(setup):
    """The taps the direct pipeline uses on the 128x128x96 workload (PSF
    FWHM 3 px at 2x supersampling, LSF FWHM 2 channels): the
    block-mean-folded kernels, built here from the oracle's PSF image and
    LSF, lose exactly the leading and trailing taps whose |sum| stays within
    2^-31 of the total (x, y: 5 per end of 42, 32 kept; z: 2 per end of 15,
    11 kept)."""
    drv, ev, sets, mock, _ = setup
    assert ev.info()['pipeline'] == 'direct'
    (tx, ty, tz), (ox, oy, oz) = ev.taps()
    p = sets[0]
    geom = P.oracle_geom(p)
    sig = p['psf_fwhm'] * P.SIG_FWHM
    psz = oracle.psf_gauss_size(sig, geom.step_hi[:2])
    kx = oracle.psf_gauss_array(sig, geom.step_hi[:2], psz).sum(axis=0)
    assert geom.edge_hi[0] == len(kx) // 2
    n, a, b = _trim(kx, p['scale'][0])
    sgz = p['lsf_fwhm'] * P.SIG_FWHM
    kz = oracle.lsf_gauss_array(sgz, geom.step_hi[2],
                                oracle.lsf_gauss_size(sgz, geom.step_hi[2]))
    assert geom.edge_hi[2] == len(kz) // 2
    nz, az, bz = _trim(kz, p['scale'][2])
    print('taps', (tx, ty, tz), 'offsets', (ox, oy, oz), 'of', (n, nz))
    assert (n, a, b) == (42, 5, 5) and (nz, az, bz) == (15, 2, 2)
    assert (tx, ty, tz) == (n - a - b, n - a - b, nz - az - bz)
    assert (ox, oy, oz) == (a, a, az)
