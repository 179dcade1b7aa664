"""Host-side guards of the batched evaluator on the GPU: a chunk whose y- and
z-pass grids exceed the device's grid limit (launched over slices of sets),
the overflow flag of an mcdisk set with more clouds than the reserved
capacity, and the int64 work cube of the global-atomic mcdisk path zeroed on
the evaluation's own stream after a release."""
import numpy as np
import pytest

import hiputil as H
import pipeline_cases as P

pytestmark = pytest.mark.gpu


def test_chunk_beyond_the_grid_limit():
    """700 sets in one chunk of a 16x16x96 cube (110 hi-res channels reach
    the outputs): the y pass would need 77 000 grid rows of (channel, set),
    beyond the 65 536 of grid z, and runs over slices of sets.  chi^2 equals
    the same sets in chunks of 100, bit for bit."""
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    p = dict(P.c1_params(), size=(16, 16, 96), rnodes=np.arange(0, 9, 1.0),
             bpt_s=3.0, rt=2.0)
    model, base = P.build_product(drv, p)
    ref, _, _ = P.oracle_cube(p)
    rng = np.random.default_rng(17)
    mock = (ref + rng.normal(0, 0.01, ref.shape)).astype(np.float32)
    sets = [P.gparams(dict(p, incl=i, posa=a))
            for i, a in zip(rng.uniform(20, 80, 700), rng.uniform(0, 360, 700))]
    big = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                         error=0.01, max_batch=700, path='direct')
    assert big.info()['fft_size'][2] == 110
    small = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                           error=0.01, max_batch=100, path='direct')
    c_big = big.evaluate(sets, use_graph=True)
    c_small = small.evaluate(sets, use_graph=False)
    print('700-set chunk chi2[0..3]', c_big[:3])
    assert np.all(np.isfinite(c_big)) and np.array_equal(c_big, c_small)
    big.close()
    small.close()


def test_mcdisk_capacity_overflow_is_reported():
    """A direct caller of gbkhip_batch_evaluate_clouds whose set has more
    clouds than the reserved capacity: the evaluation stays within the
    buffers and raises the batch's overflow flag, which check() reports
    once; an evaluation within capacity leaves it clear."""
    import ctypes
    import torch
    from gbkfit_amd import _native as N
    from gbkfit_amd.batch import BatchEvaluator
    drv = H.driver()
    p = dict(P.c1_params(), size=(32, 32, 24), size_z=24)
    model, base = P.build_mc_product(drv, p, cflux=0.01)
    ref = model.model_h(base)['scube']['d']
    mock = (ref + np.random.default_rng(2).normal(0, 0.01, ref.shape)) \
        .astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=2, template_params=base)
    cap = ev._mc_cap
    heavy = dict(base, bpt_a=base['bpt_a'] * 3.0)
    csum, _ = ev.disk.cloud_budget(heavy)
    assert csum[-1] > cap, (csum[-1], cap)
    qs = ctypes.c_double()
    N.call('gbkhip_mcdisk_fixed_scale', ctypes.c_float(ev.desc.cflux),
           int(csum[-1]), ctypes.byref(ev.grid), ctypes.byref(qs))
    dev = drv.device
    csum_d = torch.tensor(csum[None, :], dtype=torch.int32, device=dev)
    qs_d = torch.tensor([qs.value], dtype=torch.float64, device=dev)
    ev.upload_packed(ev.packer.pack([heavy]))
    ev.set_stream(torch.cuda.current_stream())
    N.call('gbkhip_batch_evaluate_clouds', ev.handle,
           ev._params_dev.data_ptr(), csum_d.data_ptr(), qs_d.data_ptr(), 1,
           ev._chi2_dev.data_ptr(), None, 0)
    with pytest.raises(N.GbkHipError, match='capacity'):
        ev.check()
    ev.check()                       # the flag was cleared
    c = ev.evaluate([base])          # within capacity: no error
    assert np.isfinite(c[0])
    ev.close()


def test_mcdisk_global_atomic_cube_zeroed_on_the_evaluation_stream():
    """More than 8192 hi-res channels: a tile spectrum does not fit LDS, so
    the batched mcdisk accumulates into an int64 work cube with global
    atomics, zero-filled when the work buffers are allocated.  After a
    release, an evaluation on a side stream allocates it again; the fill
    runs on that stream ahead of the deposits, and chi^2 and the model
    cubes equal those of the first evaluation bit for bit."""
    import torch
    from gbkfit_amd import gmodels
    from gbkfit_amd.batch import BatchEvaluator
    from gbkfit_amd.dmodels import DModelSCube
    from gbkfit_amd.objective import Model
    from gbkfit_amd.traits import make
    drv = H.driver()
    rn = np.arange(0, 4.0, 0.5)
    cmp = gmodels.SpectralMCDisk3D(
        0.001, False, False, make('bpt', 'exponential'),
        make('vpt', 'tan_arctan'), make('dpt', 'uniform'),
        make('bht', 'sech2'), rnodes=rn)
    model = Model(drv, DModelSCube((8, 8, 8200), (1, 1, 10)),
                  gmodels.GModelKinematics3D([cmp], size_z=8))
    base = dict(xpos=0.0, ypos=0.0, posa=40.0, incl=55.0, vsys=0.0,
                bpt_a=1.0, bpt_s=2.0, bht_s=0.5, vpt_rt=1.5, vpt_vt=150.0,
                dpt_a=20.0)
    ref = model.model_h(base)['scube']['d']
    assert ref.sum() > 0
    mock = (ref + np.random.default_rng(9).normal(0, 0.01, ref.shape)) \
        .astype(np.float32)
    ev = BatchEvaluator(drv, model.dmodel(), model.gmodel(), mock,
                        error=0.01, max_batch=2, template_params=base)
    assert ev.info()['pipeline'] == 'padded'
    sets = [base, dict(base, incl=70.0, posa=200.0)]
    c0, k0 = ev.evaluate(sets, want_cubes=True, use_graph=False)
    assert H.normwise(k0[0], ref) < 1e-5
    ev.release()
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        c1, k1 = ev.evaluate(sets, want_cubes=True, use_graph=True)
    torch.cuda.synchronize()
    assert np.array_equal(c0, c1) and np.array_equal(k0, k1)
    ev.close()
