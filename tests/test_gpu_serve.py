"""Serving loop on the GPU: requests from several client threads for three
resident galaxies of different sizes, batched continuously over three
streams under an HBM budget that holds the work buffers of only two
galaxies.  Every request's chi^2 and moment maps must equal, bit for bit,
those of the same request evaluated alone; batches of more than one request
must have formed and the budget must have released work buffers."""
import random
import threading

import numpy as np
import pytest

import hiputil as H
import pipeline_cases as P

pytestmark = pytest.mark.gpu

SIZES = [(32, 32, 24), (40, 36, 32), (48, 44, 24)]


@pytest.fixture(scope='module')
def galaxies():
    from gbkfit_amd.serve import GpuGalaxy, make_streams
    drv = H.driver()
    streams = make_streams(3)
    gals, sets = {}, {}
    for i, size in enumerate(SIZES):
        p = dict(P.c1_params(), size=size, rnodes=np.arange(0, 12, 1.5),
                 vt=150.0 + 20 * i)
        model, truth = P.build_product(drv, p)
        cube = model.model_h(truth)['scube']['d'].astype(np.float32)
        rng = np.random.default_rng(i)
        mock = (cube + rng.normal(0, 0.01, cube.shape)).astype(np.float32)
        gals[i] = GpuGalaxy(drv, model.dmodel(), model.gmodel(), mock,
                            error=0.01, max_batch=8, template_params=truth,
                            orders=(0, 1, 2), streams=streams)
        rs = np.random.default_rng(100 + i)
        sets[i] = [dict(truth, incl=float(rs.uniform(30, 70)),
                        posa=float(rs.uniform(0, 360)),
                        vpt_vt=float(rs.uniform(100, 250)))
                   for _ in range(20)]
    yield drv, gals, sets
    for g in gals.values():
        g.close()


def test_mixed_stream_equals_alone(galaxies):
    from gbkfit_amd.serve import Server
    drv, gals, sets = galaxies
    works = sorted(g.work_bytes() for g in gals.values())
    # room for the two largest work sets, never for all three
    srv = Server(hbm_budget=works[-1] + works[-2], slots=3, max_batch=8)
    for k, g in gals.items():
        srv.add(k, g)
    reqs = [(k, i) for k in gals for i in range(len(sets[k]))]
    random.Random(5).shuffle(reqs)
    futs = {}
    lock = threading.Lock()

    def client(part):
        rnd = random.Random(len(part))
        for k, i in part:
            f = srv.submit(k, sets[k][i])
            with lock:
                futs[(k, i)] = f
            if rnd.random() < 0.3:
                threading.Event().wait(0.001)

    srv.start()
    th = [threading.Thread(target=client, args=(reqs[j::2],))
          for j in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.stop()
    mixed = {key: f.result(60) for key, f in futs.items()}
    print('stats', {k: v for k, v in srv.stats.items() if k != 'batch_sizes'},
          srv.latency_percentiles())
    assert srv.stats['completed'] == len(reqs)
    assert srv.stats['max_batch_seen'] > 1
    assert srv.stats['releases'] > 0
    # each request alone, through the same galaxy (one set per launch)
    for (k, i), res in mixed.items():
        tok = gals[k].launch([sets[k][i]], 0)
        alone = gals[k].results(tok, 1)[0]
        assert alone['chi2'] == res['chi2'], (k, i)
        assert np.array_equal(alone['mmaps'], res['mmaps'], equal_nan=True)


def test_served_chi2_matches_direct_batch(galaxies):
    """Slot stream + graph replay gives the same chi^2 as a direct batched
    evaluation on torch's stream."""
    drv, gals, sets = galaxies
    g = gals[1]
    tok = g.launch(sets[1][:4], 0)
    res = g.results(tok, 4)
    chi2 = g.ev.evaluate(sets[1][:4], use_graph=False)
    assert [r['chi2'] for r in res] == list(chi2)


def test_slot_arenas_grouped_equal_alone(galaxies):
    """Every galaxy runs with the work arena of its slot (one arena per
    slot, shared by the galaxies launched there), several galaxies per slot
    launch: each request's chi^2 and maps equal those of the request alone
    with the galaxy's own buffers, bit for bit; no work buffers are
    released."""
    from gbkfit_amd.serve import Server, make_arenas
    drv, gals, sets = galaxies
    arena = max(g.arena_bytes() for g in gals.values())
    arenas = make_arenas(arena, 3)
    for g in gals.values():
        g.use_arenas(arenas)
    srv = Server(slots=3, max_batch=8, group=2)
    for k, g in gals.items():
        srv.add(k, g)
    reqs = [(k, i) for k in gals for i in range(len(sets[k]))]
    random.Random(9).shuffle(reqs)
    srv.start()
    futs = {key: srv.submit(key[0], sets[key[0]][key[1]]) for key in reqs}
    srv.stop()
    served = {key: f.result(60) for key, f in futs.items()}
    print('stats', {k: v for k, v in srv.stats.items() if k != 'batch_sizes'})
    assert srv.stats['completed'] == len(reqs)
    assert srv.stats['releases'] == 0
    assert srv.stats['galaxy_launches'] >= srv.stats['launches']
    for g in gals.values():
        g.use_arenas(None)
    for (k, i), res in served.items():
        alone = gals[k].results(gals[k].launch([sets[k][i]], 0), 1)[0]
        assert alone['chi2'] == res['chi2'], (k, i)
        assert np.array_equal(alone['mmaps'], res['mmaps'], equal_nan=True)


def test_arena_binding_and_rejection(galaxies):
    """A batch alternating between two arenas and its own buffers gives the
    same chi^2 each time; an arena below arena_bytes() is refused."""
    import torch
    from gbkfit_amd import _native as N
    drv, gals, sets = galaxies
    ev = gals[2].ev
    need = ev.arena_bytes()
    a = torch.empty(need, dtype=torch.uint8, device='cuda')
    b = torch.empty(need + 4096, dtype=torch.uint8, device='cuda')
    ref = ev.evaluate(sets[2][:3])
    for arena in (a, b, a, None, b):
        ev.bind_work(arena)
        assert np.array_equal(ev.evaluate(sets[2][:3]), ref)
    small = torch.empty(need - 256, dtype=torch.uint8, device='cuda')
    with pytest.raises(N.GbkHipError):
        ev.bind_work(small)
    ev.bind_work(None)
