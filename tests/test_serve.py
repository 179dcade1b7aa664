"""Scheduler of the serving loop (gbkfit_amd.serve.Server) on the CPU with
stand-in galaxies: batching on arrival, continuous admission while batches
are in flight, HBM-budget admission with least-recently-used release, one
slot per galaxy at a time, and each request resolved with its own result."""
import threading

import pytest

from gbkfit_amd.serve import Server


class FakeGalaxy:
    """Launches complete when the test says so; result = f(params)."""

    def __init__(self, work, max_batch=4):
        self.work = work
        self.max_batch = max_batch
        self._alloc = False
        self.launched = []          # list of params lists
        self.inflight = 0
        self.max_inflight = 0
        self.auto = False

    def work_bytes(self):
        return self.work

    def allocated(self):
        return self._alloc

    def release(self):
        assert self.inflight == 0, "released while running"
        self._alloc = False

    def launch(self, params, slot):
        self._alloc = True
        self.inflight += 1
        self.max_inflight = max(self.max_inflight, self.inflight)
        tok = dict(params=list(params), done=self.auto)
        self.launched.append(tok)
        return tok

    def done(self, tok):
        return tok['done']

    def results(self, tok, n):
        self.inflight -= 1
        return [dict(chi2=2.0 * p) for p in tok['params'][:n]]


def finish_all(*gals):
    for g in gals:
        for t in g.launched:
            t['done'] = True


def test_batches_form_on_arrival():
    s = Server(slots=2, max_batch=8)
    g = FakeGalaxy(10, max_batch=4)
    s.add('a', g)
    futs = [s.submit('a', float(i)) for i in range(10)]
    s.step()
    # one galaxy runs in one slot at a time, up to its max_batch
    assert [len(t['params']) for t in g.launched] == [4]
    finish_all(g)
    s.step()
    assert [len(t['params']) for t in g.launched] == [4, 4]
    finish_all(g)
    s.step()
    finish_all(g)
    s.step()
    assert [len(t['params']) for t in g.launched] == [4, 4, 2]
    assert [f.result(0) for f in futs] == [dict(chi2=2.0 * i)
                                            for i in range(10)]
    assert g.max_inflight == 1


def test_requests_join_while_others_run():
    s = Server(slots=2, max_batch=8)
    a, b = FakeGalaxy(10), FakeGalaxy(10)
    s.add('a', a)
    s.add('b', b)
    fa = s.submit('a', 1.0)
    s.step()                               # a runs alone
    fb = [s.submit('b', 2.0), s.submit('b', 3.0)]
    fa2 = s.submit('a', 4.0)               # a busy: waits
    s.step()                               # b admitted into the free slot
    assert len(b.launched) == 1 and b.launched[0]['params'] == [2.0, 3.0]
    assert len(a.launched) == 1
    finish_all(a)
    s.step()                               # a's next batch launches
    assert a.launched[1]['params'] == [4.0]
    finish_all(a, b)
    s.step()
    assert fa.result(0)['chi2'] == 2.0
    assert [f.result(0)['chi2'] for f in fb] == [4.0, 6.0]
    assert fa2.result(0)['chi2'] == 8.0


def test_oldest_request_first():
    s = Server(slots=1, max_batch=8)
    a, b = FakeGalaxy(1), FakeGalaxy(1)
    s.add('a', a)
    s.add('b', b)
    s.submit('b', 1.0)
    s.submit('a', 2.0)
    s.step()
    assert len(b.launched) == 1 and not a.launched


def test_budget_releases_least_recently_used():
    s = Server(hbm_budget=100, slots=3, max_batch=4)
    g = {k: FakeGalaxy(40) for k in 'abc'}
    for k, v in g.items():
        s.add(k, v)
    s.submit('a', 1.0)
    s.step()
    finish_all(*g.values())
    s.step()
    s.submit('b', 1.0)
    s.step()
    finish_all(*g.values())
    s.step()
    assert g['a'].allocated() and g['b'].allocated()
    s.submit('c', 1.0)
    s.step()                               # 120 > 100: release a (LRU)
    assert not g['a'].allocated() and g['b'].allocated()
    assert g['c'].allocated() and s.stats['releases'] == 1


def test_budget_never_releases_running_galaxies():
    s = Server(hbm_budget=100, slots=3, max_batch=4)
    g = {k: FakeGalaxy(40) for k in 'abc'}
    for k, v in g.items():
        s.add(k, v)
    s.submit('a', 1.0)
    s.submit('b', 1.0)
    s.step()                               # a and b running
    fc = s.submit('c', 1.0)
    s.step()                               # no room: c waits
    assert not g['c'].launched and not fc.done()
    finish_all(*g.values())
    s.step()                               # a, b done; c admitted
    assert len(g['c'].launched) == 1
    finish_all(*g.values())
    s.step()
    assert fc.result(0)['chi2'] == 2.0


def test_galaxy_larger_than_budget_is_an_error():
    s = Server(hbm_budget=10, slots=1)
    s.add('a', FakeGalaxy(40))
    s.submit('a', 1.0)
    with pytest.raises(MemoryError):
        s.step()


def test_threaded_clients_all_resolve():
    s = Server(slots=3, max_batch=5)
    gals = {k: FakeGalaxy(1, max_batch=5) for k in 'abcd'}
    for k, v in gals.items():
        v.auto = True
        s.add(k, v)
    s.start()
    futs = []
    lock = threading.Lock()

    def client(k, base):
        for i in range(50):
            f = s.submit(k, base + i)
            with lock:
                futs.append((f, base + i))

    th = [threading.Thread(target=client, args=(k, 1000 * j))
          for j, k in enumerate('abcd')]
    for t in th:
        t.start()
    for t in th:
        t.join()
    s.stop()
    assert all(f.result(1)['chi2'] == 2.0 * v for f, v in futs)
    assert s.stats['completed'] == 200
    assert all(g.max_inflight == 1 for g in gals.values())
    assert len(s.latencies) == 200


def test_start_freezes_resident_objects_until_stop():
    """start() moves the objects alive at that point (resident galaxies and
    their state) out of the cyclic collector's generations, so collections
    while serving do not walk them; stop() puts them back.  freeze_gc=False
    leaves the collector alone."""
    import gc
    gc.unfreeze()
    s = Server(slots=1)
    g = FakeGalaxy(1)
    g.auto = True
    s.add('a', g)
    s.start()
    try:
        assert gc.get_freeze_count() > 0
        assert s.submit('a', 1.0).result(5) is not None
    finally:
        s.stop()
    assert gc.get_freeze_count() == 0
    s2 = Server(slots=1, freeze_gc=False)
    g2 = FakeGalaxy(1)
    g2.auto = True
    s2.add('b', g2)
    s2.start()
    try:
        assert gc.get_freeze_count() == 0
        assert s2.submit('b', 2.0).result(5) is not None
    finally:
        s2.stop()


def test_group_launches_several_galaxies_in_one_slot():
    """group > 1: a free slot takes up to `group` galaxies, oldest request
    first, each with its own batch; the slot completes when all of them
    have."""
    s = Server(slots=1, max_batch=4, group=3)
    g = {k: FakeGalaxy(1) for k in 'abcd'}
    for k, v in g.items():
        s.add(k, v)
    futs = {k: [s.submit(k, float(i)) for i in range(2)] for k in 'dcba'}
    s.step()
    assert [len(g[k].launched) for k in 'abcd'] == [0, 1, 1, 1]
    assert s.stats['launches'] == 1 and s.stats['galaxy_launches'] == 3
    assert s.stats['batch_sizes'] == [6]
    # a waits for the slot; two of the three finished is not done
    for t in g['d'].launched + g['c'].launched:
        t['done'] = True
    s.step()
    assert not futs['d'][0].done() and not g['a'].launched
    finish_all(g['b'])
    s.step()                   # slot completes, then a launches in it
    assert [f.result(0)['chi2'] for f in futs['d']] == [0.0, 2.0]
    assert len(g['a'].launched) == 1
    finish_all(g['a'])
    s.step()
    assert all(f.result(0) for fs in futs.values() for f in fs)
    assert s.stats['completed'] == 8


def test_group_admission_per_galaxy():
    """Each galaxy of a group is admitted on its own: the ones that do not
    fit the budget wait for a later launch."""
    s = Server(hbm_budget=100, slots=2, max_batch=4, group=3)
    g = {k: FakeGalaxy(40) for k in 'abc'}
    for k, v in g.items():
        s.add(k, v)
    fs = [s.submit(k, 1.0) for k in 'abc']
    s.step()                   # a and b fit (80); c would need 120
    assert len(g['a'].launched) == 1 and len(g['b'].launched) == 1
    assert not g['c'].launched
    finish_all(*g.values())
    s.step()                   # a, b done; c admitted (LRU release)
    assert len(g['c'].launched) == 1 and s.stats['releases'] == 1
    finish_all(*g.values())
    s.step()
    assert all(f.result(0)['chi2'] == 2.0 for f in fs)
