"""Serving loop: continuous batching of model-evaluation requests.

gbkfit evaluates one parameter set per Objective.log_likelihood call
(src/gbkfit/objective/objective.py:208-226); a fitter farm or an
interactive tool instead sends a stream of requests, each a parameter set
for one of many resident galaxies.  The Server keeps every galaxy's data
cube, PSF x LSF spectra and FFT plans on the device (one BatchEvaluator per
galaxy) and runs a scheduler that

  * admits work by HBM budget: a galaxy's work buffers (per-set spectra)
    are allocated when it is scheduled and released, least recently used
    first, when another galaxy needs the room -- or, with slot arenas
    (make_arenas), every galaxy runs with the work buffers of the slot it
    is launched in, so the work memory is that of the slots and nothing is
    ever released;
  * batches on arrival: whenever an execution slot (a HIP stream) is free,
    it takes the galaxy with the oldest waiting request and launches up to
    max_batch of its waiting requests as one batched evaluation, and with
    group > 1 does the same for up to group - 1 further galaxies (oldest
    request first) in the same slot, one stream and one completion for all
    of them; requests that arrive while others are on the GPU join later
    batches as slots free up, so batches form continuously and work of
    several galaxies overlaps on the device;
  * completes requests by polling each slot's event, then resolves their
    futures with chi^2 (and moment maps when the galaxy was added with
    orders);
  * keeps Python's cyclic garbage collector off the resident state: the
    objects alive when the server starts (galaxies, models, buffers) are
    frozen out of its generations (gc.freeze) until stop(), so collections
    during serving walk only request-time objects.  Measured on 64
    resident galaxies at 1000 requests/s, collections of the whole heap
    otherwise paused the scheduler thread and set the p99 latency (2.3-3.6
    ms against 1.0 ms frozen).

A request's result does not depend on which batch it joined: each set's
chi^2 is reduced in a fixed order from its own partial sums, and the
moment maps come from its own low-resolution cube.

The scheduling logic talks to galaxies through a small interface
(work_bytes, allocated, release, launch, done, results), so it is tested on
the CPU with stand-in galaxies; GpuGalaxy implements it on BatchEvaluator.
"""
import collections
import concurrent.futures
import gc
import threading
import time

import numpy as np


class Request:
    __slots__ = ('gid', 'params', 't_submit', 't_launch', 't_done', 'future')

    def __init__(self, gid, params, t_submit):
        self.gid = gid
        self.params = params
        self.t_submit = t_submit
        self.t_launch = None
        self.t_done = None
        self.future = concurrent.futures.Future()


class _Entry:
    def __init__(self, gid, galaxy):
        self.gid = gid
        self.galaxy = galaxy
        self.pending = collections.deque()
        self.busy = False
        self.last_used = 0.0


class Server:
    """Continuous-batching request server over resident galaxies.

    hbm_budget: bytes the work buffers of all galaxies may occupy together
    (None: no limit); slots: concurrent batched evaluations; max_batch:
    requests per galaxy and launch (at most the galaxies' own max_batch);
    group: galaxies a slot launches together (their evaluations queued on
    the slot's stream back to back, completing together); freeze_gc:
    freeze the objects alive at start() out of the cyclic collector until
    stop() (a process-wide setting: pass False to leave the collector
    alone).
    """

    def __init__(self, hbm_budget=None, slots=4, max_batch=16,
                 clock=time.perf_counter, freeze_gc=True, group=1):
        if slots < 1 or max_batch < 1 or group < 1:
            raise ValueError("slots, max_batch and group must be >= 1")
        self.hbm_budget = hbm_budget
        self.nslots = slots
        self.max_batch = max_batch
        self.group = group
        self.freeze_gc = freeze_gc
        self._frozen = False
        self.clock = clock
        self._entries = collections.OrderedDict()
        self._slots = [None] * slots   # [(entry, requests, token), ...]
        self._cv = threading.Condition()
        self._thread = None
        self._stop = False
        self.stats = dict(launches=0, galaxy_launches=0, requests=0,
                          completed=0, releases=0, max_batch_seen=0,
                          batch_sizes=[])
        self.latencies = []

    # -- registry ---------------------------------------------------------
    def add(self, gid, galaxy):
        with self._cv:
            if gid in self._entries:
                raise ValueError(f"galaxy {gid!r} already added")
            self._entries[gid] = _Entry(gid, galaxy)

    def galaxy(self, gid):
        return self._entries[gid].galaxy

    # -- client side ------------------------------------------------------
    def submit(self, gid, params):
        """Queues one parameter set for galaxy `gid`; returns a Future whose
        result is what the galaxy's results() gives for it."""
        with self._cv:
            if gid not in self._entries:
                raise KeyError(f"unknown galaxy {gid!r}")
            r = Request(gid, params, self.clock())
            self._entries[gid].pending.append(r)
            self.stats['requests'] += 1
            self._cv.notify()
        return r.future

    # -- scheduler --------------------------------------------------------
    def _work_in_use(self):
        return sum(e.galaxy.work_bytes() for e in self._entries.values()
                   if e.galaxy.allocated())

    def _admit(self, entry):
        """Makes room for entry's work buffers within the budget, releasing
        idle galaxies least recently used first.  False if it cannot fit
        now (busy galaxies hold the memory)."""
        g = entry.galaxy
        if g.allocated() or self.hbm_budget is None:
            return True
        need = g.work_bytes()
        if need > self.hbm_budget:
            raise MemoryError(
                f"galaxy {entry.gid!r} needs {need} bytes of work buffers, "
                f"more than the budget of {self.hbm_budget}")
        idle = sorted((e for e in self._entries.values()
                       if e is not entry and not e.busy and
                       e.galaxy.allocated()), key=lambda e: e.last_used)
        used = self._work_in_use()
        for e in idle:
            if used + need <= self.hbm_budget:
                break
            e.galaxy.release()
            used -= e.galaxy.work_bytes()
            self.stats['releases'] += 1
        return used + need <= self.hbm_budget

    def _complete(self):
        done_any = False
        for i, s in enumerate(self._slots):
            if s is None:
                continue
            # the slot's launches share its stream: done when all are
            if not all(e.galaxy.done(token) for e, _, token in s):
                continue
            now = self.clock()
            for entry, reqs, token in s:
                results = entry.galaxy.results(token, len(reqs))
                for r, res in zip(reqs, results):
                    r.t_done = now
                    self.latencies.append(now - r.t_submit)
                    r.future.set_result(res)
                entry.busy = False
                entry.last_used = now
                self.stats['completed'] += len(reqs)
            self._slots[i] = None
            done_any = True
        return done_any

    def _launch(self):
        launched = False
        while None in self._slots:
            cands = [e for e in self._entries.values()
                     if e.pending and not e.busy]
            if not cands:
                break
            cands.sort(key=lambda e: e.pending[0].t_submit)
            slot = self._slots.index(None)
            group = []
            for e in cands:
                if len(group) == self.group:
                    break
                if not self._admit(e):
                    continue
                n = min(self.max_batch, e.galaxy.max_batch, len(e.pending))
                reqs = [e.pending.popleft() for _ in range(n)]
                now = self.clock()
                for r in reqs:
                    r.t_launch = now
                try:
                    token = e.galaxy.launch([r.params for r in reqs], slot)
                except Exception as exc:    # fail these requests, keep serving
                    for r in reqs:
                        r.future.set_exception(exc)
                    continue
                e.busy = True
                e.last_used = now
                group.append((e, reqs, token))
                self.stats['galaxy_launches'] += 1
            if not group:
                break
            self._slots[slot] = group
            n = sum(len(r) for _, r, _ in group)
            self.stats['launches'] += 1
            self.stats['batch_sizes'].append(n)
            self.stats['max_batch_seen'] = max(self.stats['max_batch_seen'], n)
            launched = True
        return launched

    def step(self):
        """One scheduling pass: complete finished batches, then launch new
        ones.  Returns True if anything happened."""
        with self._cv:
            a = self._complete()
            b = self._launch()
            return a or b

    def idle(self):
        with self._cv:
            return all(s is None for s in self._slots) and not any(
                e.pending for e in self._entries.values())

    def _loop(self):
        while True:
            with self._cv:
                if self._stop and self.idle_locked():
                    return
            if not self.step():
                with self._cv:
                    busy = any(s is not None for s in self._slots)
                    # waiting on the GPU: poll soon; nothing at all: sleep
                    # until a submit() arrives
                    self._cv.wait(timeout=0.0002 if busy else 0.05)

    def idle_locked(self):
        return all(s is None for s in self._slots) and not any(
            e.pending for e in self._entries.values())

    def start(self):
        if self._thread is None:
            if self.freeze_gc and not self._frozen:
                gc.collect()
                gc.freeze()
                self._frozen = True
            self._stop = False
            self._thread = threading.Thread(target=self._loop, daemon=True)
            self._thread.start()

    def stop(self, drain=True):
        """Stops the scheduler after the queued work (drain) is done."""
        if self._thread is None:
            return
        if not drain:
            with self._cv:
                for e in self._entries.values():
                    while e.pending:
                        e.pending.popleft().future.cancel()
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._thread.join()
        self._thread = None
        if self._frozen:
            gc.unfreeze()
            self._frozen = False

    def latency_percentiles(self):
        lat = np.asarray(self.latencies) * 1e3
        if lat.size == 0:
            return dict(p50_ms=None, p99_ms=None, mean_ms=None)
        return dict(p50_ms=float(np.percentile(lat, 50)),
                    p99_ms=float(np.percentile(lat, 99)),
                    mean_ms=float(lat.mean()))


class GpuGalaxy:
    """A resident galaxy served through a BatchEvaluator.

    Each server slot owns a HIP stream and pinned host staging; the galaxy
    owns its device parameter block, chi^2 vector and (when moment maps are
    wanted) low-resolution cubes and maps.  A galaxy runs in one slot at a
    time (the server guarantees it), so these buffers are never shared by
    two launches.
    """

    def __init__(self, driver, dmodel, gmodel, obs, error=None, mask=None,
                 weight=1.0, max_batch=16, template_params=None,
                 orders=None, streams=None, arenas=None):
        from .batch import BatchEvaluator
        import torch
        self._t = torch
        self.driver = driver
        self.ev = BatchEvaluator(driver, dmodel, gmodel, obs, error=error,
                                 mask=mask, weight=weight,
                                 max_batch=max_batch,
                                 template_params=template_params)
        self.max_batch = max_batch
        self.orders = None if orders is None else sorted(int(o)
                                                         for o in orders)
        dev = driver.device
        pst = self.ev.packer.pstride
        self._params = torch.empty((max_batch, pst), dtype=torch.float32,
                                   device=dev)
        self._chi2 = torch.empty(max_batch, dtype=torch.float64, device=dev)
        self._h_params = torch.empty((max_batch, pst), dtype=torch.float32,
                                     pin_memory=True)
        self._h_chi2 = torch.empty(max_batch, dtype=torch.float64,
                                   pin_memory=True)
        self._cubes = self._mmaps = self._h_mmaps = None
        if self.orders is not None:
            cube = dmodel.dcube()
            self._lo_size = tuple(cube.size())
            self._lo_step = tuple(cube.step())
            self._lo_zero = tuple(cube.zero())
            nx, ny, nz = self._lo_size
            self._cubes = torch.empty((max_batch, nz, ny, nx),
                                      dtype=torch.float32, device=dev)
            no = len(self.orders)
            self._mmaps = torch.empty((max_batch, no, ny, nx),
                                      dtype=torch.float32, device=dev)
            self._mmask = torch.empty((max_batch, ny, nx),
                                      dtype=torch.float32, device=dev)
            self._h_mmaps = torch.empty((max_batch, no, ny, nx),
                                        dtype=torch.float32, pin_memory=True)
            self._orders_dev = torch.tensor(self.orders, dtype=torch.int32,
                                            device=dev)
        if self.ev.mc:               # per-set cloud budgets (mcdisk)
            L = self.ev.desc.csum_len
            self._csum = torch.empty((max_batch, L), dtype=torch.int32,
                                     device=dev)
            self._qscale = torch.empty(max_batch, dtype=torch.float64,
                                       device=dev)
            self._h_csum = torch.empty((max_batch, L), dtype=torch.int32,
                                       pin_memory=True)
            self._h_qscale = torch.empty(max_batch, dtype=torch.float64,
                                         pin_memory=True)
        self._streams = streams
        # slot arenas: one work arena per slot (each at least arena_bytes()
        # of this galaxy), shared with the other galaxies of that slot
        self._arenas = arenas
        self.ev.release()            # admission allocates the work buffers
        self._allocated = False

    # -- server interface -------------------------------------------------
    def work_bytes(self):
        # with slot arenas the work memory belongs to the slots
        return 0 if self._arenas is not None else self.ev.memory()['work']

    def arena_bytes(self):
        return self.ev.arena_bytes()

    def use_arenas(self, arenas):
        """Runs with the slots' work arenas (a list, one per slot, each at
        least arena_bytes()), or with work buffers of its own (None)."""
        self.ev.release()
        self._allocated = False
        self._arenas = arenas

    def allocated(self):
        return self._allocated or self._arenas is not None

    def release(self):
        if self._arenas is not None:
            return
        self.ev.release()
        self._allocated = False

    def launch(self, params_list, slot):
        t = self._t
        n = len(params_list)
        stream = self._streams[slot]
        host = self.ev.packer.pack(params_list)
        self._h_params[:n].copy_(t.from_numpy(host))
        if self.ev.mc:
            csum, qs = self.ev.cloud_budgets(params_list)
            self._h_csum[:n].copy_(t.from_numpy(csum))
            self._h_qscale[:n].copy_(t.from_numpy(qs))
        cubes = self._cubes.data_ptr() if self._cubes is not None else None
        self.ev.bind_work(None if self._arenas is None
                          else self._arenas[slot])
        with t.cuda.stream(stream):
            self._params[:n].copy_(self._h_params[:n], non_blocking=True)
            self.ev.set_stream(stream)
            from . import _native as N
            if self.ev.mc:
                self._csum[:n].copy_(self._h_csum[:n], non_blocking=True)
                self._qscale[:n].copy_(self._h_qscale[:n], non_blocking=True)
                N.call('gbkhip_batch_evaluate_clouds', self.ev.handle,
                       self._params.data_ptr(), self._csum.data_ptr(),
                       self._qscale.data_ptr(), n, self._chi2.data_ptr(),
                       cubes, 1)
            else:
                N.call('gbkhip_batch_evaluate', self.ev.handle,
                       self._params.data_ptr(), n, self._chi2.data_ptr(),
                       cubes, 1)
            self._allocated = True
            if self.orders is not None:
                # every set's maps in one launch (the plain kernel's
                # arithmetic, cube by cube)
                N.call('gbkhip_mmaps_moments_batched',
                       self.driver.context().handle, N._i3(self._lo_size),
                       N._f3(self._lo_step), N._f3(self._lo_zero),
                       self._cubes.data_ptr(), 1e-6, len(self.orders),
                       self._orders_dev.data_ptr(), self._mmaps.data_ptr(),
                       self._mmask.data_ptr(), n)
                self._h_mmaps[:n].copy_(self._mmaps[:n], non_blocking=True)
            self._h_chi2[:n].copy_(self._chi2[:n], non_blocking=True)
            ev = t.cuda.Event()
            ev.record(stream)
        return ev

    def done(self, token):
        return token.query()

    def results(self, token, n):
        token.synchronize()
        chi2 = self._h_chi2[:n].numpy().copy()
        if self.orders is None:
            return [dict(chi2=float(c)) for c in chi2]
        maps = self._h_mmaps[:n].numpy().copy()
        return [dict(chi2=float(chi2[i]), mmaps=maps[i]) for i in range(n)]

    def close(self):
        self.ev.close()


def make_arenas(nbytes, n):
    """n device work arenas of nbytes each (one per server slot); nbytes is
    the largest arena_bytes() of the galaxies served."""
    import torch
    return [torch.empty(int(nbytes), dtype=torch.uint8, device='cuda')
            for _ in range(n)]


def make_streams(n):
    """n default-priority HIP streams for the server's slots."""
    import torch
    return [torch.cuda.Stream() for _ in range(n)]
