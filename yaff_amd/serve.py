"""Serving loop: continuous batching of force evaluations and MD segments.

Requests (a single energy/gradient/virial evaluation, or an MD segment of
n velocity-Verlet steps) arrive at any time.  The server keeps the systems of
the admitted requests resident on the device in ONE persistent batch engine
(yaff_amd.slots.SlotEngine, created once and kept) and advances them
together, one batch iteration at a time (kick/drift, neighbour list, pair
potentials, Ewald reciprocal, kick for every member in the same launches):

* an evaluation needs one iteration (time step 0: the kicks and the drift do
  nothing, the forces are computed);
* an MD segment needs 1 + n iterations (the first, with time step 0, gives the
  initial forces; the next n are Verlet steps);
* iterations run in quanta; a quantum is enqueued with the copies of the
  results of the members that finish in it and an event; arrivals are
  planned by a few host threads (k vectors and tables: native code and
  large array operations that release the interpreter lock) while the
  device works; when the event has fired, the neighbour-list status is
  checked (an overflow grows the capacities and repeats the quantum from
  its start), finished requests leave their slots and waiting ones are
  placed in free ranges: only their own slices of the device tables are
  written, the members that stay are not touched;
* admission is first come, first served, while the estimated device memory
  of the batch with the candidate stays within the budget and the batch has
  a free slot; a candidate whose planning has not finished waits for a later
  quantum unless the batch is empty.

Launch parameters are fixed for the server (not re-tuned per batch), and every
kernel of the engine is deterministic and per-system, so a request gets the
same bits alone or in any mix.

The velocity-Verlet splitting follows yaff/sampling/verlet.py:339-367.
"""
import collections
import concurrent.futures
import time

import numpy as np
import torch

from . import _lib
from .engine import _round_up, _ceil_div, ewald_kvectors_cached, ALIGN
from .batch import nbsystem, timestep, maxwell_boltzmann
from .slots import SlotEngine

__all__ = ['Request', 'Server', 'estimate_bytes', 'poisson_requests', 'latency_summary', 'SERVER_TUNING']

# Fixed launch parameters of the server: results do not depend on the batch.
SERVER_TUNING = dict(pair_lanes=32, apl=4, skb=64, fkb=64, fapl=4)


class Request(object):
    """One unit of work on one system.  ``kind`` is 'eval' or 'md';
    ``nsteps`` the number of Verlet steps of an MD segment."""

    def __init__(self, rid, system, kind='eval', nsteps=0, arrival=0.0, vel=None, dt=None):
        if kind not in ('eval', 'md'):
            raise ValueError('kind must be eval or md')
        self.rid = rid
        self.system = system
        self.kind = kind
        self.nsteps = int(nsteps) if kind == 'md' else 0
        self.arrival = float(arrival)
        self.nb = nbsystem(system)
        self.natom = system.natom
        self.vel = vel
        self.dt = timestep(system) if dt is None else dt
        if kind == 'md' and self.vel is None:
            self.vel = maxwell_boltzmann(system.masses, 120.0 if system.spec.kind == 'lj' else 300.0, rid)
        self.iters_total = 1 + self.nsteps
        self.iters_sched = 0    # iterations enqueued on the device
        self.iters_done = 0     # iterations completed
        # results
        self.energy = None
        self.gpos = None
        self.vtens = None
        self.pos = None
        self.vel_out = None
        self.t_admit = None
        self.t_done = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    @property
    def force_evals(self):
        return self.natom*self.iters_total


def _kcap_estimate(nb):
    s = nb
    if s.natom == 0:
        return 16
    if s.cell.nvec == 3:
        rho = s.natom/s.cell.volume
        n = 4.0/3.0*np.pi*s.rc**3*rho
    else:
        n = s.natom*int(np.prod(2*s.rmax + 1))
    return int(1.3*n) + 32


def _estimate_terms(s, chunk, fkb):
    """Per-system terms of estimate_bytes (padded atoms, row capacity, Ewald
    bytes, force slots), kept on the system for its launch parameters: an
    admission re-estimates the whole batch, so the terms are computed once
    per system, not once per admission."""
    key = (chunk, fkb)
    memo = getattr(s, '_estimate_terms', None)
    if memo is not None and memo[0] == key:
        return memo[1]
    ew_bytes = nslot = 0
    if s.ewald is not None and s.natom > 0:
        nk = len(ewald_kvectors_cached(s)[1])
        nkp = _round_up(max(nk, 1), 256)
        nchunk = _ceil_div(s.natom, chunk)
        ew_bytes = nkp*(32 + 16 + 48 + 64) + 16*nchunk*nkp
        # greedy row blocks: two consecutive blocks hold more than fkb k vectors
        nslot = _ceil_div(2*_ceil_div(nk, fkb) + 1, 4)
    terms = (_round_up(max(s.natom, 1), ALIGN), _kcap_estimate(s), ew_bytes, nslot)
    s._estimate_terms = (key, terms)
    return terms


def estimate_bytes(nbs, tuning=SERVER_TUNING):
    """Device memory of a batch engine over ``nbs`` (the dominant allocations
    of Engine: per-atom state, the ELL neighbour list at the batch's row
    capacity, the Ewald tables and partial sums)."""
    if not nbs:
        return 0
    chunk = 64*(tuning.get('apl') or 4)
    fkb = tuning.get('fkb') or 64
    terms = [_estimate_terms(s, chunk, fkb) for s in nbs]
    npad = sum(t[0] for t in terms)
    kcap = max(t[1] for t in terms)
    total = npad*(8*kcap + 8*3*3 + 8*3*_lib.MAX_POT + 8*3 + 4 + 2*_lib.MAX_POT*8 + 64)
    total += sum(t[2] for t in terms)
    nslot = max(t[3] for t in terms)
    if nslot:
        total += npad*(nslot*24 + 64)
    return int(total)


def poisson_requests(n, rate, natom_range=(100, 50000), md_fraction=0.5, nsteps=20, seed=0,
                     angle_range=(80.0, 100.0)):
    """n requests with exponential inter-arrival times of mean 1/rate (a
    Poisson stream), systems log-uniform in natom_range alternating LJ fluid
    and rocksalt with triclinic angles in angle_range; a fraction md_fraction
    of them MD segments of nsteps steps, the rest single evaluations."""
    from . import workloads
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0/rate, n))
    kinds = rng.uniform(0, 1, n) < md_fraction
    systems = workloads.batch_of_systems(n, natom_range, angle_range, seed=seed + 1, log_uniform=True)
    return [Request(i, systems[i], 'md' if kinds[i] else 'eval', nsteps if kinds[i] else 0, arrivals[i])
            for i in range(n)]


def latency_summary(done):
    """p50/p99/max latency (s) and throughput (atom-force-evals/s over the
    span from the first arrival to the last completion)."""
    lat = np.array([r.latency for r in done])
    span = max(r.t_done for r in done) - min(r.arrival for r in done)
    evals = sum(r.force_evals for r in done)
    return {'requests': len(done), 'p50_ms': float(np.percentile(lat, 50))*1e3,
            'p99_ms': float(np.percentile(lat, 99))*1e3, 'max_ms': float(lat.max())*1e3,
            'throughput': evals/span, 'span_s': span}


class Server(object):
    """Continuous-batching server of Requests on one device, over one
    persistent SlotEngine (created on the first run and kept)."""

    def __init__(self, precision='mixed', hbm_budget=64 << 30, max_systems=64, max_quantum=4,
                 tuning=SERVER_TUNING, planners=4, reserve_atoms=1 << 16, depth=2):
        self.planners = int(planners)
        self.depth = int(depth)                   # quanta in flight
        self.reserve_atoms = int(reserve_atoms)   # starting atom capacity of the engine
        self.precision = precision
        self.hbm_budget = int(hbm_budget)
        self.max_systems = int(max_systems)
        self.max_quantum = int(max_quantum)
        self.tuning = dict(tuning)
        self.active = []            # admitted requests
        self.waiting = collections.deque()
        self.engine = None
        self.slot_of = {}           # id(request) -> slot
        self.done = []
        self.stats = collections.Counter()
        self.peak_bytes = 0

    # -- admission policy (host only) --------------------------------------
    def _admit(self, now):
        """Moves waiting requests into the batch, first come first served,
        while the estimated device memory of the batch stays within the
        budget and a slot is free; returns the admitted requests."""
        new = []
        while self.waiting and len(self.active) < self.max_systems:
            cand = self.waiting[0]
            f = getattr(cand, '_plan', None)
            if f is not None and not f.done() and self.active:
                break     # join at a later quantum rather than keep the device waiting
            self._planned(cand)   # its k vectors, which the estimate counts
            need = estimate_bytes([r.nb for r in self.active] + [cand.nb], self.tuning)
            if need > self.hbm_budget:
                if not self.active:
                    raise RuntimeError('request %d alone needs %.1f GiB, over the budget of %.1f GiB'
                                       % (cand.rid, need/2.0**30, self.hbm_budget/2.0**30))
                break
            self.waiting.popleft()
            cand.t_admit = now
            self.active.append(cand)
            new.append(cand)
            self.peak_bytes = max(self.peak_bytes, need)
        return new

    @staticmethod
    def _planned(r):
        """The request's planning (SlotEngine.prepare, run by the planning
        threads from its arrival on); waits for it when not done yet."""
        f = getattr(r, '_plan', None)
        return None if f is None else f.result()

    # -- device side ---------------------------------------------------------
    def _engine(self):
        if self.engine is None:
            t = self.tuning
            self.engine = SlotEngine(self.max_systems, self.precision, pair_lanes=t['pair_lanes'], apl=t['apl'],
                                     skb=t['skb'], fkb=t['fkb'], fapl=t['fapl'], atom_cap=self.reserve_atoms)
            self._status_h = torch.zeros(4, dtype=torch.int32).pin_memory()
            self._total_h = torch.zeros((self.max_systems, 10), dtype=torch.float64).pin_memory()
        return self.engine

    def _place(self, new):
        """Places newly admitted requests on the engine and commits every
        change since the last commit (their slices, the slots freed by
        retired requests)."""
        eng = self.engine
        for r in new:
            self.slot_of[id(r)] = eng.admit(r.nb, vel=r.vel if r.kind == 'md' else None, prepared=self._planned(r))
        eng.commit()
        self.stats['forms'] += 1

    def _launch(self):
        """Enqueues one quantum of iterations of every member, the copies of
        the results of the members that finish in it, the neighbour-list
        status and an event, then frees the slots of those members (their
        tables are rewritten only by later commits, after the copies in
        stream order).  Returns the in-flight record; nothing here waits for
        the device."""
        eng = self.engine
        nq = min(min(r.iters_total - r.iters_sched for r in self.active), self.max_quantum)
        dts_h = torch.zeros((nq, eng.nsys), dtype=torch.float64, pin_memory=True)
        dts = dts_h.numpy()
        for r in self.active:
            sl = self.slot_of[id(r)]
            for q in range(nq):
                dts[q, sl] = r.dt if (r.kind == 'md' and r.iters_sched + q > 0) else 0.0
        dts_d = dts_h.to(eng.device, non_blocking=True)
        start = torch.cuda.Event(enable_timing=True)
        start.record()
        for q in range(nq):
            eng.dt_d.copy_(dts_d[q])
            eng.md_step()
        status_h = torch.empty(4, dtype=torch.int32, pin_memory=True)
        status_h.copy_(eng.status_d, non_blocking=True)
        eng.status_d.zero_()      # flags of this quantum's builds only
        total_h = torch.empty((eng.nsys, 10), dtype=torch.float64, pin_memory=True)
        total_h.copy_(eng.total_d, non_blocking=True)
        fin = []
        members = list(self.active)
        for r in members:
            r.iters_sched += nq
            if r.iters_sched < r.iters_total:
                continue
            sl = self.slot_of[id(r)]
            a0 = eng.atom0[sl]
            bufs = {}
            names = ('gpos_d', 'pos_d', 'vel_d') if r.kind == 'md' else ('gpos_d',)
            for name in names:
                b = torch.empty((r.natom, 3), dtype=torch.float64, pin_memory=True)
                b.copy_(getattr(eng, name)[a0:a0 + r.natom], non_blocking=True)
                bufs[name] = b
            fin.append((r, sl, bufs))
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        for r, sl, _ in fin:
            eng.retire(sl)
            del self.slot_of[id(r)]
        if fin:
            gone = set(id(r) for r, _, _ in fin)
            self.active = [r for r in self.active if id(r) not in gone]
        self.stats['iterations'] += nq
        return dict(event=ev, start=start, nq=nq, fin=fin, members=members, status=status_h, total=total_h)

    def _complete(self, fl, now):
        """After a quantum's event: its finished requests get their results
        (already on the host).  Returns False when its neighbour-list builds
        overflowed (the caller restarts the affected requests)."""
        if int(fl['status'][0]):
            return False
        # device time of the quantum: from the stream reaching its start to its end
        self.stats['device_s'] += 1e-3*fl['start'].elapsed_time(fl['event'])
        for r in fl['members']:
            r.iters_done += fl['nq']
        tot = fl['total'].numpy()
        for r, sl, bufs in fl['fin']:
            r.energy = float(tot[sl, 0])
            r.vtens = tot[sl, 1:].reshape(3, 3).copy()
            r.gpos = bufs['gpos_d'].numpy().copy()
            if r.kind == 'md':
                r.pos = bufs['pos_d'].numpy().copy()
                r.vel_out = bufs['vel_d'].numpy().copy()
            r.t_done = now()
            r._plan = None    # its planning (page-locked copies of its arrays) is no longer needed
            self.done.append(r)
        return True

    def _recover(self, inflight, first):
        """An overflow in the quantum `first` of the in-flight records: wait
        for the device, grow the capacities the flags name, empty the batch
        and put every request of that quantum and the later ones back at the
        head of the queue, to run again from its inputs (every result is a
        function of the request alone, so the rerun gives the bits a run
        without the overflow gives)."""
        eng = self.engine
        torch.cuda.synchronize()
        flags = 0
        for fl in inflight:
            flags |= int(fl['status'][0])
        if flags & 1:
            longest = int(eng.cnt_d[:eng.npad].max().item())
            eng.grow(1, max(longest, 2*eng.kcap))
        if flags & 4:
            eng._fullest_bin = max(int(eng.status_d[1].item()), 2*eng.bin_cap)
            eng.grow(4, 0)
        if flags & 10:
            eng.grow(flags & 10, 0)
        eng.status_d.zero_()
        redo = {}
        for fl in inflight:
            for r in fl['members']:
                redo[id(r)] = r
        for r in self.active:
            eng.retire(self.slot_of.pop(id(r)))
        self.active = []
        for r in sorted(redo.values(), key=lambda r: (r.arrival, r.rid), reverse=True):
            r.iters_sched = r.iters_done = 0
            r.t_admit = None
            self.waiting.appendleft(r)
        eng.commit()
        self.stats['regrow'] += 1

    # -- driver ------------------------------------------------------------
    def run(self, requests, realtime=True):
        """Serves ``requests`` (arrival times in seconds from the start when
        realtime, else all at once) and returns them completed, in completion
        order.  Latency = completion - arrival, on the host clock, with the
        results on the host.  Up to `depth` quanta are in flight: the next
        quantum is placed and enqueued while the device runs the previous
        one; arrivals are planned by a thread pool meanwhile."""
        self._engine()
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0
        if not realtime:
            for r in pending:
                r.arrival = 0.0
        for r in pending:
            r.iters_sched = r.iters_done = 0
        self.done = []
        inflight = collections.deque()
        pool = concurrent.futures.ThreadPoolExecutor(self.planners)

        def arrivals():
            t = now()
            while pending and pending[0].arrival <= t:
                r = pending.popleft()
                r._plan = pool.submit(self.engine.prepare, r.nb)
                self.waiting.append(r)
        try:
            while pending or self.waiting or self.active or inflight:
                arrivals()
                if inflight and (len(inflight) >= self.depth or not (self.active or self.waiting)):
                    t1 = time.perf_counter()
                    while not inflight[0]['event'].query():
                        arrivals()
                        time.sleep(0)
                    self.stats['wait_s'] += time.perf_counter() - t1
                    if not self._complete(inflight[0], now):
                        self._recover(inflight, 0)
                        inflight.clear()
                    else:
                        inflight.popleft()
                    continue
                if not self.active and not self.waiting:
                    if pending:
                        wait = pending[0].arrival - now()
                        if wait > 0:
                            time.sleep(min(wait, 0.001))
                    continue
                t1 = time.perf_counter()
                new = self._admit(now())
                if new or self.engine.changed():
                    self._place(new)
                self.stats['form_s'] += time.perf_counter() - t1
                if self.active:
                    inflight.append(self._launch())
        finally:
            pool.shutdown()
        return self.done
