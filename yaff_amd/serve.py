"""Serving loop: continuous batching of force evaluations and MD segments.

Requests (a single energy/gradient/virial evaluation, or an MD segment of
n velocity-Verlet steps) arrive at any time.  The server keeps the systems of
the admitted requests resident on the device in ONE batch engine and advances
them together, one batch iteration at a time (kick/drift, neighbour list,
pair potentials, Ewald reciprocal, kick for every member in the same
launches):

* an evaluation needs one iteration (time step 0: the kicks and the drift do
  nothing, the forces are computed);
* an MD segment needs 1 + n iterations (the first, with time step 0, gives the
  initial forces; the next n are Verlet steps);
* between quanta of iterations, finished requests leave the batch and waiting
  ones join it: the batch is re-formed (per-system tables are cached, so this
  is a concatenation), and the positions, velocities and gradients of the
  members that stay are carried over on the device;
* admission is first come, first served, while the estimated device memory
  of the batch with the candidate stays within the budget and the batch has
  a free slot.

Launch parameters are fixed for the server (not re-tuned per batch), and every
kernel of the engine is deterministic and per-system, so a request gets the
same bits alone or in any mix.

The velocity-Verlet splitting follows yaff/sampling/verlet.py:339-367.
"""
import collections
import time

import numpy as np
import torch

from . import _lib
from .engine import Engine, _round_up, _ceil_div, ewald_kvectors_cached, ALIGN
from .batch import nbsystem, timestep, maxwell_boltzmann

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
        self.iters_done = 0
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


def estimate_bytes(nbs, tuning=SERVER_TUNING):
    """Device memory of a batch engine over ``nbs`` (the dominant allocations
    of Engine: per-atom state, the ELL neighbour list at the batch's row
    capacity, the Ewald tables and partial sums)."""
    if not nbs:
        return 0
    npad = sum(_round_up(max(s.natom, 1), ALIGN) for s in nbs)
    kcap = max(_kcap_estimate(s) for s in nbs)
    total = npad*(8*kcap + 8*3*3 + 8*3*_lib.MAX_POT + 8*3 + 4 + 2*_lib.MAX_POT*8 + 64)
    chunk = 64*(tuning.get('apl') or 4)
    fkb = tuning.get('fkb') or 64
    nslot = 0
    for s in nbs:
        if s.ewald is None or s.natom == 0:
            continue
        nk = len(ewald_kvectors_cached(s)[1])
        nkp = _round_up(max(nk, 1), 256)
        nchunk = _ceil_div(s.natom, chunk)
        total += nkp*(32 + 16 + 48 + 64) + 16*nchunk*nkp
        # greedy row blocks: two consecutive blocks hold more than fkb k vectors
        nslot = max(nslot, _ceil_div(2*_ceil_div(nk, fkb) + 1, 4))
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
    """Continuous-batching server of Requests on one device."""

    def __init__(self, precision='mixed', hbm_budget=64 << 30, max_systems=64, max_quantum=4,
                 tuning=SERVER_TUNING):
        self.precision = precision
        self.hbm_budget = int(hbm_budget)
        self.max_systems = int(max_systems)
        self.max_quantum = int(max_quantum)
        self.tuning = dict(tuning)
        self.active = []            # admitted requests, in batch order
        self.waiting = collections.deque()
        self.engine = None
        self.engine_members = []
        self.done = []
        self.stats = collections.Counter()
        self.peak_bytes = 0

    # -- admission ---------------------------------------------------------
    def _admit(self, now):
        changed = False
        while self.waiting and len(self.active) < self.max_systems:
            cand = self.waiting[0]
            need = estimate_bytes([r.nb for r in self.active] + [cand.nb], self.tuning)
            if need > self.hbm_budget:
                if not self.active:
                    raise RuntimeError('request %d alone needs %.1f GiB, over the budget of %.1f GiB'
                                       % (cand.rid, need/2.0**30, self.hbm_budget/2.0**30))
                break
            self.waiting.popleft()
            cand.t_admit = now
            self.active.append(cand)
            self.peak_bytes = max(self.peak_bytes, need)
            changed = True
        return changed

    # -- batch formation ----------------------------------------------------
    def _form(self):
        """Builds the engine of the current members; carries the device state
        of members that were in the previous batch."""
        t0 = time.perf_counter()
        old = self.engine
        old_members = self.engine_members
        eng = Engine([r.nb for r in self.active], precision=self.precision, upload=False,
                     pair_lanes=self.tuning['pair_lanes'], apl=self.tuning['apl'], skb=self.tuning['skb'],
                     fkb=self.tuning['fkb'], fapl=self.tuning['fapl'])
        host_pos = np.zeros((eng.npad, 3))
        host_vel = np.zeros((eng.npad, 3))
        src, dst = [], []
        old_index = {id(r): i for i, r in enumerate(old_members)}
        for i, r in enumerate(self.active):
            a0 = eng.atom0[i]
            j = old_index.get(id(r))
            if j is not None and old is not None and r.iters_done > 0:
                b0 = old.atom0[j]
                src.append(np.arange(b0, b0 + r.natom))
                dst.append(np.arange(a0, a0 + r.natom))
            else:
                host_pos[a0:a0 + r.natom] = r.nb.pos
                if r.vel is not None:
                    host_vel[a0:a0 + r.natom] = r.vel
        eng.pos_d.copy_(torch.from_numpy(host_pos))
        eng.vel_d.copy_(torch.from_numpy(host_vel))
        if src:
            si = torch.from_numpy(np.concatenate(src)).to(eng.device)
            di = torch.from_numpy(np.concatenate(dst)).to(eng.device)
            for name in ('pos_d', 'vel_d', 'gpos_d'):
                getattr(eng, name).index_copy_(0, di, getattr(old, name).index_select(0, si))
        self.engine = eng
        self.engine_members = list(self.active)
        self.stats['forms'] += 1
        self.stats['form_s'] += time.perf_counter() - t0

    # -- iterations --------------------------------------------------------
    def _quantum(self):
        eng = self.engine
        nq = min(min(r.iters_total - r.iters_done for r in self.active), self.max_quantum)
        dts = np.zeros((nq, eng.nsys))
        for i, r in enumerate(self.active):
            for q in range(nq):
                it = r.iters_done + q
                dts[q, i] = r.dt if (r.kind == 'md' and it > 0) else 0.0
        dts_d = torch.from_numpy(dts).to(eng.device)
        snap = [t.clone() for t in (eng.pos_d, eng.vel_d, eng.gpos_d)]
        for attempt in range(3):
            for q in range(nq):
                eng.dt_d.copy_(dts_d[q])
                eng.md_step()
            flags, longest = eng.nlist_status()   # synchronises
            if not flags:
                break
            # a row or a bin overflowed: grow and redo the quantum from the snapshot
            eng.grow(flags, longest)
            for t, v in zip((eng.pos_d, eng.vel_d, eng.gpos_d), snap):
                t.copy_(v)
            self.stats['regrow'] += 1
        else:
            raise RuntimeError('neighbour list capacities still exceeded')
        self.stats['iterations'] += nq
        for r in self.active:
            r.iters_done += nq
        return nq

    def _retire(self, now):
        eng = self.engine
        fin = [i for i, r in enumerate(self.active) if r.iters_done >= r.iters_total]
        if not fin:
            return False
        tot = eng.total_d.cpu().numpy()
        for i in fin:
            r = self.active[i]
            a0 = eng.atom0[i]
            r.energy = float(tot[i, 0])
            r.vtens = tot[i, 1:].reshape(3, 3).copy()
            r.gpos = eng.gpos_d[a0:a0 + r.natom].cpu().numpy()
            if r.kind == 'md':
                r.pos = eng.pos_d[a0:a0 + r.natom].cpu().numpy()
                r.vel_out = eng.vel_d[a0:a0 + r.natom].cpu().numpy()
            r.t_done = now()
            self.done.append(r)
        keep = [r for i, r in enumerate(self.active) if i not in set(fin)]
        self.active = keep
        return True

    # -- driver ------------------------------------------------------------
    def run(self, requests, realtime=True):
        """Serves ``requests`` (arrival times in seconds from the start when
        realtime, else all at once) and returns them completed, in completion
        order.  Latency = completion - arrival, on the host clock, with the
        results on the host."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0
        if not realtime:
            for r in pending:
                r.arrival = 0.0
        self.done = []
        while pending or self.waiting or self.active:
            t = now()
            while pending and pending[0].arrival <= t:
                self.waiting.append(pending.popleft())
            if not self.active and not self.waiting:
                wait = pending[0].arrival - now()
                if wait > 0:
                    time.sleep(min(wait, 0.001))
                continue
            self._admit(t)
            if self.engine is None or [id(r) for r in self.engine_members] != [id(r) for r in self.active]:
                self._form()
            self._quantum()
            self._retire(now)
            if not self.active:
                self.engine = None
                self.engine_members = []
        return self.done
