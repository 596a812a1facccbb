"""Persistent batch engine: systems join and leave without a new engine.

The engine owns device tables sized by capacities that only grow.  Every
system occupies one of ``nslots`` slots and ranges of the batch's index
spaces, each handed out by a first-fit allocator:

    atoms     padded atom ranges (multiples of 256 atoms)
    bins      cell bins of the neighbour-list build
    k         k vectors (multiples of 256)
    rows      k-vector rows
    skblk     row blocks of the structure-factor pass
    fblk      row blocks of the force pass
    chunk     atom chunks of the structure-factor pass
    work      structure-factor work items
    rec       force records (rows padded to 4, plus 4 spare zero records)
    spart     partial structure factors (chunk x k entries)

Admission places the system in free ranges, writes only its own slices of
the device tables and its slot's ynb_sys; retirement marks its ranges free
(owner -1 in the wavefront, bin, k-block and work-item tables: every kernel
skips them) and empties its slot.  Launches cover the used prefix of each
space.  The small tables that are concatenations over systems (scaling rows,
ExpRep cross tables, the Ewald-correction tables) are rebuilt on a change.

Every kernel computes per system, in an order fixed by the system alone and
the engine's launch parameters, so a system gets the same bits in any slot,
at any offset and next to any other systems.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import YnbSys, YnbBatch, YnbNlist, YnbPairpar, YnbEwald, MAX_POT, SCALING_DTYPE, check
from .engine import (Engine, CorrTables, corr_plan, system_entry, shift_offsets, _ewald_local_plan, _round_up,
                     _ceil_div, ALIGN, KBLOCK)

__all__ = ['Ranges', 'SlotEngine']

GROWTH = 2.0


class Ranges(object):
    """First-fit allocator over [0, cap) in units; every range starts at a
    multiple of ``align`` and holds a multiple of it."""

    def __init__(self, cap, align=1):
        self.cap, self.align = int(cap), int(align)
        self.used = []    # sorted (offset, size)

    def size_of(self, n):
        return _round_up(max(int(n), 1), self.align)

    def alloc(self, n):
        """Offset of a new range of n units, or None when none fits."""
        n = self.size_of(n)
        pos = 0
        for i, (o, m) in enumerate(self.used):
            if o - pos >= n:
                self.used.insert(i, (pos, n))
                return pos
            pos = o + m
        if self.cap - pos >= n:
            self.used.append((pos, n))
            return pos
        return None

    def need(self, n):
        """Capacity with which alloc(n) succeeds (at the end if not before)."""
        return max(self.cap, self.high + self.size_of(n))

    def free(self, off):
        for i, (o, _) in enumerate(self.used):
            if o == off:
                del self.used[i]
                return
        raise KeyError(off)

    @property
    def high(self):
        """End of the last range in use."""
        return self.used[-1][0] + self.used[-1][1] if self.used else 0


SEG_BYTES = 1 << 16   # bytes per scatter segment: one block each
_NP_DTYPE = {torch.float64: np.float64, torch.int32: np.int32, torch.int64: np.int64}


def _pinned(nbytes):
    """A page-locked host byte buffer (torch's caching host allocator keeps it
    from reuse until the copies enqueued from it have run)."""
    return torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)


class _Stager(object):
    """Copies and fills of device tables collected for one ynb_scatter
    launch.  A destination is an engine attribute (a device tensor) and an
    element offset, resolved when the batch is sent, after any reallocation
    of the tables."""

    def __init__(self):
        self.ops = []
        self.blobs = []

    def put_blob(self, buf, items):
        """Copies out of one page-locked float64 host tensor buf: items are
        (attr, element offset in attr, element offset in buf, elements).
        The buffer goes to the device staging tensor with one asynchronous
        copy of its own, after the staged data, and the scatter launch
        places its pieces."""
        if buf.numel():
            self.blobs.append((buf, list(items)))

    def put(self, attr, off, arr):
        a = np.ascontiguousarray(arr)
        if a.size:
            self.ops.append((attr, int(off), a.reshape(-1).view(np.uint8), None))

    def fill(self, attr, off, n, value):
        if n > 0:
            self.ops.append((attr, int(off), int(n), value))

    def flush(self, eng):
        """One pinned buffer (segment table, then data) -> one copy to the
        engine's staging tensor, each put_blob buffer -> one copy after it ->
        one ynb_scatter launch."""
        if not self.ops and not self.blobs:
            return
        segs, blobs = [], []
        off = 0
        for attr, o, data, value in self.ops:
            t = getattr(eng, attr)
            es = t.element_size()
            elem = 8 if es in (1, 8) else 4
            dst = t.data_ptr() + o*es
            if value is None:
                nbytes = len(data)
                if (nbytes % elem) or o*es + nbytes > t.numel()*es:
                    raise ValueError('staged copy into %s out of range or not in %d-byte elements' % (attr, elem))
                for c in range(0, nbytes, SEG_BYTES):
                    segs.append((dst + c, off + c, min(SEG_BYTES, nbytes - c), 0, elem))
                blobs.append((off, data))
                off += _round_up(nbytes, 8)
            else:
                nbytes = data*es
                if o*es + nbytes > t.numel()*es:
                    raise ValueError('staged fill of %s out of range' % attr)
                if es not in (4, 8):
                    raise ValueError('staged fills need 4- or 8-byte elements (%s)' % attr)
                bits = int(np.array(value, dtype=_NP_DTYPE[t.dtype]).view({4: np.uint32, 8: np.uint64}[es]))
                for c in range(0, nbytes, SEG_BYTES):
                    segs.append((dst + c, -1, min(SEG_BYTES, nbytes - c), bits, elem))
        # put_blob buffers after the staged data (offsets relative to the data start)
        boff = _round_up(off, 256)
        placed = []
        for buf, items in self.blobs:
            for attr, o, so, n in items:
                t = getattr(eng, attr)
                if t.dtype != torch.float64 or o + n > t.numel():
                    raise ValueError('blob copy into %s out of range or not float64' % attr)
                dst, nbytes = t.data_ptr() + 8*o, 8*n
                for c in range(0, nbytes, SEG_BYTES):
                    segs.append((dst + c, boff + 8*so + c, min(SEG_BYTES, nbytes - c), 0, 8))
            placed.append((boff, buf))
            boff += _round_up(8*buf.numel(), 256)
        table = np.zeros(len(segs), _lib.SEGMENT_DTYPE)
        cols = list(zip(*segs))
        for k, name in enumerate(('dst', 'src', 'nbytes', 'fill', 'elem')):
            table[name] = np.array(cols[k], dtype=table.dtype[name])
        # the blocks of one launch run in any order: no two segments may overlap
        order = np.argsort(table['dst'], kind='stable')
        d, n = table['dst'][order], table['nbytes'][order].astype(np.uint64)
        if len(d) > 1 and np.any(d[:-1] + n[:-1] > d[1:]):
            raise ValueError('staged table updates overlap')
        head = _round_up(table.nbytes, 256)
        pinned = _pinned(head + off)
        buf = pinned.numpy()
        buf[:table.nbytes] = table.view(np.uint8)
        for o, data in blobs:
            buf[head + o:head + o + len(data)] = data
        if eng._stage_d.numel() < head + boff:
            eng._stage_d = torch.empty(_round_up(int((head + boff)*GROWTH), 256), dtype=torch.uint8,
                                       device=eng.device)
        eng._stage_d[:len(buf)].copy_(pinned, non_blocking=True)
        for o, b in placed:
            eng._stage_d[head + o:head + o + 8*b.numel()].view(torch.float64).copy_(b, non_blocking=True)
        base = eng._stage_d.data_ptr()
        _lib.check(eng.lib.ynb_scatter(base, len(table), base + head, eng.stream()))
        self.ops = []
        self.blobs = []


def _pin_member_arrays(s, entry, plan):
    """The arrays of a member that do not depend on its placement
    (positions, per-atom records and parameter rows, charges, masses and
    their inverses, k vectors), copied into one page-locked buffer: the
    planning threads pay for the host copies, and the commit sends the
    buffer with one asynchronous copy.  Returns the float64 buffer and
    name -> (element offset, elements)."""
    n = s.natom
    masses = np.ones(n) if s.masses is None else np.asarray(s.masses, dtype=np.float64)
    items = [('pos', s.pos), ('rec', entry.rec)]
    par = entry.par.reshape(MAX_POT*3, n)
    items += [('par%d' % row, par[row]) for row in range(3*len(s.pots))]
    items += [('charges', np.zeros(n) if s.charges is None else s.charges), ('masses', masses),
              ('minv', 1.0/masses)]
    if plan is not None:
        items.append(('kvec', plan['kpad']))
    arrs = [(name, np.ascontiguousarray(a, dtype=np.float64).reshape(-1)) for name, a in items]
    buf = torch.empty(sum(a.size for _, a in arrs), dtype=torch.float64, pin_memory=True)
    host = buf.numpy()
    where, off = {}, 0
    for name, a in arrs:
        host[off:off + a.size] = a
        where[name] = (off, a.size)
        off += a.size
    return buf, where


def _nlist_kcap(s):
    """Row capacity that fits a system's full-list rows (with a margin)."""
    if s.natom == 0:
        return 16
    if s.cell.nvec == 3:
        n = 4.0/3.0*np.pi*s.rc**3*s.natom/s.cell.volume
    else:
        n = s.natom*(int(np.prod(2*s.rmax + 1)) if len(s.rmax) else 1)
    return int(1.3*n) + 32


class _Member(object):
    """A system in a slot: its entry (offsets of its ranges filled in, scaling
    and cross offsets system-local) and its ranges."""

    def __init__(self, s, entry, ranges, ew_plan, kcap):
        self.s, self.entry, self.ranges, self.ew_plan, self.kcap = s, entry, ranges, ew_plan, kcap


class SlotEngine(Engine):
    """A batch Engine whose members change in place (admit / retire, then
    commit) instead of a new Engine per membership.  Launch parameters are
    fixed at construction.  Results, MD steps, neighbour-list status and
    capacity growth work as in Engine, over the members."""

    def __init__(self, nslots=64, precision='mixed', pair_lanes=32, apl=4, skb=64, fkb=64, fapl=4, device='cuda',
                 atom_cap=ALIGN):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError('yaff_amd.SlotEngine needs a HIP device (torch.cuda.is_available() is False)')
        if apl not in (2, 4, 8) or fapl not in (1, 2, 4) or pair_lanes not in (16, 32, 64):
            raise ValueError('apl must be 2, 4 or 8; fapl 1, 2 or 4; pair_lanes 16, 32 or 64')
        self.device = torch.device(device)
        self.precision = {'fp64': _lib.FP64, 'mixed': _lib.MIXED}[precision] if isinstance(precision, str) \
            else int(precision)
        self.nsys = int(nslots)
        self.pair_lanes, self.apl, self.skb, self.fkb, self.fapl = int(pair_lanes), int(apl), int(skb), int(fkb), \
            int(fapl)
        self.systems = [None]*self.nsys
        self.atom0 = [None]*self.nsys
        self.members = [None]*self.nsys
        self.hsys = (YnbSys*self.nsys)()
        # small starting capacities: the first members size the spaces
        self.sp = {'atoms': Ranges(_round_up(atom_cap, ALIGN), ALIGN), 'bins': Ranges(64),
                   'k': Ranges(KBLOCK, KBLOCK), 'rows': Ranges(64), 'skblk': Ranges(64),
                   'fblk': Ranges(64), 'chunk': Ranges(16), 'work': Ranges(256), 'rec': Ranges(256, 4),
                   'spart': Ranges(4096)}
        self.kcap = 64
        self.bin_cap = 16
        self.npad = ALIGN
        self.nbin = 0
        self.fslots_cap = 1
        self.nbuilds = 0
        self.gpos_slots_d = None
        self.corr_tables = None
        self.has_ewald = self.has_corr = self.use_skin = False
        self.uncorrected = []
        self.groups = 0
        self.fslots = 0
        self._pending = []
        self._stage = _Stager()
        self._stage_free = _Stager()
        self._work_gen = []     # structure-factor work items written on the device at the commit
        self._alloc_fixed()
        self._alloc_atoms(self.sp['atoms'].cap)
        self._alloc_bins(self.sp['bins'].cap)
        self._alloc_k(self.sp['k'].cap)
        for name in ('rows', 'skblk', 'fblk', 'chunk', 'work', 'rec', 'spart'):
            self._alloc_space(name, self.sp[name].cap)
        self._alloc_gpart(self.fslots_cap)
        self._commit_structs()

    # ---- allocation of the capacity-sized tables ------------------------------
    def _z(self, n, dtype):
        return torch.zeros(n, dtype=dtype, device=self.device)

    @staticmethod
    def _grown(old, new_numel, fill=0):
        """A tensor of new_numel elements that starts with old's content."""
        t = torch.full((new_numel,), fill, dtype=old.dtype, device=old.device) if fill else \
            torch.zeros(new_numel, dtype=old.dtype, device=old.device)
        n = min(old.numel(), new_numel)
        t[:n].copy_(old.reshape(-1)[:n])
        return t

    def _alloc_fixed(self):
        n = self.nsys
        f64 = torch.float64
        self.sys_d = self._z(ctypes.sizeof(YnbSys)*n, torch.uint8)
        self.pair_out_d = torch.zeros((n, MAX_POT, 10), dtype=f64, device=self.device)
        self.ewald_out_d = torch.zeros((n, 10), dtype=f64, device=self.device)
        self.total_d = torch.zeros((n, 10), dtype=f64, device=self.device)
        self.dt_d = self._z(n, f64)
        self.ekin2_d = self._z(n, f64)
        self.skin_d = self._z(n, f64)
        self.rebuild_d = torch.ones(n, dtype=torch.int32, device=self.device)
        self.status_d = self._z(4, torch.int32)
        self.cross_d = self._z(1, f64)
        self.stab_d = self._z(SCALING_DTYPE.itemsize, torch.uint8)
        self._stage_d = self._z(1 << 16, torch.uint8)

    def _alloc_atoms(self, cap):
        """Per-atom tables at capacity cap (state carried over on growth)."""
        old = getattr(self, 'pos_d', None) is not None
        f64 = torch.float64
        if old:
            self.pos_d = self._grown(self.pos_d, 3*cap).view(cap, 3)
            self.vel_d = self._grown(self.vel_d, 3*cap).view(cap, 3)
            self.gpos_d = self._grown(self.gpos_d, 3*cap).view(cap, 3)
            self.pos_ref_d = self._grown(self.pos_ref_d, 3*cap, 1e150).view(cap, 3)
            self.rec_d = self._grown(self.rec_d, 4*cap).view(cap, 4)
            self.charges_d = self._grown(self.charges_d, cap)
            self.masses_d = self._grown(self.masses_d, cap, 1.0)
            self.minv_d = self._grown(self.minv_d, cap, 1.0)
            self.wave_sys_d = self._grown(self.wave_sys_d, cap//64, -1)
            self.par_d = self._grown(self.par_d, MAX_POT*3*cap)
        else:
            self.pos_d = torch.zeros((cap, 3), dtype=f64, device=self.device)
            self.vel_d = torch.zeros((cap, 3), dtype=f64, device=self.device)
            self.gpos_d = torch.zeros((cap, 3), dtype=f64, device=self.device)
            self.pos_ref_d = torch.full((cap, 3), 1e150, dtype=f64, device=self.device)
            self.rec_d = torch.zeros((cap, 4), dtype=f64, device=self.device)
            self.charges_d = self._z(cap, f64)
            self.masses_d = torch.ones(cap, dtype=f64, device=self.device)
            self.minv_d = torch.ones(cap, dtype=f64, device=self.device)
            self.wave_sys_d = torch.full((cap//64,), -1, dtype=torch.int32, device=self.device)
            self.par_d = self._z(MAX_POT*3*cap, f64)
        self.acap = cap
        self.sp['atoms'].cap = cap
        self.cnt_d = self._z(cap, torch.int32)
        self.partial_d = self._z((cap//4)*MAX_POT*8, f64)
        self.md_partial_d = self._z(cap//64, f64)
        # [nsys] pair, [nsys] unused, [nsys] kick, [cap/64] Ewald force blocks (zero, left at zero)
        self.ctr_d = self._z(3*self.nsys + cap//64 + 1, torch.int32)
        self.ew_frac_d = self._z(4*cap, f64)
        self.ew_step_d = self._z(4*cap, f64)
        self._alloc_nlist(self.kcap)
        self._alloc_nlist_work()
        if hasattr(self, 'ew_gpart_d'):
            self._alloc_gpart(self.fslots_cap)

    def _alloc_nlist(self, kcap):
        self.kcap = int(kcap)
        self.nbr_j_d = self._z(self.kcap*self.acap, torch.int32)
        self.nbr_img_d = self._z(self.kcap*self.acap, torch.int32)
        self.pos_ref_d.fill_(1e150)   # new capacity: the next skin test rebuilds everything

    def _alloc_nlist_work(self):
        # sized for the capacities; zeroed (the build keeps its bin counters at zero)
        self.nl_work_bytes = self.lib.ynb_nlist_workspace_bytes(self.acap, self.sp['bins'].cap, self.bin_cap)
        self.nl_work_d = self._z(self.nl_work_bytes, torch.uint8)

    def _alloc_bins(self, cap):
        old = getattr(self, 'bin_sys_d', None)
        self.bin_sys_d = self._grown(old, cap, -1) if old is not None else \
            torch.full((cap,), -1, dtype=torch.int32, device=self.device)
        self.sp['bins'].cap = cap
        if hasattr(self, 'acap'):
            self._alloc_nlist_work()

    def _alloc_k(self, cap):
        old = getattr(self, 'kvec_d', None) is not None
        f64 = torch.float64
        if old:
            self.kvec_d = self._grown(self.kvec_d, 4*cap)
            self.kpos_d = self._grown(self.kpos_d, cap, -1)
            self.kblk_sys_d = self._grown(self.kblk_sys_d, cap//KBLOCK, -1)
        else:
            self.kvec_d = self._z(4*cap, f64)
            self.kpos_d = torch.full((cap,), -1, dtype=torch.int32, device=self.device)
            self.kblk_sys_d = torch.full((cap//KBLOCK,), -1, dtype=torch.int32, device=self.device)
        self.ew_coef_d = self._z(6*cap, f64)
        self.ew_kpart_d = self._z(8*(cap//32), f64)
        self.sp['k'].cap = cap

    _SPACE_TABLES = {'rows': (('krow_d', 4, 0), ('krow_k_d', 1, 0), ('krow_f_d', 1, 0)),
                     'skblk': (('skblk_d', 2, 0),), 'fblk': (('fblk_d', 2, 0),), 'chunk': (('chunk_d', 4, 0),),
                     'work': (('work_d', 4, -1),)}

    def _alloc_space(self, name, cap):
        if name == 'rec':
            # 2 doubles per record; a system's padding records must read zero
            old = getattr(self, 'ew_frec_d', None)
            self.ew_frec_d = self._grown(old, 2*cap) if old is not None else self._z(2*cap, torch.float64)
        elif name == 'spart':
            self.ew_spart_d = self._z(2*cap, torch.float64)
        else:
            for attr, width, fill in self._SPACE_TABLES[name]:
                old = getattr(self, attr, None)
                if old is not None:
                    t = self._grown(old, width*cap, fill)
                else:
                    t = torch.full((width*cap,), fill, dtype=torch.int32, device=self.device) if fill else \
                        self._z(width*cap, torch.int32)
                setattr(self, attr, t)
        self.sp[name].cap = cap

    def _alloc_gpart(self, nslot):
        self.fslots_cap = int(nslot)
        self.ew_gpart_d = self._z(self.fslots_cap*self.acap*3, torch.float64)

    def _take(self, name, n):
        """Offset of a new range of n units of space `name`; the space grows
        (by GROWTH at least, tables carried over) when no free range fits."""
        r = self.sp[name]
        off = r.alloc(n)
        if off is not None:
            return off
        cap = _round_up(max(r.need(n), int(r.cap*GROWTH)), r.align)
        if name == 'atoms':
            self._alloc_atoms(cap)
        elif name == 'bins':
            self._alloc_bins(cap)
        elif name == 'k':
            self._alloc_k(cap)
        else:
            self._alloc_space(name, cap)
        self._commit_structs()   # launches before the next commit see the moved tables
        return r.alloc(n)

    def device_bytes(self):
        """Device memory held by the engine's tables."""
        total = 0
        for v in self.__dict__.values():
            if isinstance(v, torch.Tensor) and v.is_cuda:
                total += v.numel()*v.element_size()
        if self.corr_tables is not None:
            for v in self.corr_tables.__dict__.values():
                if isinstance(v, torch.Tensor):
                    total += v.numel()*v.element_size()
        return total

    # ---- membership -----------------------------------------------------------
    def free_slots(self):
        return [i for i in range(self.nsys) if self.members[i] is None]

    def prepare(self, s):
        """The host-side planning of NBSystem s that does not depend on the
        engine's state (its ynb_sys before placement, its Ewald tables, its
        row capacity); safe to run in another thread while the engine works,
        one call per system."""
        if s.ewald is not None and s.natom > 0 and s.cell.nvec != 3:
            raise ValueError('Ewald reciprocal needs a 3D periodic cell')
        entry = system_entry(self.lib, s, 0)
        plan = _ewald_local_plan(s, self.apl, self.skb, self.fkb) if s.ewald is not None and s.natom > 0 else None
        return entry, plan, _nlist_kcap(s), _pin_member_arrays(s, entry, plan)

    def admit(self, s, vel=None, prepared=None):
        """Places NBSystem s (planned by prepare(s), which admit calls when
        ``prepared`` is None) in a free slot, growing capacities as needed,
        and returns the slot.  Its tables reach the device at the next
        commit(); its positions, velocities (default zero) and gradient
        (zero) then start its MD state."""
        free = self.free_slots()
        if not free:
            raise RuntimeError('no free slot')
        slot = free[0]
        entry, plan, kcap, pinned = self.prepare(s) if prepared is None else prepared
        ranges = {}
        napad = _round_up(max(s.natom, 1), ALIGN)
        ranges['atoms'] = (self._take('atoms', napad), napad)
        h = entry.h
        h.atom0 = ranges['atoms'][0]
        if h.nbin > 0:
            ranges['bins'] = (self._take('bins', h.nbin), h.nbin)
            h.bin0 = ranges['bins'][0]
        if plan is not None:
            sizes = {'k': plan['nkp'], 'rows': plan['nrow'], 'skblk': len(plan['skblk']),
                     'fblk': len(plan['fblk']), 'chunk': len(plan['chunks']), 'work': plan['nwork'],
                     'rec': plan['nrec'] + 4, 'spart': len(plan['chunks'])*plan['nkp']}
            for name, n in sizes.items():
                ranges[name] = (self._take(name, n), n)
            ew = s.ewald
            h.alpha, h.dielectric = ew.alpha, ew.dielectric
            h.k0, h.nk = ranges['k'][0], plan['nkp']
            h.frb0, h.nfrb = ranges['fblk'][0], len(plan['fblk'])
            h.nslot = _ceil_div(h.nfrb, 4)
            h.chunk0, h.nchunk = ranges['chunk'][0], len(plan['chunks'])
            h.spart0 = ranges['spart'][0]
            h.nchunk_low, h.nchunk_high = plan['nchunk_low'], plan['nchunk_high']
        m = _Member(s, entry, ranges, plan, kcap)
        m.pinned = pinned
        self.members[slot] = m
        self.systems[slot] = s
        self.atom0[slot] = ranges['atoms'][0]
        self._pending.append((slot, vel))
        return slot

    def changed(self):
        """Whether admissions or retirements wait for a commit()."""
        return bool(self._pending or self._stage_free.ops)

    def retire(self, slot):
        """Frees slot's ranges and empties the slot (takes effect on the
        device at the next commit())."""
        m = self.members[slot]
        if m is None:
            raise KeyError(slot)
        self._pending = [(sl, v) for sl, v in self._pending if sl != slot]
        for name, (off, n) in m.ranges.items():
            self.sp[name].free(off)
        a0, napad = m.ranges['atoms']
        st = self._stage_free   # sent before the new members' slices, which may reuse these ranges
        st.fill('wave_sys_d', a0//64, napad//64, -1)
        if 'bins' in m.ranges:
            b0, nb = m.ranges['bins']
            st.fill('bin_sys_d', b0, nb, -1)
        if 'k' in m.ranges:
            k0, nk = m.ranges['k']
            st.fill('kblk_sys_d', k0//KBLOCK, nk//KBLOCK, -1)
            w0, nw = m.ranges['work']
            st.fill('work_d', 4*w0, 4*nw, -1)
        self.members[slot] = None
        self.systems[slot] = None
        self.atom0[slot] = None
        self.hsys[slot] = YnbSys()

    # ---- commit ---------------------------------------------------------------
    def _up(self, dst, arr):
        """Asynchronous upload of a host array (same element type and number
        of elements) into a device tensor or slice, through pinned memory."""
        a = np.ascontiguousarray(arr)
        if a.size == 0:
            return
        if not a.flags.writeable:
            a = a.copy()
        src = torch.from_numpy(a).pin_memory()
        dst.copy_(src.reshape(dst.shape), non_blocking=True)

    def commit(self):
        """Sends the changes since the last commit to the device: the owner
        entries of freed ranges (one scatter launch), then new members'
        slices, the system table and the concatenated small tables (a
        second), and the launch extents.  Asynchronous on the current
        stream."""
        self._stage_free.flush(self)
        act = [i for i in range(self.nsys) if self.members[i] is not None]
        npad = max(ALIGN, self.sp['atoms'].high)
        nbin = self.sp['bins'].high
        kcap = max([self.kcap] + [self.members[i].kcap for i in act])
        if kcap > self.kcap:
            self._alloc_nlist(_round_up(kcap, 16))
        cb = self.bin_cap
        for i in act:
            h = self.members[i].entry.h
            if h.nbin > 0:
                cb = max(cb, int(2*self.members[i].s.natom/h.nbin) + 16)
        cb = min(_round_up(cb, 16), 1024)
        if cb != self.bin_cap:
            self.bin_cap = cb
            self._alloc_nlist_work()
        elif nbin > self.nbin:
            # the workspace starts with one counter per bin (zero between
            # builds); the counters of bins past the old extent lay in the
            # region that followed them
            self.nl_work_d[4*self.nbin:4*nbin].zero_()
        if npad != self.npad:
            # par is [MAX_POT][3][npad]: move it to the new stride
            m = min(npad, self.npad)
            old = self.par_d[:MAX_POT*3*self.npad].view(MAX_POT*3, self.npad)[:, :m].clone()
            self.par_d[:MAX_POT*3*npad].view(MAX_POT*3, npad)[:, :m].copy_(old)
        self.npad, self.nbin = npad, nbin
        for slot, vel in self._pending:
            self._write_member(slot, vel)
        self._pending = []
        # concatenated tables, system table
        stabs, cross = [], []
        stab_off = cross_off = 0
        flags = 0
        fslots = 0
        self.has_ewald = self.has_corr = self.use_skin = False
        self.uncorrected = []
        self.groups = 0
        corr_entries = [None]*self.nsys
        for i in act:
            m = self.members[i]
            s, e = m.s, m.entry
            h = YnbSys.from_buffer_copy(e.h)
            shift_offsets(h, stab_off, cross_off)
            stabs.extend(e.stabs)
            stab_off += sum(len(st) for st in e.stabs)
            cross.extend(e.cross)
            cross_off += sum(len(c) for c in e.cross)
            self.hsys[i] = h
            flags |= self._sys_flags(h)
            if s.skin > 0:
                self.use_skin = True
            if s.ewald is not None:
                self.has_ewald = True
                fslots = max(fslots, h.nslot)
                if m.ew_plan is not None and m.ew_plan['grouped']:
                    self.groups = 1
                if s.ewald.correction and len(s.stab) > 0:
                    self.has_corr = True
                    corr_entries[i] = dict(atom0=self.atom0[i], natom=s.natom, alpha=s.ewald.alpha,
                                           dielectric=s.ewald.dielectric, nlow=s.ewald.nlow, nhigh=s.ewald.nhigh,
                                           stab=s.stab, charges=s.charges)
                elif not s.ewald.correction and len(s.stab) > 0:
                    self.uncorrected.append(i)
        stg = self._stage
        stg.put('sys_d', 0, np.frombuffer(bytes(self.hsys), np.uint8))
        if stabs:
            st = np.concatenate(stabs)
            if self.stab_d.numel() < st.nbytes:
                self.stab_d = torch.empty(st.nbytes, dtype=torch.uint8, device=self.device)
            stg.put('stab_d', 0, st)
        if cross:
            c = np.concatenate(cross)
            if self.cross_d.numel() < len(c):
                self.cross_d = torch.empty(len(c), dtype=torch.float64, device=self.device)
            stg.put('cross_d', 0, c)
        stg.flush(self)
        for w0, nblk, nch, slot, sb0, c0 in self._work_gen:
            if 4*(w0 + nblk*nch) > self.work_d.numel():
                raise RuntimeError('work items beyond the work table')
            check(self.lib.ynb_ewald_work_device(self.work_d.data_ptr() + 16*w0, nblk, nch, slot, sb0, c0,
                                                 self.stream()))
        self._work_gen = []
        self.corr_tables = CorrTables(corr_plan(corr_entries, npad), self.device) if self.has_corr else None
        if fslots > self.fslots_cap:
            self._alloc_gpart(fslots)
        self.fslots = fslots
        self.batch_flags = flags
        self._commit_structs()

    @staticmethod
    def _sys_flags(h):
        f = 0
        if h.natom > 0 and h.nbin == 0:
            f |= _lib.BATCH_ALLPAIRS
        if h.nbin > 0 and not h.box_safe:
            f |= _lib.BATCH_EXACT_BUILD
        if h.nbin > 0 and _lib.load().ynb_build_search_bins(_lib.C.byref(h)) > _lib.BUILD_NB_SMALL:
            f |= _lib.BATCH_WIDE_SEARCH
        f |= {_lib.FORM_GENERIC: _lib.BATCH_FORM_GENERIC, _lib.FORM_LJ: _lib.BATCH_FORM_LJ,
              _lib.FORM_IONIC: _lib.BATCH_FORM_IONIC}[h.pair_form]
        return f

    def _write_member(self, slot, vel):
        """Device slices of a newly admitted member (staged for the commit's
        scatter launch)."""
        m = self.members[slot]
        s, e = m.s, m.entry
        st = self._stage
        a0, napad = m.ranges['atoms']
        n = s.natom
        pbuf, pin = m.pinned   # the slot-independent arrays, page-locked by prepare()
        m.pinned = None
        blob = [('pos_d', 3*a0) + pin['pos']]
        if vel is not None:
            st.put('vel_d', 3*a0, np.asarray(vel, dtype=np.float64))
        else:
            st.fill('vel_d', 3*a0, 3*napad, 0.0)
        st.fill('gpos_d', 3*a0, 3*napad, 0.0)
        st.fill('pos_ref_d', 3*a0, 3*napad, 1e150)
        blob.append(('rec_d', 4*a0) + pin['rec'])
        for row in range(3*len(s.pots)):   # the kernels read no row of a potential the system lacks
            blob.append(('par_d', row*self.npad + a0) + pin['par%d' % row])
        blob += [('charges_d', a0) + pin['charges'], ('masses_d', a0) + pin['masses'],
                 ('minv_d', a0) + pin['minv']]
        ws = np.full(napad//64, -1, np.int32)
        ws[:(max(n, 1) + 63)//64] = slot
        st.put('wave_sys_d', a0//64, ws)
        st.put('skin_d', slot, np.array([s.skin], np.float64))
        st.fill('rebuild_d', slot, 1, 1)
        if 'bins' in m.ranges:
            b0, nb = m.ranges['bins']
            st.fill('bin_sys_d', b0, nb, slot)
        p = m.ew_plan
        if p is not None:
            k0 = m.ranges['k'][0]
            r0 = m.ranges['rows'][0]
            sb0 = m.ranges['skblk'][0]
            fb0 = m.ranges['fblk'][0]
            c0 = m.ranges['chunk'][0]
            w0 = m.ranges['work'][0]
            rec0, nrec = m.ranges['rec']
            nkp = p['nkp']
            blob.append(('kvec_d', 4*k0) + pin['kvec'])
            st.put('kpos_d', k0, np.where(p['kpos'] >= 0, p['kpos'] + rec0, -1).astype(np.int32))
            st.fill('kblk_sys_d', k0//KBLOCK, nkp//KBLOCK, slot)
            if p['nrow']:
                st.put('krow_d', 4*r0, p['krow'])
                st.put('krow_k_d', r0, (p['krow_k'] + k0).astype(np.int32))
                st.put('krow_f_d', r0, (p['krow_f'] + rec0).astype(np.int32))
            st.put('skblk_d', 2*sb0, (p['skblk'] + r0).astype(np.int32))
            st.put('fblk_d', 2*fb0, (p['fblk'] + r0).astype(np.int32))
            ch = p['chunks'].copy()
            ch[:, 0] = slot
            st.put('chunk_d', 4*c0, ch)
            self._work_gen.append((w0, len(p['skblk']), len(p['chunks']), slot, sb0, c0))
            st.fill('ew_frec_d', 2*rec0, 2*nrec, 0.0)
        st.put_blob(pbuf, blob)

    def _commit_structs(self):
        """The ctypes structures of the launches, pointing at the current tables."""
        self.batch = YnbBatch(self.sys_d.data_ptr(), self.wave_sys_d.data_ptr(), self.bin_sys_d.data_ptr(),
                              self.nsys, self.nbin, self.npad, self.bin_cap, getattr(self, 'batch_flags', 0))
        self.pairpar = YnbPairpar(self.par_d.data_ptr(), self.cross_d.data_ptr(), self.stab_d.data_ptr(),
                                  self.rec_d.data_ptr(), self._ctr(0), self.pair_lanes, 0)
        self.nlist = YnbNlist(self.nbr_j_d.data_ptr(), self.nbr_img_d.data_ptr(), self.cnt_d.data_ptr(),
                              self.status_d.data_ptr(), None, self.kcap, 0)
        self.ewald = YnbEwald(self.kvec_d.data_ptr(), self.krow_d.data_ptr(), self.krow_k_d.data_ptr(),
                              self.skblk_d.data_ptr(), self.fblk_d.data_ptr(), self.chunk_d.data_ptr(),
                              self.work_d.data_ptr(), self.kblk_sys_d.data_ptr(), self.kpos_d.data_ptr(),
                              self.krow_f_d.data_ptr(), self._ctr(3*self.nsys), self.sp['work'].high,
                              self.sp['k'].high, self.apl, self.fslots, self.groups, self.fapl, 0, 0,
                              self.ew_frac_d.data_ptr(), self.ew_step_d.data_ptr(), self.ew_spart_d.data_ptr(),
                              self.ew_coef_d.data_ptr(), self.ew_frec_d.data_ptr(), self.ew_kpart_d.data_ptr(),
                              self.ew_gpart_d.data_ptr())

    # ---- Engine hooks ---------------------------------------------------------
    def grow(self, flags, longest):
        """Enlarges the capacities named by the flags of a failed build."""
        if flags & 2:
            raise RuntimeError('neighbour list: an image translation exceeds 511 cells')
        if flags & 8:
            raise RuntimeError('neighbour list: partner-bin tables smaller than a system needs (batch flags)')
        if flags & 4:
            cap = _round_up(int(self._fullest_bin*1.25) + 16, 16)
            if cap > 1024:
                raise RuntimeError('neighbour list: a bin holds more than 1024 atoms')
            self.bin_cap = cap
            self._alloc_nlist_work()
        if flags & 1:
            self._alloc_nlist(int(longest*1.1) + 8)
        self.pos_ref_d.fill_(1e150)
        self._commit_structs()

    def set_tuning(self, **kw):
        raise RuntimeError('the launch parameters of a SlotEngine are fixed at construction')

    def upload_positions(self, pos_list=None):
        for i in range(self.nsys):
            if self.members[i] is not None:
                p = self.systems[i].pos if pos_list is None else pos_list[i]
                self._up(self.pos_d[self.atom0[i]:self.atom0[i] + len(p)], p)

    def results(self, i=None):
        """(energy, gpos, vtens) of slot i (or a dict over the members); synchronises."""
        tot = self.total_d.cpu().numpy()
        slots = [i] if i is not None else [k for k in range(self.nsys) if self.members[k] is not None]
        out = {}
        for s in slots:
            a0, n = self.atom0[s], self.systems[s].natom
            out[s] = (tot[s, 0], self.gpos_d[a0:a0 + n].cpu().numpy(), tot[s, 1:].reshape(3, 3).copy())
        return out[i] if i is not None else out
