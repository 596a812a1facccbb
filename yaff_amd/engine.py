"""Batched device engine: lays out systems in HBM and drives libyaff_nb.so.

A batch is any number of independent systems.  Their atoms are concatenated in
padded ranges (each system starts at a multiple of 256 atoms, so neither a
wavefront nor an atom block of the Ewald force pass spans two systems); every
launch covers the whole batch.  A single
system -- the drop-in ForceField path -- is a batch of one.

Device memory is held in torch tensors (plumbing only); every kernel is a
hand-written HIP kernel of libyaff_nb.so launched on torch's current stream,
so a sequence of engine calls can be captured in a HIP graph.
"""
import math

import numpy as np
import torch

from . import _lib
from ._lib import (YnbSys, YnbBatch, YnbNlist, YnbPairpar, YnbEwald, MAX_POT, FP64, MIXED, TR_NONE, POT_LJ,
                   ROW_DTYPE, SCALING_DTYPE, check)

__all__ = ['PotSpec', 'EwaldSpec', 'NBSystem', 'Engine', 'ewald_kvectors']

TWO_PI = 6.2831853071795864769
FOUR_PI = 12.5663706143591729538
ALIGN = 256          # systems start at multiples of 256 atoms (k_ewald_force takes blocks of up to 256 atoms)
KBLOCK = 256         # k vectors per reduction block


class PotSpec(object):
    """One pair potential of a system.  p0/p1/p2 are per-atom parameters:
    LJ/MM3 sigma, epsilon (MM3: p2 = onlypauli; LJ: p2 is filled with
    sqrt(epsilon) by the engine); EI charge, radius; ExpRep
    ffatype id with amp_cross/b_cross tables.  ``stab`` is this potential's
    scaling table (default: the system's)."""

    def __init__(self, kind, rcut, trunc=TR_NONE, trunc_par=0.0, p0=None, p1=None, p2=None, alpha=0.0,
                 dielectric=1.0, amp_cross=None, b_cross=None, stab=None):
        self.kind, self.rcut, self.trunc, self.trunc_par = int(kind), float(rcut), int(trunc), float(trunc_par)
        self.stab = None if stab is None else np.asarray(stab, dtype=SCALING_DTYPE)
        self.p0, self.p1, self.p2 = p0, p1, p2
        self.alpha, self.dielectric = float(alpha), float(dielectric)
        self.amp_cross, self.b_cross = amp_cross, b_cross


class EwaldSpec(object):
    """Reciprocal-space settings; gmax defaults to ff.py:386 for the cell.
    nlow/nhigh are the Ewald parts' own atom groups (ff.py:331-532), not the
    neighbour list's.  With ``correction`` the engine's evaluations and MD
    steps also add ForcePartEwaldCorrection (ewald.c:278-349: the self term
    and the erf term of the system's scaled pairs) with the same alpha,
    dielectric and groups."""

    def __init__(self, alpha, gcut, dielectric=1.0, nlow=0, nhigh=-1, gmax=None, correction=False):
        self.alpha, self.gcut, self.dielectric = float(alpha), float(gcut), float(dielectric)
        self.nlow, self.nhigh = int(nlow), int(nhigh)
        self.gmax = None if gmax is None else np.asarray(gmax, dtype=np.int64)
        self.correction = bool(correction)


class NBSystem(object):
    """Input of the engine for one system.  ``rc`` is the neighbour-list radius
    (default: largest pair cutoff plus ``skin``); rmax defaults to
    nlist.py:132.  With skin > 0 the engine's MD steps rebuild the list of the
    system only when an atom has moved by skin/2 (ynb_nlist_skin_check)."""

    def __init__(self, pos, cell, pots, rc=None, rmax=None, nlow=0, nhigh=-1, stab=None, ewald=None,
                 charges=None, masses=None, skin=0.0):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.cell = cell
        self.pots = list(pots)
        if len(self.pots) > MAX_POT:
            raise ValueError('at most %d pair potentials per system' % MAX_POT)
        if skin < 0:
            raise ValueError('The skin parameter must be positive.')
        self.skin = float(skin)
        self.rc = float(rc if rc is not None else max([p.rcut for p in self.pots] + [0.0]) + self.skin)
        if rmax is None:
            rmax = np.ceil(self.rc/cell.rspacings - 0.5).astype(np.int64) if cell.nvec > 0 else np.zeros(0, int)
        self.rmax = np.asarray(rmax, dtype=np.int64)
        self.nlow = int(nlow)
        self.nhigh = len(self.pos) if nhigh == -1 else int(nhigh)
        self.stab = np.zeros(0, SCALING_DTYPE) if stab is None else np.asarray(stab, dtype=SCALING_DTYPE)
        self.ewald = ewald
        self.charges = None if charges is None else np.asarray(charges, dtype=np.float64)
        self.masses = None if masses is None else np.asarray(masses, dtype=np.float64)

    @property
    def natom(self):
        return len(self.pos)


def ewald_kplan(cell, alpha, gcut, gmax=None):
    """The half-space k vectors compute_ewald_reci visits (ewald.c:38-58),
    from the library's host planner (ynb_ewald_kvectors: the same FP64
    expressions for k and |k|^2, so the cutoff test picks the same set), in
    row order (g2, g1 major, g0 ascending) with c_k = 4 pi/V
    exp(-k^2/4 alpha^2)/k^2: (gmax, g [nk][3], kv [nk][4], krow [nrow][4] =
    first g0, g1, g2, count, krow_k [nrow] = first k of each row)."""
    if gmax is None:
        gmax = np.ceil(gcut/cell.gspacings - 0.5).astype(np.int64)
    gmax = np.ascontiguousarray(gmax, dtype=np.int64)
    cap = int((2*gmax[0] + 1)*(2*gmax[1] + 1)*(gmax[2] + 1))
    kv = np.empty((cap, 4))
    g = np.empty((cap, 3), np.int32)
    krow = np.empty((cap, 4), np.int32)
    krow_k = np.empty(cap, np.int32)
    nk, nrow = _lib.C.c_int64(0), _lib.C.c_int64(0)
    gv = np.ascontiguousarray(cell.gvecs_full, dtype=np.float64)
    L = _lib.load()
    check(L.ynb_ewald_kvectors(gv.ctypes.data, float(cell.volume), float(alpha), float(gcut), gmax.ctypes.data,
                               cap, kv.ctypes.data, g.ctypes.data, krow.ctypes.data, krow_k.ctypes.data,
                               _lib.C.byref(nk), _lib.C.byref(nrow)))
    nk, nrow = nk.value, nrow.value
    return gmax, g[:nk].astype(np.int64), kv[:nk].copy(), krow[:nrow].copy(), krow_k[:nrow].copy()


def ewald_kvectors(cell, alpha, gcut, gmax=None):
    """(gmax, g [nk][3], kv [nk][4]) of ewald_kplan."""
    return ewald_kplan(cell, alpha, gcut, gmax)[:3]


def ewald_kvectors_cached(s):
    """ewald_kplan of an NBSystem, cached on it."""
    ew = s.ewald
    cache = s.__dict__.setdefault('_plan_cache', {})
    kkey = ('kv', s.cell.rvecs_full.tobytes(), ew.alpha, ew.gcut, None if ew.gmax is None else tuple(ew.gmax))
    if kkey not in cache:
        cache.clear()
        cache[kkey] = ewald_kplan(s.cell, ew.alpha, ew.gcut, ew.gmax)
    return cache[kkey]


SQRT_PI = 1.77245385090551602729


def corr_self_energy(alpha, charges, nlow, nhigh):
    """-(alpha/sqrt(pi)) sum_{nlow <= i < nhigh} q_i^2, summed in atom order
    with the operations of ewald.c:286-290 (a cumulative sum is sequential)."""
    q = np.asarray(charges, dtype=np.float64)[nlow:nhigh]
    terms = -((alpha/SQRT_PI*q)*q)
    return float(np.cumsum(terms)[-1]) if len(terms) else 0.0


def corr_plan(entries, npad):
    """Host tables of ynb_ewald_corr (include/yaff_nb.h: ynb_corr) for a
    batch.  ``entries``: one per system, None (no correction: no rows, no
    self term) or a dict with atom0, natom, alpha, dielectric, nlow, nhigh
    (-1: natom), stab (scaling rows, local indices) and charges.  The self
    term is summed in atom order as ewald.c:286-290 does; each atom lists its
    rows in row order (2 r: it is b, 2 r + 1: it is a)."""
    csys = np.zeros(len(entries), dtype=np.dtype([('alpha', 'f8'), ('dielectric', 'f8'), ('self_energy', 'f8'),
                                                  ('nlow', 'i4'), ('nhigh', 'i4'), ('row0', 'i4'), ('nrow', 'i4'),
                                                  ('blk0', 'i4'), ('nblk', 'i4')]))
    stabs, blk_sys, atoms, refs = [], [], [], []
    row0 = blk0 = 0
    for i, e in enumerate(entries):
        c = csys[i]
        c['dielectric'] = 1.0
        nrow = 0
        if e is not None:
            st = np.asarray(e['stab'] if e['stab'] is not None else np.zeros(0, SCALING_DTYPE), dtype=SCALING_DTYPE)
            nlow = int(e['nlow'])
            nhigh = e['natom'] if e['nhigh'] == -1 else int(e['nhigh'])
            c['alpha'], c['dielectric'] = e['alpha'], e['dielectric']
            c['self_energy'] = corr_self_energy(e['alpha'], e['charges'], nlow, nhigh)
            c['nlow'], c['nhigh'] = nlow, nhigh
            nrow = len(st)
            if nrow:
                stabs.append(st)
                r = row0 + np.arange(nrow, dtype=np.int64)
                atoms.append(np.stack([e['atom0'] + st['b'], e['atom0'] + st['a']], axis=1).ravel())
                refs.append(np.stack([2*r, 2*r + 1], axis=1).ravel())
        nblk = max(1, _ceil_div(nrow, 256))
        c['row0'], c['nrow'], c['blk0'], c['nblk'] = row0, nrow, blk0, nblk
        blk_sys.append(np.full(nblk, i, np.int32))
        row0 += nrow
        blk0 += nblk
    atoms = np.concatenate(atoms) if atoms else np.zeros(0, np.int64)
    refs = np.concatenate(refs) if refs else np.zeros(0, np.int64)
    order = np.argsort(atoms, kind='stable')
    ref0 = np.zeros(npad + 1, np.int32)
    ref0[1:] = np.cumsum(np.bincount(atoms, minlength=npad)[:npad])
    return dict(csys=csys, stab=np.concatenate(stabs) if stabs else np.zeros(1, SCALING_DTYPE),
                blk_sys=np.concatenate(blk_sys) if blk_sys else np.zeros(1, np.int32),
                atom_ref0=ref0, atom_ref=refs[order].astype(np.int32) if len(refs) else np.zeros(1, np.int32),
                nrow_tot=row0, nblk_tot=blk0)


class CorrTables(object):
    """Device copy of a corr_plan plus its work buffers, and the YnbCorr
    structure that points at them."""

    def __init__(self, plan, device):
        def t(a):
            return torch.from_numpy(np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint8).copy()).to(device)
        self.plan = plan
        self.csys_d = t(plan['csys'])
        self.stab_d = t(plan['stab'])
        self.blk_sys_d = t(plan['blk_sys'])
        self.ref0_d = t(plan['atom_ref0'])
        self.ref_d = t(plan['atom_ref'])
        self.rowg_d = torch.zeros(3*max(plan['nrow_tot'], 1), dtype=torch.float64, device=device)
        self.partial_d = torch.zeros(8*plan['nblk_tot'], dtype=torch.float64, device=device)
        self.ctr_d = torch.zeros(len(plan['csys']), dtype=torch.int32, device=device)
        self.c = _lib.YnbCorr(self.csys_d.data_ptr(), self.stab_d.data_ptr(), self.blk_sys_d.data_ptr(),
                              self.ref0_d.data_ptr(), self.ref_d.data_ptr(), self.rowg_d.data_ptr(),
                              self.partial_d.data_ptr(), self.ctr_d.data_ptr(), plan['nrow_tot'],
                              plan['nblk_tot'])


def work_table(plan, sys, blk0, chunk0):
    """The structure-factor work items of a system's plan (every row block
    x every chunk, block major) with batch offsets: [nwork][4] int32
    (ynb_ewald_work)."""
    work = np.empty((max(plan['nwork'], 1), 4), np.int32)
    check(_lib.load().ynb_ewald_work(len(plan['skblk']), len(plan['chunks']), int(sys), int(blk0), int(chunk0),
                                     work.ctypes.data))
    return work[:plan['nwork']]


def _row_blocks(krow, kmax):
    """Consecutive rows grouped into blocks of at most kmax k vectors (a
    longer row is a block of its own): [n][2] local row ranges
    (ynb_row_blocks)."""
    krow = np.ascontiguousarray(krow, dtype=np.int32)
    blk = np.empty((max(len(krow), 1), 2), np.int32)
    n = _lib.C.c_int64(0)
    check(_lib.load().ynb_row_blocks(krow.ctypes.data, len(krow), int(kmax), blk.ctypes.data, _lib.C.byref(n)))
    return blk[:n.value].copy()


def _ewald_local_plan(s, apl, skb, fkb):
    """The Ewald tables of one system with system-local offsets, cached on
    the NBSystem (keyed by the cell, the Ewald settings and the launch
    parameters) so that batches can be re-formed cheaply: k vectors, rows,
    row blocks of the structure-factor pass (<= skb k each) and of the force
    pass (<= fkb k each), atom chunks of 64 * apl atoms and the
    structure-factor work items (every row block x every chunk)."""
    ew = s.ewald
    gmax, g, kv, krow, krow_k = ewald_kvectors_cached(s)
    cache = s._plan_cache
    kkey = next(k for k in cache if k[0] == 'kv')
    pkey = ('plan', kkey, apl, skb, fkb)
    if pkey in cache:
        return cache[pkey]
    nk = len(g)
    nkp = max(_round_up(nk, KBLOCK), KBLOCK)
    kpad = np.zeros((nkp, 4))
    kpad[:nk] = kv
    skblk = _row_blocks(krow, skb)
    fblk = _row_blocks(krow, fkb)
    # force records: rows padded to multiples of 4 records
    krow_f = np.empty(max(len(krow), 1), np.int32)
    kpos = np.empty(nkp, np.int32)
    nrec = _lib.C.c_int64(0)
    krow = np.ascontiguousarray(krow, dtype=np.int32)
    check(_lib.load().ynb_ewald_records(krow.ctypes.data, len(krow), nkp, krow_f.ctypes.data, kpos.ctypes.data,
                                        _lib.C.byref(nrec)))
    krow_f = krow_f[:len(krow)]
    # atom chunks: below nlow (group 1), [nlow, nhigh) (group 0), above nhigh (group 2)
    nhigh = s.natom if ew.nhigh == -1 else ew.nhigh
    nlow = ew.nlow
    csize = 64*apl
    parts = []
    for (lo, hi, grp) in ((0, nlow, 1), (nlow, nhigh, 0), (nhigh, s.natom, 2)):
        a = np.arange(lo, hi, csize)
        parts.append(np.stack([np.zeros_like(a), a, np.minimum(csize, hi - a), np.full_like(a, grp)], axis=1))
    chunks = np.concatenate(parts).astype(np.int32)
    plan = dict(gmax=gmax, nk=nk, nkp=nkp, kpad=kpad, krow=krow, krow_k=krow_k.astype(np.int32), nrow=len(krow),
                skblk=skblk, fblk=fblk, chunks=chunks, nwork=len(skblk)*len(chunks), krow_f=krow_f, kpos=kpos,
                nrec=int(nrec.value), grouped=bool(nlow > 0 or nhigh < s.natom),
                nchunk_low=int((chunks[:, 3] == 1).sum()), nchunk_high=int((chunks[:, 3] == 2).sum()))
    cache[pkey] = plan
    return plan


class SystemEntry(object):
    """One system's part of the batch tables: its ynb_sys (offsets into the
    batch tables left at their system-local values, see shift_offsets), per-atom
    pair parameters [MAX_POT][3][natom] and records [natom][4], its ExpRep
    cross tables and scaling tables (one per potential, in potential order)."""

    def __init__(self, h, par, rec, cross, stabs):
        self.h, self.par, self.rec, self.cross, self.stabs = h, par, rec, cross, stabs


def system_entry(lib, s, a0):
    """SystemEntry of NBSystem ``s`` placed at padded atom offset ``a0``:
    cell, cutoff, image range and groups; ynb_sys_plan's bin grid; each
    potential's kind, truncation and parameters; the specialised pair form
    and its per-atom record; the Ewald groups and, for a correction without
    scaled pairs, its self term as ew_const."""
    h = YnbSys()
    cell = s.cell
    for k in range(9):
        h.rvecs[k] = float(cell.rvecs_full.ravel()[k])
        h.gvecs[k] = float(cell.gvecs_full.ravel()[k])
    h.rc = s.rc
    h.volume = cell.volume
    h.nvec = cell.nvec
    h.natom = s.natom
    h.atom0 = a0
    h.npot = len(s.pots)
    for k in range(3):
        h.rmax[k] = int(s.rmax[k]) if k < cell.nvec else 0
    h.nlow, h.nhigh = s.nlow, s.nhigh
    h.ew_nlow, h.ew_nhigh = 0, s.natom
    h.ew_const = 0.0
    if s.ewald is not None:
        h.ew_nlow = s.ewald.nlow
        h.ew_nhigh = s.natom if s.ewald.nhigh == -1 else s.ewald.nhigh
        if s.ewald.correction and len(s.stab) == 0:
            # the correction is its self term alone: a constant of the reciprocal out
            h.ew_const = corr_self_energy(s.ewald.alpha, s.charges, h.ew_nlow, h.ew_nhigh)/s.ewald.dielectric
    check(lib.ynb_sys_plan(h))
    par = np.zeros((MAX_POT, 3, s.natom), np.float64)
    cross, stabs = [], []
    stab_off = cross_off = 0
    for p, pot in enumerate(s.pots):
        hp = h.pot[p]
        st = s.stab if pot.stab is None else pot.stab
        hp.stab0, hp.nstab = stab_off, len(st)
        stab_off += len(st)
        if len(st):
            stabs.append(np.asarray(st, dtype=SCALING_DTYPE))
            h.any_stab = 1
        hp.kind, hp.trunc, hp.rcut, hp.trunc_par = pot.kind, pot.trunc, pot.rcut, pot.trunc_par
        hp.alpha, hp.dielectric = pot.alpha, pot.dielectric
        for q, arr in enumerate((pot.p0, pot.p1, pot.p2)):
            if arr is not None:
                par[p, q] = np.asarray(arr, dtype=np.float64)
        if pot.kind == POT_LJ and pot.p1 is not None:
            # sqrt(epsilon) per atom: the pair kernel forms eps_ab = se_a se_b
            par[p, 2] = np.sqrt(np.asarray(pot.p1, dtype=np.float64))
        if pot.amp_cross is not None:
            hp.nffatype = pot.amp_cross.shape[0]
            hp.cross0 = cross_off
            c = np.concatenate([np.ascontiguousarray(pot.amp_cross, dtype=float).ravel(),
                                np.ascontiguousarray(pot.b_cross, dtype=float).ravel()])
            cross.append(c)
            cross_off += len(c)
    h.pair_form, rec = _pair_form(s)
    return SystemEntry(h, par, rec, cross, stabs)


def shift_offsets(h, stab_off, cross_off):
    """Moves a SystemEntry's scaling-table and cross-table offsets to where
    its tables start in the batch tables."""
    for p in range(h.npot):
        h.pot[p].stab0 += stab_off
        if h.pot[p].nffatype:
            h.pot[p].cross0 += cross_off


def _pair_form(s):
    """pair_form of a system (a specialised pair kernel for the potential
    sets of the LJ fluid and of the rocksalt-type ionic force field) and the
    per-atom record [natom][4] that kernel reads."""
    n = s.natom
    rec = np.zeros((n, 4))
    pots = s.pots
    if any(len(s.stab if p.stab is None else p.stab) for p in pots) or n == 0:
        return _lib.FORM_GENERIC, rec
    kinds = [p.kind for p in pots]
    if kinds == [_lib.POT_LJ] and pots[0].trunc in (_lib.TR_NONE, _lib.TR_SWITCH3) and pots[0].p1 is not None:
        rec[:, 0] = np.asarray(pots[0].p0, dtype=float)
        rec[:, 1] = np.sqrt(np.asarray(pots[0].p1, dtype=float))
        return _lib.FORM_LJ, rec
    if (kinds == [_lib.POT_EXPREP, _lib.POT_LJ, _lib.POT_EI] and all(p.trunc == TR_NONE for p in pots)
            and pots[2].alpha > 0 and (pots[2].p1 is None or not np.any(np.asarray(pots[2].p1)))
            and pots[1].p1 is not None):
        rec[:, 0] = np.asarray(pots[1].p0, dtype=float)
        rec[:, 1] = np.sqrt(np.asarray(pots[1].p1, dtype=float))
        rec[:, 2] = np.asarray(pots[2].p0, dtype=float)/math.sqrt(pots[2].dielectric)
        rec[:, 3] = np.asarray(pots[0].p0, dtype=float)
        return _lib.FORM_IONIC, rec
    return _lib.FORM_GENERIC, rec


def _round_up(x, m):
    return ((x + m - 1)//m)*m


def _ceil_div(a, b):
    return -(-a//b)


class Engine(object):
    """Device state and launches for a batch of NBSystem objects."""

    overlap = True   # MD steps run the Ewald chain on a second stream (see _md_forces)
    reci_first = False  # with overlap: issue the Ewald chain before the list and pair chain
    _side = None
    _fork = None

    def __init__(self, systems, precision='fp64', apl=4, skb=None, fkb=None, fapl=4, kcap=None, device='cuda',
                 pair_lanes=32, upload=True):
        self.lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError('yaff_amd.Engine needs a HIP device (torch.cuda.is_available() is False)')
        self.device = torch.device(device)
        self.systems = list(systems)
        self.precision = {'fp64': FP64, 'mixed': MIXED}[precision] if isinstance(precision, str) else int(precision)
        self.nsys = len(self.systems)
        self.pair_lanes = int(pair_lanes)
        self.apl, self.skb, self.fkb, self.fapl = int(apl), skb, fkb, int(fapl)
        self._layout()
        self._plan_sys()
        self._alloc(kcap)
        if upload:
            self.upload_positions()
        self.nbuilds = 0

    # ------------------------------------------------------------------
    def _layout(self):
        self.atom0 = []
        off = 0
        for s in self.systems:
            self.atom0.append(off)
            off += _round_up(max(s.natom, 1), ALIGN)
        self.npad = max(off, ALIGN)
        self.wave_sys = np.full(self.npad//64, -1, np.int32)
        for i, s in enumerate(self.systems):
            a0 = self.atom0[i]
            self.wave_sys[a0//64:(a0 + max(s.natom, 1) + 63)//64] = i

    def _plan_sys(self):
        self.hsys = (YnbSys*self.nsys)()
        par = np.zeros((MAX_POT, 3, self.npad), np.float64)
        rec = np.zeros((self.npad, 4))
        cross = []
        stabs = []
        nbin = 0
        stab_off = cross_off = 0
        self.has_ewald = any(s.ewald is not None for s in self.systems)
        # systems whose correction has scaled pairs: ynb_ewald_corr (the others' self term is ew_const)
        self.has_corr = any(s.ewald is not None and s.ewald.correction and len(s.stab) > 0 for s in self.systems)
        # systems whose scaled pairs would lose the Ewald correction in compute() and MD steps
        self.uncorrected = [i for i, s in enumerate(self.systems)
                            if s.ewald is not None and not s.ewald.correction and len(s.stab) > 0]
        self.groups = 0
        for i, s in enumerate(self.systems):
            e = system_entry(self.lib, s, self.atom0[i])
            h = e.h
            shift_offsets(h, stab_off, cross_off)
            h.bin0 = nbin
            nbin += h.nbin
            stabs.extend(e.stabs)
            stab_off += sum(len(st) for st in e.stabs)
            cross.extend(e.cross)
            cross_off += sum(len(c) for c in e.cross)
            a0 = self.atom0[i]
            par[:, :, a0:a0 + s.natom] = e.par
            rec[a0:a0 + s.natom] = e.rec
            self.hsys[i] = h
        self.nbin = nbin
        self.bin_sys = np.zeros(max(nbin, 1), np.int32)
        for i in range(self.nsys):
            self.bin_sys[self.hsys[i].bin0:self.hsys[i].bin0 + self.hsys[i].nbin] = i
        self.par_h = par
        self.rec_h = rec
        self.cross_h = np.concatenate(cross) if cross else np.zeros(1)
        self.stab_h = np.concatenate(stabs) if stabs else np.zeros(1, SCALING_DTYPE)
        if self.has_ewald:
            self._plan_ewald()

    def _plan_ewald(self, apl=None, skb=None, fkb=None, fapl=None):
        """Rows, row blocks, atom chunks and work items of the Ewald passes:
        the cached per-system tables joined with their batch offsets.  Unset
        parameters get defaults that give both passes a few thousand
        wavefronts."""
        ew_sys = [i for i, s in enumerate(self.systems) if s.ewald is not None and s.natom > 0]
        for i in ew_sys:
            if self.systems[i].cell.nvec != 3:
                raise ValueError('Ewald reciprocal needs a 3D periodic cell')
        apl = self.apl if apl is None else int(apl)
        if apl not in (2, 4, 8):
            raise ValueError('apl must be 2, 4 or 8')
        self.apl = apl
        fapl = self.fapl if fapl is None else int(fapl)
        if fapl not in (1, 2, 4):
            raise ValueError('fapl must be 1, 2 or 4')
        self.fapl = fapl
        nks = {i: len(ewald_kvectors_cached(self.systems[i])[1]) for i in ew_sys}
        if skb is None:
            skb = self.skb
        if skb is None:
            # structure-factor wavefronts: sum over systems of blocks x chunks
            skb = 512
            while skb > 16 and sum(max(1, nks[i]//skb)*_ceil_div(self.systems[i].natom, 64*apl)
                                   for i in ew_sys) < 4096:
                skb //= 2
        if fkb is None:
            fkb = self.fkb
        if fkb is None:
            # force wavefronts: sum over systems of atom blocks x row blocks
            fkb = 512
            while fkb > 16 and sum(max(1, nks[i]//fkb)*_ceil_div(self.systems[i].natom, 64*fapl)
                                   for i in ew_sys) < 4096:
                fkb //= 2
        self.skb, self.fkb = int(skb), int(fkb)
        kvec, krow, krow_k, skblk, fblk, chunk, work, kblk_sys, krow_f, kpos = [], [], [], [], [], [], [], [], [], []
        k0 = row0 = sb0 = fb0 = chunk0 = rec0 = 0
        spart0 = 0
        fslots = 0
        self.gmax = [None]*self.nsys
        self.groups = 0
        for i, s in enumerate(self.systems):
            h = self.hsys[i]
            if s.ewald is None or s.natom == 0:
                h.nk = 0
                h.k0 = k0
                h.nfrb = h.nslot = 0
                continue
            ew = s.ewald
            p = _ewald_local_plan(s, self.apl, self.skb, self.fkb)
            self.gmax[i] = p['gmax']
            nkp = p['nkp']
            kvec.append(p['kpad'])
            kblk_sys.append(np.full(nkp//KBLOCK, i, np.int32))
            krow.append(p['krow'])
            krow_k.append(p['krow_k'] + k0)
            skblk.append(p['skblk'] + row0)
            fblk.append(p['fblk'] + row0)
            krow_f.append(p['krow_f'] + rec0)
            kpos.append(np.where(p['kpos'] >= 0, p['kpos'] + rec0, -1).astype(np.int32))
            ch = p['chunks'].copy()
            ch[:, 0] = i
            chunk.append(ch)
            work.append(work_table(p, i, sb0, chunk0))
            if p['grouped']:
                self.groups = 1
            nch = len(ch)
            h.alpha, h.dielectric = ew.alpha, ew.dielectric
            h.k0, h.nk = k0, nkp
            h.frb0, h.nfrb = fb0, len(p['fblk'])
            h.nslot = _ceil_div(h.nfrb, 4)
            fslots = max(fslots, h.nslot)
            h.chunk0, h.nchunk = chunk0, nch
            h.spart0 = spart0
            h.nchunk_low = p['nchunk_low']
            h.nchunk_high = p['nchunk_high']
            spart0 += nch*nkp
            k0 += nkp
            row0 += p['nrow']
            sb0 += len(p['skblk'])
            fb0 += len(p['fblk'])
            chunk0 += nch
            rec0 += p['nrec']

        def cat(parts, empty):
            parts = [x for x in parts if len(x)]
            return np.ascontiguousarray(np.concatenate(parts)) if parts else empty
        self.ewald_h = dict(
            kvec=cat(kvec, np.zeros((KBLOCK, 4))),
            krow=cat(krow, np.zeros((1, 4), np.int32)),
            krow_k=cat(krow_k, np.zeros(1, np.int32)),
            skblk=cat(skblk, np.zeros((1, 2), np.int32)),
            fblk=cat(fblk, np.zeros((1, 2), np.int32)),
            chunk=cat(chunk, np.zeros((1, 4), np.int32)),
            work=cat(work, np.zeros((1, 4), np.int32)),
            kblk_sys=cat(kblk_sys, np.zeros(1, np.int32)),
            kpos=cat(kpos, np.full(KBLOCK, -1, np.int32)),
            krow_f=cat(krow_f, np.zeros(1, np.int32)),
        )
        self.nfrec = rec0 + 4  # 4 spare records: the force pass reads one group ahead
        self.nk_tot = k0
        self.nwork = int(sum(len(w) for w in work))
        self.spart_entries = max(spart0, 1)
        self.fslots = fslots

    # ------------------------------------------------------------------
    def _t(self, arr, dtype=None):
        a = np.ascontiguousarray(arr)
        t = torch.from_numpy(a.copy())
        return t.to(self.device)

    def _alloc(self, kcap):
        L = self.lib
        dev = self.device
        self.sys_d = self._t(np.frombuffer(bytes(self.hsys), np.uint8))
        self.wave_sys_d = self._t(self.wave_sys)
        self.bin_sys_d = self._t(self.bin_sys)
        self.bin_cap = self._max_bin_atoms()
        self.batch = YnbBatch(self.sys_d.data_ptr(), self.wave_sys_d.data_ptr(), self.bin_sys_d.data_ptr(),
                              self.nsys, self.nbin, self.npad, self.bin_cap, self._batch_flags())
        self.pos_d = torch.zeros((self.npad, 3), dtype=torch.float64, device=dev)
        self.gpos_d = torch.zeros((self.npad, 3), dtype=torch.float64, device=dev)
        self.par_d = self._t(self.par_h)
        self.cross_d = self._t(self.cross_h)
        self.stab_d = self._t(np.frombuffer(self.stab_h.tobytes(), np.uint8))
        self.rec_d = self._t(self.rec_h)
        # block counters of the last-block reductions (zero; kernels leave them at zero):
        # [nsys] pair, [nsys] unused, [nsys] kick, [npad/64] Ewald force blocks
        self.ctr_d = torch.zeros(3*self.nsys + self.npad//64 + 1, dtype=torch.int32, device=dev)
        self.pairpar = YnbPairpar(self.par_d.data_ptr(), self.cross_d.data_ptr(), self.stab_d.data_ptr(),
                                  self.rec_d.data_ptr(), self._ctr(0), self.pair_lanes, 0)
        # neighbour list
        self._alloc_nlist_work()
        self.cnt_d = torch.zeros(self.npad, dtype=torch.int32, device=dev)
        self.status_d = torch.zeros(4, dtype=torch.int32, device=dev)
        skins = np.array([s.skin for s in self.systems], dtype=np.float64)
        self.use_skin = bool(np.any(skins > 0))
        self.skin_d = self._t(skins if len(skins) else np.zeros(1))
        self.rebuild_d = torch.ones(max(self.nsys, 1), dtype=torch.int32, device=dev)
        self.pos_ref_d = torch.full((self.npad, 3), 1e150, dtype=torch.float64, device=dev)
        self._alloc_nlist(kcap if kcap is not None else self._estimate_kcap())
        self.partial_d = torch.zeros((self.npad//4)*MAX_POT*8, dtype=torch.float64, device=dev)
        self.pair_out_d = torch.zeros((self.nsys, MAX_POT, 10), dtype=torch.float64, device=dev)
        self.ewald_out_d = torch.zeros((self.nsys, 10), dtype=torch.float64, device=dev)
        self.total_d = torch.zeros((self.nsys, 10), dtype=torch.float64, device=dev)
        self.gpos_slots_d = None
        charges = np.zeros(self.npad)
        for i, s in enumerate(self.systems):
            if s.charges is not None:
                charges[self.atom0[i]:self.atom0[i] + s.natom] = s.charges
        self.charges_d = self._t(charges)
        if self.has_ewald:
            self._alloc_ewald()
        self.corr_tables = None
        if self.has_corr:
            entries = [dict(atom0=self.atom0[i], natom=s.natom, alpha=s.ewald.alpha, dielectric=s.ewald.dielectric,
                            nlow=s.ewald.nlow, nhigh=s.ewald.nhigh, stab=s.stab, charges=s.charges)
                       if s.ewald is not None and s.ewald.correction and len(s.stab) > 0 else None
                       for i, s in enumerate(self.systems)]
            self.corr_tables = CorrTables(corr_plan(entries, self.npad), dev)
        # MD state
        masses = np.ones(self.npad)
        for i, s in enumerate(self.systems):
            if s.masses is not None:
                masses[self.atom0[i]:self.atom0[i] + s.natom] = s.masses
        self.masses_d = self._t(masses)
        self.minv_d = self._t(1.0/masses)
        self.vel_d = torch.zeros((self.npad, 3), dtype=torch.float64, device=dev)
        self.dt_d = torch.zeros(self.nsys, dtype=torch.float64, device=dev)
        self.ekin2_d = torch.zeros(self.nsys, dtype=torch.float64, device=dev)
        self.md_partial_d = torch.zeros(self.npad//64, dtype=torch.float64, device=dev)

    def _batch_flags(self):
        """Build paths the batch needs: all-pairs (systems without bins), the
        exact in-build re-test (binned systems that are not box_safe) and the
        large partner-bin tables (more than BUILD_NB_SMALL partner bins per
        home group)."""
        f = 0
        for i in range(self.nsys):
            h = self.hsys[i]
            if h.natom > 0 and h.nbin == 0:
                f |= _lib.BATCH_ALLPAIRS
            if h.nbin > 0 and not h.box_safe:
                f |= _lib.BATCH_EXACT_BUILD
            if self.lib.ynb_build_search_bins(_lib.C.byref(h)) > _lib.BUILD_NB_SMALL:
                f |= _lib.BATCH_WIDE_SEARCH
            f |= {_lib.FORM_GENERIC: _lib.BATCH_FORM_GENERIC, _lib.FORM_LJ: _lib.BATCH_FORM_LJ,
                  _lib.FORM_IONIC: _lib.BATCH_FORM_IONIC}[h.pair_form]
        return f

    def _max_bin_atoms(self):
        """Bin capacity: twice the largest mean occupancy plus a margin."""
        m = 16
        for i, s in enumerate(self.systems):
            if self.hsys[i].nbin > 0:
                m = max(m, int(2*s.natom/self.hsys[i].nbin) + 16)
        return min(_round_up(m, 16), 1024)

    def _ctr(self, off):
        """Address of counter `off` of ctr_d."""
        return self.ctr_d.data_ptr() + 4*off

    def _alloc_nlist_work(self):
        self.nl_work_bytes = self.lib.ynb_nlist_workspace_bytes(self.npad, self.nbin, self.bin_cap)
        # zeroed once: the build keeps its bin counters at zero between builds
        self.nl_work_d = torch.zeros(self.nl_work_bytes, dtype=torch.uint8, device=self.device)

    def _alloc_ewald(self):
        dev = self.device
        e = self.ewald_h
        self.ew_tabs = {k: self._t(v) for k, v in e.items()}
        self.ew_frac_d = torch.zeros(4*self.npad, dtype=torch.float64, device=dev)
        self.ew_step_d = torch.zeros(4*self.npad, dtype=torch.float64, device=dev)
        self.ew_spart_d = torch.zeros(2*self.spart_entries, dtype=torch.float64, device=dev)
        self.ew_coef_d = torch.zeros(6*max(self.nk_tot, KBLOCK), dtype=torch.float64, device=dev)
        self.ew_frec_d = torch.zeros(2*self.nfrec, dtype=torch.float64, device=dev)
        self.ew_kpart_d = torch.zeros(8*max(self.nk_tot//32, 1), dtype=torch.float64, device=dev)
        self.ew_gpart_d = torch.zeros(max(self.fslots, 1)*self.npad*3, dtype=torch.float64, device=dev)
        t = self.ew_tabs
        self.ewald = YnbEwald(t['kvec'].data_ptr(), t['krow'].data_ptr(), t['krow_k'].data_ptr(),
                              t['skblk'].data_ptr(), t['fblk'].data_ptr(), t['chunk'].data_ptr(),
                              t['work'].data_ptr(), t['kblk_sys'].data_ptr(), t['kpos'].data_ptr(),
                              t['krow_f'].data_ptr(), self._ctr(3*self.nsys), self.nwork,
                              self.nk_tot, self.apl, self.fslots, self.groups, self.fapl, 0, 0,
                              self.ew_frac_d.data_ptr(),
                              self.ew_step_d.data_ptr(), self.ew_spart_d.data_ptr(), self.ew_coef_d.data_ptr(),
                              self.ew_frec_d.data_ptr(), self.ew_kpart_d.data_ptr(), self.ew_gpart_d.data_ptr())

    def tuning(self):
        """The launch parameters an autotuner may change."""
        e = self.has_ewald
        return {'pair_lanes': self.pair_lanes, 'apl': self.apl if e else None, 'skb': self.skb if e else None,
                'fkb': self.fkb if e else None, 'fapl': self.fapl if e else None}

    def set_tuning(self, pair_lanes=None, apl=None, skb=None, fkb=None, fapl=None):
        """Changes launch parameters in place (results are unchanged up to
        summation order).  Ewald tables are re-planned when their parameters
        change."""
        if pair_lanes is not None:
            if pair_lanes not in (16, 32, 64):
                raise ValueError('pair_lanes must be 16, 32 or 64')
            self.pair_lanes = int(pair_lanes)
            self.pairpar.lanes = self.pair_lanes
        if self.has_ewald and (apl is not None or skb is not None or fkb is not None or fapl is not None):
            self._plan_ewald(apl, skb, fkb, fapl)
            self.sys_d = self._t(np.frombuffer(bytes(self.hsys), np.uint8))
            self.batch.sys = self.sys_d.data_ptr()
            self._alloc_ewald()

    def _estimate_kcap(self):
        est = 16
        for s in self.systems:
            if s.natom == 0:
                continue
            if s.cell.nvec == 3:
                rho = s.natom/s.cell.volume
                n = 4.0/3.0*math.pi*s.rc**3*rho
            else:
                nimg = int(np.prod(2*s.rmax + 1)) if len(s.rmax) else 1
                n = s.natom*nimg
            est = max(est, int(1.3*n) + 32)
        return est

    def _alloc_nlist(self, kcap):
        self.kcap = int(kcap)
        self.nbr_j_d = torch.zeros(self.kcap*self.npad, dtype=torch.int32, device=self.device)
        self.nbr_img_d = torch.zeros(self.kcap*self.npad, dtype=torch.int32, device=self.device)
        self.nlist = YnbNlist(self.nbr_j_d.data_ptr(), self.nbr_img_d.data_ptr(), self.cnt_d.data_ptr(),
                              self.status_d.data_ptr(), None, self.kcap, 0)
        if hasattr(self, 'pos_ref_d'):
            self.pos_ref_d.fill_(1e150)   # new capacity: the next skin test rebuilds everything

    # ------------------------------------------------------------------
    @staticmethod
    def stream():
        """The current stream of the current device, as the raw handle the
        library takes (read without building a torch Stream object, which
        costs ~10 us per call: a served MD iteration makes ~20 such calls)."""
        return C_stream(_raw_stream())

    def upload_positions(self, pos_list=None):
        """Copies host positions (list per system, default: the systems' own)."""
        host = np.zeros((self.npad, 3))
        for i, s in enumerate(self.systems):
            p = s.pos if pos_list is None else pos_list[i]
            host[self.atom0[i]:self.atom0[i] + s.natom] = p
        self.pos_d.copy_(torch.from_numpy(host), non_blocking=False)

    def set_system_positions(self, i, pos):
        a0 = self.atom0[i]
        self.pos_d[a0:a0 + len(pos)].copy_(torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float64)))

    def build_nlist(self, force=True):
        """Builds the list of every system (force) or, with a skin, of the
        systems whose atoms moved by skin/2 since their last build (decided
        on the device, so the step can be captured in a graph)."""
        if force or not self.use_skin:
            self.nlist.rebuild = None
            check(self.lib.ynb_nlist_build(self.batch, self.pos_d.data_ptr(), self.nl_work_d.data_ptr(),
                                           self.nl_work_bytes, self.nlist, self.stream()))
            if self.use_skin:
                self.pos_ref_d.copy_(self.pos_d)
        else:
            check(self.lib.ynb_nlist_skin_check(self.batch, self.pos_d.data_ptr(), self.pos_ref_d.data_ptr(),
                                                self.skin_d.data_ptr(), self.rebuild_d.data_ptr(), self.stream()))
            self.nlist.rebuild = self.rebuild_d.data_ptr()
            check(self.lib.ynb_nlist_build(self.batch, self.pos_d.data_ptr(), self.nl_work_d.data_ptr(),
                                           self.nl_work_bytes, self.nlist, self.stream()))
            self.nlist.rebuild = None
        self.nbuilds += 1

    def nlist_status(self):
        """(flags, longest row) of the builds since the last call; synchronises
        and clears the (sticky) flags.  Flags: 1 a row exceeded kcap, 2 an
        image shift exceeded 511 cells, 4 a bin exceeded its capacity, 8 the
        partner-bin tables were too small for a system (the batch flags do
        not match its systems).  The longest row is read only after a flag."""
        flags = int(self.status_d[0].item())
        longest = 0
        if flags & 1:
            longest = int(self.cnt_d.max().item())
        if flags & 4:
            self._fullest_bin = int(self.status_d[1].item())
        if flags:
            self.status_d.zero_()
        return flags, longest

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
            self.batch.max_bin_atoms = cap
            self._alloc_nlist_work()
        if flags & 1:
            self._alloc_nlist(int(longest*1.1) + 8)
        self.pos_ref_d.fill_(1e150)   # lists of the failed build are partial: rebuild every system next

    def ensure_nlist(self):
        """Builds the list, growing capacities until every row and bin fits."""
        for _ in range(4):
            self.build_nlist()
            flags, mx = self.nlist_status()
            if not flags:
                return
            self.grow(flags, mx)
        raise RuntimeError('neighbour list: capacities still exceeded')

    def pair(self, mask=0xF, gpos=True, accumulate=False, slots=False):
        if slots and self.gpos_slots_d is None:
            self.gpos_slots_d = torch.zeros((MAX_POT, self.npad, 3), dtype=torch.float64, device=self.device)
        check(self.lib.ynb_pair_compute(self.batch, self.pairpar, self.pos_d.data_ptr(), self.nlist, mask,
                                        self.precision, self.gpos_d.data_ptr() if gpos else None,
                                        1 if accumulate else 0,
                                        self.gpos_slots_d.data_ptr() if slots else None,
                                        self.partial_d.data_ptr(), self.pair_out_d.data_ptr(), self.stream()))

    def ewald_reci(self, gpos=True, accumulate=True):
        if not self.has_ewald:
            self.ewald_out_d.zero_()
            return
        check(self.lib.ynb_ewald_reci(self.batch, self.ewald, self.pos_d.data_ptr(), self.charges_d.data_ptr(),
                                      self.precision, self.gpos_d.data_ptr() if gpos else None,
                                      1 if accumulate else 0, self.ewald_out_d.data_ptr(), self.stream()))

    def ewald_corr(self, gpos=True, accumulate=True, out=None, out_accumulate=True):
        """ForcePartEwaldCorrection of every system whose EwaldSpec asks for
        it: gradient into gpos_d, energy/virial into ``out`` (default: added
        to ewald_out_d, so the system totals include it)."""
        if self.corr_tables is None:
            return
        out = self.ewald_out_d if out is None else out
        check(self.lib.ynb_ewald_corr(self.batch, self.corr_tables.c, self.pos_d.data_ptr(),
                                      self.charges_d.data_ptr(), self.gpos_d.data_ptr() if gpos else None,
                                      1 if accumulate else 0, out.data_ptr(), 1 if out_accumulate else 0,
                                      self.stream()))

    def _check_corrected(self):
        if self.uncorrected:
            raise ValueError('systems %s have scaled pairs and an Ewald reciprocal part but no Ewald correction: '
                             'their energies and forces would miss ewald.c:278-349 (use '
                             'EwaldSpec(..., correction=True))' % self.uncorrected)

    def total(self):
        check(self.lib.ynb_total(self.batch, self.pair_out_d.data_ptr(),
                                 self.ewald_out_d.data_ptr() if self.has_ewald else None,
                                 self.total_d.data_ptr(), self.stream()))

    def compute(self, build=True, gpos=True):
        """Neighbour list, pair potentials and Ewald reciprocal for every
        system; energy+virial totals in total_d, gradient in gpos_d.
        Asynchronous (call nlist_status() to validate the list)."""
        self._check_corrected()
        if build:
            self.build_nlist(force=False)
        self.pair(gpos=gpos, accumulate=False)
        self.ewald_reci(gpos=gpos, accumulate=True)
        self.ewald_corr(gpos=gpos)
        self.total()

    # ---- MD (velocity Verlet, yaff/sampling/verlet.py:339-367) ----------
    def set_velocities(self, vel_list):
        host = np.zeros((self.npad, 3))
        for i, s in enumerate(self.systems):
            host[self.atom0[i]:self.atom0[i] + s.natom] = vel_list[i]
        self.vel_d.copy_(torch.from_numpy(host))

    def set_timestep(self, dt):
        dt = np.array(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.nsys,)))
        self.dt_d.copy_(torch.from_numpy(dt))

    def md_step(self):
        """kick, drift, neighbour list, forces, kick: one Verlet step of every
        system, on device, without synchronisation."""
        self.md_steps(1)

    def md_steps(self, nsteps):
        """nsteps Verlet steps of every system, on device, without
        synchronisation.  The kick that ends a step and the kick and drift
        that start the next are one kernel (bitwise the same values as
        md_step repeated, dt_d being the same for both halves: it must not
        change inside a sequence -- the server, which changes it between
        iterations, calls md_step); the sequence starts with a kick and drift
        and ends with a kick, so the state between calls is that of
        md_step."""
        self._check_corrected()
        self._md_kick_drift()
        for k in range(nsteps):
            self._md_forces()
            self._md_kick(join=k + 1 < nsteps)

    def _md_prebin(self):
        # the drift also prepares the Ewald pass and bins the atoms for the
        # list build (not with a skin: the skin test decides which systems
        # are rebuilt, after the drift)
        return not self.use_skin and self.batch.nbin > 0

    def _md_kick_drift(self):
        prebin = self._md_prebin()
        ew = self.ewald if self.has_ewald else None
        check(self.lib.ynb_verlet_kick_drift(self.batch, self.pos_d.data_ptr(), self.vel_d.data_ptr(),
                                             self.gpos_d.data_ptr(), self.minv_d.data_ptr(), self.dt_d.data_ptr(),
                                             ew, self.charges_d.data_ptr(),
                                             self.nl_work_d.data_ptr() if prebin else None, self.nl_work_bytes,
                                             self.nlist if prebin else None, self.stream()))

    def _md_forces(self):
        # The Ewald chain (structure factors, k reduction, forces) needs only
        # the positions and the set-up the drift made, not the list or the
        # pair pass: with `overlap` it runs on a second stream (forked after
        # the drift, joined before the correction and the kick; recorded as
        # a fork and a join in a captured graph), so the two chains share the
        # chip.  Their buffers are disjoint; results are the same bits.
        # The list and pair chain is issued first (unless reci_first): a
        # captured graph replays the first-issued branch of a fork on the
        # queue of the node before it, and a dependency within one queue
        # costs less than one across queues, so the longer chain goes there.
        overlap = self.has_ewald and self.overlap
        if overlap:
            main = torch.cuda.current_stream()
            side = self._side_stream()
            if self._fork is None:
                self._fork = torch.cuda.Event()
            self._fork.record(main)
            if self.reci_first:
                self._md_side(side)
        if self._md_prebin():
            self.nlist.flags = _lib.NLIST_PREBINNED
            try:
                self.build_nlist(force=False)
            finally:
                self.nlist.flags = 0
        else:
            self.build_nlist(force=False)
        self.pair(gpos=True, accumulate=False)
        if self.has_ewald:
            if overlap:
                if not self.reci_first:
                    self._md_side(side)
                main.wait_stream(side)
            else:
                self._md_reci()
            # the correction adds its gradient to gpos_d (after the pair pass
            # has written it) and its energy/virial to ewald_out_d (after the
            # reciprocal pass has written it)
            self.ewald_corr()

    def _md_side(self, side):
        # the Ewald chain on the side stream, after the fork event
        side.wait_event(self._fork)
        with torch.cuda.stream(side):
            self._md_reci()

    def _md_reci(self):
        # the Ewald partial gradients stay in their slots and are summed by
        # the kick, which leaves the total gradient in gpos_d
        ew = self.ewald
        ew.flags = _lib.EWALD_PREPARED
        try:
            self.ewald_reci(gpos=False)
        finally:
            ew.flags = 0

    def _side_stream(self):
        """The second stream of the MD step (default priority), made once."""
        if getattr(self, '_side', None) is None:
            self._side = torch.cuda.Stream(device=self.device)
        return self._side

    def _md_kick(self, join=False):
        # the kick sums the pair gradient and the Ewald slots and totals the
        # energy and virial (in place of total()); with join, the next step's
        # kick and drift follow in the same kernel
        common = (self.vel_d.data_ptr(), self.gpos_d.data_ptr(),
                  self.ew_gpart_d.data_ptr() if self.has_ewald else None,
                  self.minv_d.data_ptr(), self.masses_d.data_ptr(), self.dt_d.data_ptr(),
                  self.md_partial_d.data_ptr(), self._ctr(2*self.nsys), self.ekin2_d.data_ptr(),
                  self.pair_out_d.data_ptr(), self.ewald_out_d.data_ptr() if self.has_ewald else None,
                  self.total_d.data_ptr())
        if not join:
            check(self.lib.ynb_verlet_kick(self.batch, *common, self.stream()))
            return
        prebin = self._md_prebin()
        check(self.lib.ynb_verlet_kick_kick_drift(self.batch, self.pos_d.data_ptr(), *common,
                                                  self.ewald if self.has_ewald else None, self.charges_d.data_ptr(),
                                                  self.nl_work_d.data_ptr() if prebin else None, self.nl_work_bytes,
                                                  self.nlist if prebin else None, self.stream()))

    def capture_md(self, nsteps=1):
        """Captures nsteps Verlet steps of every system in a HIP graph (all
        launches are kernels of libyaff_nb.so on the capture stream: no
        memsets, no library sorts, no synchronisation).  Capacities must
        already fit (run ensure_nlist / a few eager steps first); check
        nlist_status() after replays."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            pass
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.md_steps(nsteps)
        return graph

    # ---- host views ---------------------------------------------------
    def results(self, i=None):
        """(energy, gpos, vtens) of system i (or lists for all); synchronises."""
        tot = self.total_d.cpu().numpy()
        g = self.gpos_d.cpu().numpy()
        out = []
        for s in range(self.nsys) if i is None else [i]:
            a0, n = self.atom0[s], self.systems[s].natom
            out.append((tot[s, 0], g[a0:a0 + n].copy(), tot[s, 1:].reshape(3, 3).copy()))
        return out if i is None else out[0]

    def rows(self, i=0):
        """Yaff rows (neigh_dtype) of system i in nlist_build_low's order."""
        L = self.lib
        natom = self.systems[i].natom
        count = _lib.C.c_int64(0)
        wbytes = L.ynb_rows_workspace_bytes(natom, 0)
        work = torch.empty(max(wbytes, 1), dtype=torch.uint8, device=self.device)
        check(L.ynb_nlist_rows(self.batch, i, self.pos_d.data_ptr(), self.nlist, work.data_ptr(), wbytes, None, 0,
                               _lib.C.byref(count), self.stream()))
        n = count.value
        wbytes = L.ynb_rows_workspace_bytes(natom, n)
        work = torch.empty(max(wbytes, 1), dtype=torch.uint8, device=self.device)
        rows = torch.empty(max(n, 1)*ROW_DTYPE.itemsize, dtype=torch.uint8, device=self.device)
        check(L.ynb_nlist_rows(self.batch, i, self.pos_d.data_ptr(), self.nlist, work.data_ptr(), wbytes,
                               rows.data_ptr(), n, _lib.C.byref(count), self.stream()))
        return np.frombuffer(rows.cpu().numpy().tobytes()[:n*ROW_DTYPE.itemsize], ROW_DTYPE).copy()


def C_stream(handle):
    return _lib.C.c_void_p(handle)


if hasattr(torch._C, '_cuda_getCurrentRawStream') and hasattr(torch._C, '_cuda_getDevice'):
    def _raw_stream():
        return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())
else:  # the public, slower route
    def _raw_stream():
        return torch.cuda.current_stream().cuda_stream
