"""Per-system device context behind the drop-in classes.

Every ForcePartPair, NeighborList and Ewald part of one System registers with
the same context; the context owns a one-system Engine (rebuilt when the set
of parts or the cell changes), uploads positions, and evaluates all pair
potentials of the system in one fused pass, handing each part its own slot.
"""
import numpy as np
import torch

from . import _lib
from ._lib import MAX_POT, POT_LJ, POT_MM3, POT_EXPREP, POT_EI, TR_NONE, TR_SWITCH3, TR_HAMMER, SCALING_DTYPE
from .engine import Engine, NBSystem, PotSpec, EwaldSpec, CorrTables, corr_plan

__all__ = ['SystemDevice', 'set_default_precision']

_DEFAULT_PRECISION = ['fp64']


def set_default_precision(precision):
    """'fp64' (bitwise pair sets, 1e-10 forces) or 'mixed' (FP32 arithmetic,
    FP64 accumulation) for contexts created afterwards."""
    if precision not in ('fp64', 'mixed'):
        raise ValueError(precision)
    _DEFAULT_PRECISION[0] = precision


def _trunc_code(tr):
    if tr is None:
        return TR_NONE, 0.0
    from .pair_pot import Switch3, Hammer
    if isinstance(tr, Switch3):
        return TR_SWITCH3, tr.width
    if isinstance(tr, Hammer):
        return TR_HAMMER, tr.tau
    raise TypeError('unsupported truncation %r' % tr)


def pot_spec(pair_pot, stab):
    """PotSpec of a drop-in PairPot object."""
    from . import pair_pot as pp
    code, par = _trunc_code(pair_pot.get_truncation())
    if isinstance(pair_pot, pp.PairPotLJ):
        return PotSpec(POT_LJ, pair_pot.rcut, code, par, pair_pot.sigmas, pair_pot.epsilons, stab=stab)
    if isinstance(pair_pot, pp.PairPotMM3):
        return PotSpec(POT_MM3, pair_pot.rcut, code, par, pair_pot.sigmas, pair_pot.epsilons,
                       pair_pot.onlypaulis.astype(float), stab=stab)
    if isinstance(pair_pot, pp.PairPotExpRep):
        return PotSpec(POT_EXPREP, pair_pot.rcut, code, par, pair_pot.ffatype_ids.astype(float),
                       amp_cross=pair_pot.amp_cross, b_cross=pair_pot.b_cross, stab=stab)
    if isinstance(pair_pot, pp.PairPotEI):
        return PotSpec(POT_EI, pair_pot.rcut, code, par, pair_pot.charges, pair_pot.radii, alpha=pair_pot.alpha,
                       dielectric=pair_pot.dielectric, stab=stab)
    raise TypeError('unsupported pair potential %r' % pair_pot)


class SystemDevice(object):
    @staticmethod
    def of(system):
        dev = getattr(system, '_ynb_device', None)
        if dev is None:
            dev = SystemDevice(system)
            system._ynb_device = dev
        return dev

    def __init__(self, system):
        self.system = system
        self.pairs = []           # (pair_pot, stab) in slot order
        self.nlist = None
        self.ewald = None         # ForcePartEwaldReciprocal
        self.precision = _DEFAULT_PRECISION[0]
        self.engine = None
        self.dirty = True
        self.token = 0            # bumped when positions or the list change
        self._pair_cache = (-1, None)
        self._ewald_cache = (-1, None)
        self.list_token = -1

    # -- registration ------------------------------------------------------
    def add_pair(self, nlist, pair_pot, stab):
        if self.nlist is not None and nlist is not self.nlist:
            raise ValueError('all pair parts of a system must share one NeighborList')
        if len(self.pairs) >= MAX_POT:
            raise ValueError('at most %d ForcePartPair objects per system' % MAX_POT)
        self.nlist = nlist
        self.pairs.append((pair_pot, stab))
        self.dirty = True
        return len(self.pairs) - 1

    def set_ewald(self, part):
        self.ewald = part
        self.dirty = True

    def invalidate(self):
        self.dirty = True

    # -- engine --------------------------------------------------------------
    def get_engine(self):
        if self.engine is None or self.dirty:
            s = self.system
            pots = [pot_spec(p, st) for p, st in self.pairs]
            if self.nlist is not None:
                rc = self.nlist.rcut + self.nlist.skin
                rmax = self.nlist.rmax if self.nlist.rmax is not None else None
                nlow, nhigh = self.nlist.nlow, self.nlist.nhigh
            else:
                rc, rmax, nlow, nhigh = 1.0, None, 0, -1
            ew = None
            if self.ewald is not None:
                e = self.ewald
                ew = EwaldSpec(e.alpha, e.gcut, e.dielectric, e.nlow, e.nhigh, gmax=e.gmax)
            nb = NBSystem(s.pos, s.cell, pots, rc=max(rc, 1e-3), rmax=rmax, nlow=nlow, nhigh=nhigh, ewald=ew,
                          charges=s.charges, masses=s.masses)
            self.engine = Engine([nb], precision=self.precision)
            self.dirty = False
            self.list_token = -1
        return self.engine

    def upload_positions(self):
        eng = self.get_engine()
        eng.set_system_positions(0, self.system.pos)
        self.token += 1

    def rebuild_list(self):
        eng = self.get_engine()
        eng.ensure_nlist()
        self.token += 1
        self.list_token = self.token

    # -- evaluation ------------------------------------------------------------
    def pair_results(self):
        """Per slot: (energy, gpos (natom,3), vtens (3,3)) of all pair parts,
        from one fused launch per position state."""
        tok, res = self._pair_cache
        if tok == self.token and res is not None:
            return res
        eng = self.get_engine()
        if self.list_token < 0:
            raise RuntimeError('the neighbour list was not built: call NeighborList.update() first')
        eng.pair(slots=True, gpos=False)
        n = self.system.natom
        out = eng.pair_out_d[0].cpu().numpy()
        gs = eng.gpos_slots_d[:, :n].cpu().numpy()
        res = [(float(out[p, 0]), gs[p].copy(), out[p, 1:].reshape(3, 3).copy()) for p in range(len(self.pairs))]
        self._pair_cache = (self.token, res)
        return res

    def ewald_results(self):
        tok, res = self._ewald_cache
        if tok == self.token and res is not None:
            return res
        eng = self.get_engine()
        eng.ewald_reci(gpos=True, accumulate=False)
        n = self.system.natom
        out = eng.ewald_out_d[0].cpu().numpy()
        g = eng.gpos_d[:n].cpu().numpy().copy()
        res = (float(out[0]), g, out[1:].reshape(3, 3).copy())
        self._ewald_cache = (self.token, res)
        return res

    def ewald_corr(self, alpha, stab, dielectric, nlow, nhigh):
        """(energy, gpos, vtens) of a ForcePartEwaldCorrection with these
        settings, by ynb_ewald_corr on the system's engine.  The tables and
        buffers are made once per set of settings (and engine)."""
        eng = self.get_engine()
        s = self.system
        stab = np.zeros(0, SCALING_DTYPE) if stab is None else np.asarray(stab, dtype=SCALING_DTYPE)
        key = (id(eng), float(alpha), float(dielectric), int(nlow), int(nhigh), stab.tobytes())
        cache = self.__dict__.setdefault('_corr_cache', {})
        if key not in cache:
            cache.clear()
            plan = corr_plan([dict(atom0=0, natom=s.natom, alpha=alpha, dielectric=dielectric, nlow=nlow,
                                   nhigh=nhigh, stab=stab, charges=s.charges)], eng.npad)
            cache[key] = (CorrTables(plan, eng.device),
                          torch.zeros(10, dtype=torch.float64, device=eng.device),
                          torch.zeros((eng.npad, 3), dtype=torch.float64, device=eng.device))
        tabs, out, gpos = cache[key]
        _lib.check(eng.lib.ynb_ewald_corr(eng.batch, tabs.c, eng.pos_d.data_ptr(), eng.charges_d.data_ptr(),
                                          gpos.data_ptr(), 0, out.data_ptr(), 0, Engine.stream()))
        o = out.cpu().numpy()
        return float(o[0]), gpos[:s.natom].cpu().numpy(), o[1:].reshape(3, 3).copy()
