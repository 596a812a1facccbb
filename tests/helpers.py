"""Shared builders for the parity tests: the same system and force field fed
to the oracle (oracle/) and to the HIP engine (yaff_amd)."""
import os

import numpy as np

from oracle import oracle as orc
from yaff_amd.units import angstrom
from yaff_amd import workloads
from yaff_amd.system import System

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')

KIND = {'lj': orc.LJ, 'mm3': orc.MM3, 'exprep': orc.EXPREP, 'ei': orc.EI}
TRUNC = {None: orc.TR_NONE, 'switch3': orc.TR_SWITCH3, 'hammer': orc.TR_HAMMER}


class Case(object):
    """A system plus pair potentials (dicts) plus optional Ewald settings."""

    def __init__(self, name, system, pots, ewald=None, skin=0.0, stab=None):
        self.name, self.system, self.pots, self.ewald, self.skin = name, system, pots, ewald, skin
        self.stab = np.zeros(0, orc.SROW_DTYPE) if stab is None else stab

    @property
    def rc(self):
        return max(p['rcut'] for p in self.pots) + self.skin


def oracle_pot(p):
    return orc.Pot(KIND[p['kind']], p['rcut'], TRUNC[p.get('trunc')], p.get('trunc_par', 0.0),
                   p0=p.get('p0'), p1=p.get('p1'), onlypauli=p.get('onlypauli'),
                   ffatype_ids=p.get('ffatype_ids'), amp_cross=p.get('amp_cross'), b_cross=p.get('b_cross'),
                   alpha=p.get('alpha', 0.0), dielectric=p.get('dielectric', 1.0))


def engine_system(case):
    from yaff_amd.engine import PotSpec, EwaldSpec, NBSystem
    from yaff_amd import _lib
    kinds = {'lj': _lib.POT_LJ, 'mm3': _lib.POT_MM3, 'exprep': _lib.POT_EXPREP, 'ei': _lib.POT_EI}
    trs = {None: _lib.TR_NONE, 'switch3': _lib.TR_SWITCH3, 'hammer': _lib.TR_HAMMER}
    specs = []
    for p in case.pots:
        if p['kind'] == 'exprep':
            p0, p1, p2 = p['ffatype_ids'].astype(float), None, None
        elif p['kind'] == 'mm3':
            p0, p1, p2 = p['p0'], p['p1'], p['onlypauli'].astype(float)
        elif p['kind'] == 'ei':
            p0, p1, p2 = p['p0'], p.get('p1', np.zeros(case.system.natom)), None
        else:
            p0, p1, p2 = p['p0'], p['p1'], None
        specs.append(PotSpec(kinds[p['kind']], p['rcut'], trs[p.get('trunc')], p.get('trunc_par', 0.0), p0, p1, p2,
                             alpha=p.get('alpha', 0.0), dielectric=p.get('dielectric', 1.0),
                             amp_cross=p.get('amp_cross'), b_cross=p.get('b_cross')))
    ew = None
    if case.ewald is not None:
        ew = EwaldSpec(case.ewald['alpha'], case.ewald['gcut'], case.ewald.get('dielectric', 1.0))
    return NBSystem(case.system.pos, case.system.cell, specs, rc=case.rc, stab=case.stab, ewald=ew,
                    charges=case.system.charges, masses=case.system.masses)


def oracle_results(O, case):
    """rows, [(E, gpos, vtens) per pot], ewald (E, gpos, vtens) or None."""
    s = case.system
    cell = O.cell(s.cell.rvecs if s.cell.nvec else None)
    rmax = orc.rmax_of(cell, case.rc) if cell.nvec else np.zeros(0, np.int64)
    rows = O.nlist_build(s.pos, cell, case.rc, rmax)
    pots = [O.pair_compute(rows, case.stab, oracle_pot(p), s.natom) for p in case.pots]
    ew = None
    if case.ewald is not None:
        gmax = orc.gmax_of(cell, case.ewald['gcut'])
        ew = O.ewald_reci(s.pos, s.charges, cell, case.ewald['alpha'], gmax, case.ewald['gcut'],
                          case.ewald.get('dielectric', 1.0))
    return rows, pots, ew


# ---------------------------------------------------------------------------
def lj_case(reps=(3, 3, 3), seed=0, angles=(90, 90, 90), name='lj'):
    s = workloads.lj_fluid(reps, seed=seed, angles=angles)
    sp = s.spec
    pots = [dict(kind='lj', rcut=sp.rcut, trunc='switch3', trunc_par=sp.width, p0=np.full(s.natom, sp.sigma),
                 p1=np.full(s.natom, sp.epsilon))]
    return Case(name, s, pots)


def ionic_case(reps=(2, 2, 2), seed=0, angles=(90, 90, 90), name='rocksalt', gcut=None):
    s = workloads.rocksalt(reps, seed=seed, angles=angles)
    sp = s.spec
    n = s.natom
    amp = np.full((2, 2), sp.exprep_a)
    bb = np.full((2, 2), sp.exprep_b)
    pots = [dict(kind='exprep', rcut=sp.rcut, ffatype_ids=s.ffatype_ids, amp_cross=amp, b_cross=bb),
            dict(kind='lj', rcut=sp.rcut, p0=np.full(n, sp.lj_sigma), p1=np.full(n, sp.lj_epsilon)),
            dict(kind='ei', rcut=sp.rcut, p0=s.charges, alpha=sp.alpha)]
    return Case(name, s, pots, ewald=dict(alpha=sp.alpha, gcut=sp.gcut if gcut is None else gcut))


def load_fixture_system(name):
    d = np.load(os.path.join(GOLDEN, 'systems.npz'))
    pos = d[name + '_pos']
    rvecs = d[name + '_rvecs'] if (name + '_rvecs') in d.files else None
    charges = d[name + '_charges'] if (name + '_charges') in d.files else None
    bonds = d[name + '_bonds'] if (name + '_bonds') in d.files else None
    numbers = d[name + '_numbers']
    return System(numbers, pos, rvecs=rvecs, charges=charges, bonds=bonds,
                  ffatype_ids=d[name + '_ffatype_ids'] if (name + '_ffatype_ids') in d.files else None)


def rel_err(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max()/scale
