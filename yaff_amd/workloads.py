"""Synthetic workloads: the LJ fluid and rocksalt systems the engine is
benchmarked and tested on, built from seeds (numpy default_rng).

* ``lj_fluid``   -- FCC argon-like fluid: nx*ny*nz cells x 4 atoms, lattice
  constant 34.7/6 A, uniform jitter +-0.1 A, PairPotLJ sigma=3.405 A,
  epsilon=3.79e-4 Ha, rcut 12 A, Switch3(4 A), mass 39.948 amu.
* ``rocksalt``   -- NaCl: nx*ny*nz conventional cells x 8 atoms, a=5.64 A,
  jitter +-0.05 A, charges +-1, ExpRep (A=0.5 Ha, B=1.9/A), LJ (sigma=3 A,
  epsilon=1.5e-4 Ha), EI real space alpha=0.28/A, all rcut 11 A, no
  truncation, Ewald reciprocal with gcut 0.32/A, masses Na 22.990 / Cl 35.453.
* Both accept cell angles (alpha, beta, gamma in degrees) for triclinic boxes;
  positions are fractional lattice sites times the cell plus jitter.

The 864-atom fluid (6x6x6), the 6912-atom rocksalt (12x12x6), the 55296-atom
triclinic rocksalt (24x24x12), 64-system batches with 200-6000 atoms and
Poisson request streams with 100-50000 atoms are built from these.
"""
import math

import numpy as np

from .units import angstrom, amu
from .system import System

__all__ = ['cell_from_lengths_angles', 'lj_fluid', 'rocksalt', 'ForceFieldSpec', 'lj_spec',
           'rocksalt_spec', 'batch_of_systems', 'natom_to_reps']


def cell_from_lengths_angles(a, b, c, alpha=90.0, beta=90.0, gamma=90.0):
    """Row-major cell vectors: a along x, b in the xy plane."""
    al, be, ga = (math.radians(x) for x in (alpha, beta, gamma))
    va = np.array([a, 0.0, 0.0])
    vb = np.array([b*math.cos(ga), b*math.sin(ga), 0.0])
    cx = c*math.cos(be)
    cy = c*(math.cos(al) - math.cos(be)*math.cos(ga))/math.sin(ga)
    cz = math.sqrt(max(c*c - cx*cx - cy*cy, 0.0))
    vc = np.array([cx, cy, cz])
    return np.array([va, vb, vc])


_FCC = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
_NA = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
_CL = _NA + np.array([0.5, 0.0, 0.0])


def _lattice(basis, reps, a, angles, jitter, rng):
    nx, ny, nz = reps
    rvecs = cell_from_lengths_angles(nx*a, ny*a, nz*a, *angles)
    cells = np.array([(i, j, k) for i in range(nx) for j in range(ny) for k in range(nz)], float)
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)/np.array([nx, ny, nz], float)
    pos = frac.dot(rvecs)
    pos += rng.uniform(-jitter, jitter, pos.shape)
    return pos, rvecs


class ForceFieldSpec(object):
    """Parameters of the nonbonded force field of a synthetic system."""

    def __init__(self, kind, **kw):
        self.kind = kind
        self.__dict__.update(kw)


def lj_spec():
    return ForceFieldSpec('lj', sigma=3.405*angstrom, epsilon=3.79e-4, rcut=12.0*angstrom,
                          width=4.0*angstrom, mass=39.948*amu)


def rocksalt_spec():
    return ForceFieldSpec('ionic', exprep_a=0.5, exprep_b=1.9/angstrom, lj_sigma=3.0*angstrom,
                          lj_epsilon=1.5e-4, alpha=0.28/angstrom, rcut=11.0*angstrom,
                          gcut=0.32/angstrom, mass_na=22.990*amu, mass_cl=35.453*amu)


def lj_fluid(reps=(6, 6, 6), seed=0, angles=(90.0, 90.0, 90.0), a=34.7/6):
    rng = np.random.default_rng(seed)
    pos, rvecs = _lattice(_FCC, reps, a*angstrom, angles, 0.1*angstrom, rng)
    n = len(pos)
    spec = lj_spec()
    s = System(np.full(n, 18), pos, ffatype_ids=np.zeros(n, np.int64), rvecs=rvecs, bonds=np.zeros((0, 2), int),
               masses=np.full(n, spec.mass))
    s.ffatypes = np.array(['Ar'])
    s.spec = spec
    return s


def rocksalt(reps=(12, 12, 6), seed=0, angles=(90.0, 90.0, 90.0)):
    rng = np.random.default_rng(seed)
    basis = np.concatenate([_NA, _CL])
    pos, rvecs = _lattice(basis, reps, 5.64*angstrom, angles, 0.05*angstrom, rng)
    n = len(pos)
    ids = np.tile(np.array([0]*4 + [1]*4), n//8).astype(np.int64)
    spec = rocksalt_spec()
    charges = np.where(ids == 0, 1.0, -1.0)
    masses = np.where(ids == 0, spec.mass_na, spec.mass_cl)
    s = System(np.where(ids == 0, 11, 17), pos, ffatype_ids=ids, rvecs=rvecs, bonds=np.zeros((0, 2), int),
               charges=charges, masses=masses)
    s.ffatypes = np.array(['Na', 'Cl'])
    s.spec = spec
    return s


def natom_to_reps(natom, per_cell):
    """Nearest whole n1 x n2 x n3 supercell (n1 >= n2 >= n3 >= 1) to natom."""
    ncell = max(1, int(round(natom/float(per_cell))))
    best = None
    n = max(1, int(round(ncell**(1.0/3.0))))
    for n1 in range(max(1, n - 2), n + 3):
        for n2 in range(max(1, n - 2), n1 + 1):
            n3 = max(1, int(round(ncell/float(n1*n2))))
            if n3 > n2:
                continue
            err = abs(n1*n2*n3 - ncell) + 0.01*(n1 - n3)
            if best is None or err < best[0]:
                best = (err, (n1, n2, n3))
    return best[1]


def batch_of_systems(nsys=64, natom_range=(200, 6000), angle_range=(80.0, 100.0), seed=0,
                     log_uniform=False):
    """Even index: LJ fluid; odd index: rocksalt; triclinic angles drawn from
    ``angle_range``; sizes uniform (or log-uniform) in ``natom_range``."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nsys):
        if log_uniform:
            target = int(round(math.exp(rng.uniform(math.log(natom_range[0]), math.log(natom_range[1])))))
        else:
            target = int(rng.integers(natom_range[0], natom_range[1] + 1))
        angles = tuple(rng.uniform(angle_range[0], angle_range[1], 3))
        sub = int(rng.integers(0, 2**31 - 1))
        if i % 2 == 0:
            out.append(lj_fluid(natom_to_reps(target, 4), seed=sub, angles=angles))
        else:
            out.append(rocksalt(natom_to_reps(target, 8), seed=sub, angles=angles))
    return out
