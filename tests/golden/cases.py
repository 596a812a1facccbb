"""Definitions of the golden parity cases (inputs only, all seeded).

Each case is a system with pair potentials (and optionally Ewald settings and a
scaling table) whose reference outputs -- computed with molmod/yaff's own C
kernels compiled into oracle/_ref -- are stored in ref_outputs.npz by
make_golden.py.  The systems are the coordinate fixtures of the reference's
test suite (yaff/test/common.py), kept as data in systems.npz, plus synthetic
LJ and rocksalt boxes.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ANGSTROM = 1.889726133921252


def fixture_arrays(name):
    d = np.load(os.path.join(HERE, 'systems.npz'))
    out = {}
    for k in d.files:
        if k.startswith(name + '_'):
            out[k[len(name) + 1:]] = d[k]
    return out


def scaling_rows(natom, bonds, s1, s2, s3, s4=1.0):
    from oracle.oracle import scaling_table
    return scaling_table(natom, bonds, s1, s2, s3, s4)


def oracle_pot(p):
    """The oracle's description of a pair-potential dict of a case."""
    from oracle import oracle as orc
    kinds = {'lj': orc.LJ, 'mm3': orc.MM3, 'exprep': orc.EXPREP, 'ei': orc.EI}
    trs = {None: orc.TR_NONE, 'switch3': orc.TR_SWITCH3, 'hammer': orc.TR_HAMMER}
    return orc.Pot(kinds[p['kind']], p['rcut'], trs[p.get('trunc')], p.get('trunc_par', 0.0),
                   p0=p.get('p0'), p1=p.get('p1'), onlypauli=p.get('onlypauli'),
                   ffatype_ids=p.get('ffatype_ids'), amp_cross=p.get('amp_cross'), b_cross=p.get('b_cross'),
                   alpha=p.get('alpha', 0.0), dielectric=p.get('dielectric', 1.0))


def _params(natom, seed):
    rng = np.random.default_rng(seed)
    sig = rng.uniform(2.5, 3.5, natom)*ANGSTROM
    eps = rng.uniform(0.1, 0.5, natom)*1e-3
    pauli = rng.integers(0, 2, natom).astype(np.int32)
    return sig, eps, pauli


def build_cases():
    """name -> dict(pos, rvecs (or None), charges, pots (list of dicts), rc, ewald, stab)."""
    cases = {}
    # -- water32: periodic cubic, bonded (scalings 0, 0, 0.5) -------------
    w = fixture_arrays('water32')
    n = len(w['pos'])
    sig, eps, pauli = _params(n, 1)
    stab = scaling_rows(n, w['bonds'], 0.0, 0.0, 0.5)
    ids = np.array([0, 1, 1]*32, np.int64)
    amp = np.array([[4.1, 2.2], [2.2, 1.3]])
    bb = np.array([[4.5, 3.9], [3.9, 3.1]])/ANGSTROM
    cases['water32'] = dict(
        pos=w['pos'], rvecs=w['rvecs'], charges=w['charges'], stab=stab, rc=9.0*ANGSTROM,
        pots=[dict(kind='lj', rcut=9.0*ANGSTROM, trunc='switch3', trunc_par=4.0*ANGSTROM, p0=sig, p1=eps),
              dict(kind='mm3', rcut=8.0*ANGSTROM, trunc='hammer', trunc_par=1.0*ANGSTROM, p0=sig*0.5, p1=eps,
                   onlypauli=pauli),
              dict(kind='exprep', rcut=7.5*ANGSTROM, ffatype_ids=ids, amp_cross=amp, b_cross=bb),
              dict(kind='ei', rcut=9.0*ANGSTROM, p0=w['charges'], alpha=0.2, dielectric=1.44)],
        ewald=dict(alpha=0.2, gcut=0.4, dielectric=1.44))
    # -- water32 with the long real-space cutoff of the alpha=0.05 table ---
    cases['water32_long'] = dict(
        pos=w['pos'], rvecs=w['rvecs'], charges=w['charges'], stab=stab, rc=5.5/0.05,
        pots=[dict(kind='ei', rcut=5.5/0.05, p0=w['charges'], alpha=0.05)],
        ewald=dict(alpha=0.05, gcut=0.1))
    # -- quartz 2x2x2: skewed cell ----------------------------------------
    q = fixture_arrays('quartz222')
    n = len(q['pos'])
    sig, eps, pauli = _params(n, 2)
    cases['quartz222'] = dict(
        pos=q['pos'], rvecs=q['rvecs'], charges=q['charges'], stab=scaling_rows(n, q['bonds'], 0.2, 0.4, 0.7),
        rc=5.5/0.2,
        pots=[dict(kind='ei', rcut=5.5/0.2, p0=q['charges'], alpha=0.2),
              dict(kind='lj', rcut=6.0*ANGSTROM, trunc='switch3', trunc_par=2.0*ANGSTROM, p0=sig, p1=eps)],
        ewald=dict(alpha=0.2, gcut=0.4))
    # -- lower-dimensional and isolated systems -----------------------------
    for name, seed in (('graphene8', 3), ('polyethylene4', 4), ('caffeine', 5), ('water', 6)):
        f = fixture_arrays(name)
        n = len(f['pos'])
        sig, eps, pauli = _params(n, seed)
        rng = np.random.default_rng(seed + 100)
        charges = f['charges'] if 'charges' in f else rng.uniform(-0.5, 0.5, n)
        radii = rng.uniform(0.5, 1.5, n)*ANGSTROM
        cases[name] = dict(
            pos=f['pos'], rvecs=f.get('rvecs'), charges=charges, stab=None, rc=7.0*ANGSTROM,
            pots=[dict(kind='lj', rcut=7.0*ANGSTROM, trunc='switch3', trunc_par=2.0*ANGSTROM, p0=sig, p1=eps),
                  dict(kind='ei', rcut=7.0*ANGSTROM, p0=charges, alpha=0.0),
                  dict(kind='ei', rcut=6.0*ANGSTROM, p0=charges, p1=radii, alpha=0.0),
                  dict(kind='ei', rcut=6.5*ANGSTROM, p0=charges, p1=radii, alpha=0.3)],
            ewald=None)
    # -- synthetic: triclinic LJ fluid (rmax = 1) and rocksalt with Ewald ----
    from yaff_amd import workloads
    s = workloads.lj_fluid((3, 3, 3), seed=11, angles=(84.0, 97.0, 92.0))
    sp = s.spec
    cases['lj108_tri'] = dict(
        pos=s.pos, rvecs=s.cell.rvecs, charges=None, stab=None, rc=sp.rcut,
        pots=[dict(kind='lj', rcut=sp.rcut, trunc='switch3', trunc_par=sp.width, p0=np.full(s.natom, sp.sigma),
                   p1=np.full(s.natom, sp.epsilon))], ewald=None)
    s = workloads.rocksalt((3, 3, 2), seed=12, angles=(88.0, 93.0, 86.0))
    sp = s.spec
    n = s.natom
    cases['rocksalt144_tri'] = dict(
        pos=s.pos, rvecs=s.cell.rvecs, charges=s.charges, stab=None, rc=sp.rcut,
        pots=[dict(kind='exprep', rcut=sp.rcut, ffatype_ids=s.ffatype_ids, amp_cross=np.full((2, 2), sp.exprep_a),
                   b_cross=np.full((2, 2), sp.exprep_b)),
              dict(kind='lj', rcut=sp.rcut, p0=np.full(n, sp.lj_sigma), p1=np.full(n, sp.lj_epsilon)),
              dict(kind='ei', rcut=sp.rcut, p0=s.charges, alpha=sp.alpha)],
        ewald=dict(alpha=sp.alpha, gcut=sp.gcut))
    return cases
