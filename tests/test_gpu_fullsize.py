"""Parity at the benchmark's full sizes.

* The 864-atom LJ fluid (FCC 6x6x6 cells, L = 34.7 A, sigma 3.405 A,
  rcut 12 A, Switch3 of 4 A) and the 6912-atom rocksalt (12x12x6 cells,
  ExpRep + LJ + EI, rcut 11 A, Ewald gcut 0.32/A): neighbour rows
  bit-identical to the oracle's, energy, gradient and virial to 1e-10
  (FP64); the mixed mode within its contract against FP64 (RMS force
  1e-5 relative, energy 1e-6).
* The 55,296-atom triclinic rocksalt (24x24x12 cells, angles drawn in
  85-95 degrees): the oracle would take minutes, so properties that hold at
  any size.  The gradients sum to zero (pair and Ewald terms conserve
  momentum); a rigid translation of every atom leaves energy, gradient and
  virial unchanged up to rounding (the same pairs, and the Ewald phases
  shift by one factor per k that cancels in |S(k)|^2); the mixed mode stays
  within its contract against FP64.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]

TOL = 1e-10


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max()/np.abs(b).max()


def _gpu(s, precision):
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import Engine
    eng = Engine([nbsystem(s)], precision=precision)
    eng.ensure_nlist()
    eng.compute(build=False)
    return eng


def _oracle(s):
    """Rows and the summed (energy, gradient, virial) of the pair parts and
    the Ewald reciprocal and correction terms, from the oracle."""
    from oracle import oracle as orc
    O = orc.Oracle()
    sp = s.spec
    n = s.natom
    cell = O.cell(s.cell.rvecs)
    rows = O.nlist_build(s.pos, cell, sp.rcut, orc.rmax_of(cell, sp.rcut))
    if sp.kind == 'lj':
        pots = [orc.Pot(orc.LJ, sp.rcut, orc.TR_SWITCH3, sp.width, p0=np.full(n, sp.sigma), p1=np.full(n, sp.epsilon))]
    else:
        pots = [orc.Pot(orc.EXPREP, sp.rcut, ffatype_ids=s.ffatype_ids, amp_cross=np.full((2, 2), sp.exprep_a),
                        b_cross=np.full((2, 2), sp.exprep_b)),
                orc.Pot(orc.LJ, sp.rcut, p0=np.full(n, sp.lj_sigma), p1=np.full(n, sp.lj_epsilon)),
                orc.Pot(orc.EI, sp.rcut, p0=s.charges, alpha=sp.alpha)]
    parts = [O.pair_compute(rows, None, p, n) for p in pots]
    if sp.kind != 'lj':
        parts.append(O.ewald_reci(s.pos, s.charges, cell, sp.alpha, orc.gmax_of(cell, sp.gcut), sp.gcut))
        parts.append(O.ewald_corr(s.pos, s.charges, cell, sp.alpha, None))  # the self term
    return rows, (sum(p[0] for p in parts), sum(p[1] for p in parts), sum(p[2] for p in parts))


@pytest.mark.parametrize('name', ['lj864', 'rocksalt6912'])
def test_benchmark_system_matches_oracle(name):
    from yaff_amd import workloads
    print('case %s' % name, flush=True)
    s = workloads.lj_fluid() if name == 'lj864' else workloads.rocksalt((12, 12, 6), seed=0)
    assert s.natom == {'lj864': 864, 'rocksalt6912': 6912}[name]
    rows, (e_ref, g_ref, v_ref) = _oracle(s)
    eng = _gpu(s, 'fp64')
    assert eng.nlist_status()[0] == 0
    got = eng.rows(0)
    assert len(got) == len(rows) and got.tobytes() == rows.tobytes()
    e, g, v = eng.results(0)
    assert abs(e - e_ref) <= TOL*abs(e_ref), (e, e_ref)
    assert _rel(g, g_ref) <= TOL and _rel(v, v_ref) <= TOL
    em, gm, vm = _gpu(s, 'mixed').results(0)
    rms = np.sqrt(((gm - g)**2).sum(axis=1).mean()/(g**2).sum(axis=1).mean())
    print(name, 'mixed: rel dE %.2e, rms rel dG %.2e' % (abs(em - e)/abs(e), rms), flush=True)
    assert abs(em - e) <= 1e-6*abs(e) and rms <= 1e-5


def test_rocksalt55296_triclinic_invariants():
    from yaff_amd import workloads
    print('case rocksalt55296 triclinic', flush=True)
    rng = np.random.default_rng(2)
    angles = tuple(float(a) for a in rng.uniform(85.0, 95.0, 3))
    s = workloads.rocksalt((24, 24, 12), seed=2, angles=angles)
    assert s.natom == 55296
    eng = _gpu(s, 'fp64')
    assert eng.nlist_status()[0] == 0
    e, g, v = eng.results(0)
    # momentum: the gradients cancel to rounding
    assert np.abs(g.sum(axis=0)).max() <= 1e-12*np.abs(g).sum()
    # rigid translation by a vector that is no lattice vector
    shift = np.array([0.37, -0.21, 0.53])*1.889726133921252
    eng.upload_positions([s.pos + shift])
    eng.ensure_nlist()
    eng.compute(build=False)
    e2, g2, v2 = eng.results(0)
    print('translation: rel dE %.2e, rel dG %.2e, rel dV %.2e' % (abs(e2 - e)/abs(e), _rel(g2, g), _rel(v2, v)),
          flush=True)
    assert abs(e2 - e) <= TOL*abs(e) and _rel(g2, g) <= 1e-9 and _rel(v2, v) <= TOL
    # mixed precision against FP64
    em, gm, vm = _gpu(s, 'mixed').results(0)
    rms = np.sqrt(((gm - g)**2).sum(axis=1).mean()/(g**2).sum(axis=1).mean())
    print('mixed: rel dE %.2e, rms rel dG %.2e, rel dV %.2e' % (abs(em - e)/abs(e), rms, _rel(vm, v)), flush=True)
    assert abs(em - e) <= 1e-6*abs(e) and rms <= 1e-5 and _rel(vm, v) <= 1e-6


def test_rocksalt55296_matches_reference_fixture():
    """The precision workload against molmod/yaff's own C kernels (fixture
    tests/golden/rocksalt55296_ref.npz, written by
    make_rocksalt55296_golden.py): FP64 rows bit-identical (count and
    SHA-256), energy and virial to 1e-10, the gradient of 8192 seeded atoms
    to 1e-10 and 8 random projections of the whole gradient to 1e-10 of
    their bound; mixed precision within its contract (energy and virial
    1e-6, RMS force error over the sample 1e-5)."""
    import hashlib
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, 'golden', 'rocksalt55296_ref.npz')
    ref = np.load(path)
    sys.path.insert(0, os.path.join(here, 'golden'))
    from make_rocksalt55296_golden import workload, sample_and_weights
    s, angles = workload()
    assert s.natom == int(ref['natom']) == 55296 and np.allclose(angles, ref['angles'])
    print('case rocksalt55296 vs reference fixture', flush=True)
    idx, w = sample_and_weights(s.natom)
    assert np.array_equal(idx, ref['sample'])
    eng = _gpu(s, 'fp64')
    assert eng.nlist_status()[0] == 0
    rows = eng.rows(0)
    assert len(rows) == int(ref['nrows'])
    assert hashlib.sha256(rows.tobytes()).hexdigest() == str(ref['rows_sha256'])
    del rows
    e, g, v = eng.results(0)
    gnorm = np.sqrt(float(ref['gpos_sumsq']))
    print('fp64: rel dE %.2e, rel dV %.2e, sample rel dG %.2e' % (
        abs(e - ref['e'])/abs(ref['e']), _rel(v, ref['vtens']), _rel(g[idx], ref['gpos_sample'])), flush=True)
    assert abs(e - ref['e']) <= TOL*abs(ref['e'])
    assert _rel(v, ref['vtens']) <= TOL
    assert _rel(g[idx], ref['gpos_sample']) <= TOL
    proj = np.einsum('kij,ij->k', w, g)
    bound = np.sqrt((w*w).sum(axis=(1, 2)))*gnorm
    assert (np.abs(proj - ref['gpos_proj']) <= TOL*bound).all()
    del eng
    em, gm, vm = _gpu(s, 'mixed').results(0)
    gs = ref['gpos_sample']
    rms = np.sqrt(((gm[idx] - gs)**2).sum(axis=1).mean()/(gs**2).sum(axis=1).mean())
    print('mixed vs reference: rel dE %.2e, rms rel dG (sample) %.2e, rel dV %.2e' % (
        abs(em - ref['e'])/abs(ref['e']), rms, _rel(vm, ref['vtens'])), flush=True)
    assert abs(em - ref['e']) <= 1e-6*abs(ref['e']) and rms <= 1e-5 and _rel(vm, ref['vtens']) <= 1e-6
