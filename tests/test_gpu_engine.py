"""GPU parity of the HIP engine against the reference outputs.

Neighbour rows must be bit-identical to the rows of molmod/yaff's C kernels
(SHA-256 of the 72-byte records plus count); energies, gradients and virials
of every pair potential and of the Ewald reciprocal and correction terms
must agree to a relative error of 1e-10 (FP64 mode).
"""
import hashlib
import os
import sys

import numpy as np
import pytest

from conftest import gpu_available

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]

TOL = 1e-10
REF = np.load(os.path.join(HERE, 'golden', 'ref_outputs.npz'))


def _cases():
    import cases
    return cases.build_cases()


CASES = _cases()


def _engine(c, precision='fp64', correction=True):
    from yaff_amd.cell import Cell
    from yaff_amd.engine import NBSystem, PotSpec, EwaldSpec
    from yaff_amd import _lib
    kinds = {'lj': _lib.POT_LJ, 'mm3': _lib.POT_MM3, 'exprep': _lib.POT_EXPREP, 'ei': _lib.POT_EI}
    trs = {None: _lib.TR_NONE, 'switch3': _lib.TR_SWITCH3, 'hammer': _lib.TR_HAMMER}
    n = len(c['pos'])
    specs = []
    for p in c['pots']:
        if p['kind'] == 'exprep':
            p0, p1, p2 = p['ffatype_ids'].astype(float), None, None
        elif p['kind'] == 'mm3':
            p0, p1, p2 = p['p0'], p['p1'], p['onlypauli'].astype(float)
        else:
            p0, p1, p2 = p['p0'], p.get('p1'), None
        specs.append(PotSpec(kinds[p['kind']], p['rcut'], trs[p.get('trunc')], p.get('trunc_par', 0.0), p0, p1, p2,
                             alpha=p.get('alpha', 0.0), dielectric=p.get('dielectric', 1.0),
                             amp_cross=p.get('amp_cross'), b_cross=p.get('b_cross')))
    ew = None
    if c['ewald'] is not None:
        ew = EwaldSpec(c['ewald']['alpha'], c['ewald']['gcut'], c['ewald'].get('dielectric', 1.0), correction=correction)
    cell = Cell(c['rvecs'])
    from yaff_amd._lib import SCALING_DTYPE
    stab = None if c['stab'] is None else np.asarray(c['stab']).astype(SCALING_DTYPE)
    s = NBSystem(c['pos'], cell, specs, rc=c['rc'], stab=stab, ewald=ew, charges=c['charges'])
    from yaff_amd.engine import Engine
    return Engine([s], precision=precision)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max()/max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize('name', sorted(CASES))
def test_rows_bit_exact(name):
    print('case', name, flush=True)
    eng = _engine(CASES[name])
    eng.ensure_nlist()
    rows = eng.rows(0)
    assert len(rows) == int(REF[name + '/nrows'])
    if name + '/rows' in REF.files:
        ref = REF[name + '/rows']
        for f in ref.dtype.names:
            np.testing.assert_array_equal(rows[f], ref[f], err_msg=f)
    assert hashlib.sha256(rows.tobytes()).hexdigest() == str(REF[name + '/rows_sha256'])


@pytest.mark.parametrize('name', sorted(CASES))
def test_pair_fp64(name):
    print('case', name, flush=True)
    c = CASES[name]
    eng = _engine(c)
    eng.ensure_nlist()
    eng.pair(slots=True)
    out = eng.pair_out_d.cpu().numpy()[0]
    gs = eng.gpos_slots_d.cpu().numpy()
    n = len(c['pos'])
    for ip in range(len(c['pots'])):
        e = REF['%s/pot%d/e' % (name, ip)]
        g = REF['%s/pot%d/gpos' % (name, ip)]
        v = REF['%s/pot%d/vtens' % (name, ip)]
        assert abs(out[ip, 0] - e) <= TOL*max(abs(e), 1e-300), (ip, out[ip, 0], e)
        assert rel(gs[ip, :n], g) <= TOL, ip
        assert rel(out[ip, 1:].reshape(3, 3), v) <= TOL, ip


@pytest.mark.parametrize('name', sorted(k for k in CASES if CASES[k]['ewald'] is not None))
def test_ewald_reci_fp64(name):
    print('case', name, flush=True)
    c = CASES[name]
    eng = _engine(c, correction=False)   # the reciprocal term alone (no self term in its out)
    eng.gpos_d.zero_()
    eng.ewald_reci(gpos=True, accumulate=False)
    out = eng.ewald_out_d.cpu().numpy()[0]
    g = eng.gpos_d.cpu().numpy()[:len(c['pos'])]
    e = float(REF[name + '/reci/e'])
    assert abs(out[0] - e) <= TOL*abs(e), (out[0], e)
    assert rel(g, REF[name + '/reci/gpos']) <= TOL
    assert rel(out[1:].reshape(3, 3), REF[name + '/reci/vtens']) <= TOL


# Mixed precision: FP64 arithmetic with FP32 exp/erfc/sincos on FP64-reduced
# arguments.  Tolerances: RMS force error < 1e-5 of the RMS force (the
# north-star bound), energy and virial within 1e-6 relative.
MIXED_FORCE_RMS_TOL = 1e-5
MIXED_REL_TOL = 1e-6


@pytest.mark.parametrize('name', ['rocksalt144_tri', 'water32', 'lj108_tri', 'quartz222'])
def test_mixed_precision_error(name):
    print('case', name, flush=True)
    c = CASES[name]
    n = len(c['pos'])
    eng = _engine(c, precision='mixed')
    eng.ensure_nlist()
    eng.compute(build=False)
    e, g, v = eng.results(0)
    e_ref = sum(float(REF['%s/pot%d/e' % (name, ip)]) for ip in range(len(c['pots'])))
    g_ref = sum(REF['%s/pot%d/gpos' % (name, ip)] for ip in range(len(c['pots'])))
    v_ref = sum(REF['%s/pot%d/vtens' % (name, ip)] for ip in range(len(c['pots'])))
    if c['ewald'] is not None:
        for part in ('reci', 'corr'):
            e_ref += float(REF[name + '/%s/e' % part])
            g_ref = g_ref + REF[name + '/%s/gpos' % part]
            v_ref = v_ref + REF[name + '/%s/vtens' % part]
    rms = np.sqrt(((g - g_ref)**2).sum(axis=1).mean())/np.sqrt((g_ref**2).sum(axis=1).mean())
    print('mixed', name, 'energy', abs(e - e_ref)/abs(e_ref), 'force rms', rms, flush=True)
    assert abs(e - e_ref) <= MIXED_REL_TOL*abs(e_ref)
    assert rms <= MIXED_FORCE_RMS_TOL
    assert np.abs(v - v_ref).max() <= MIXED_REL_TOL*np.abs(v_ref).max()


@pytest.mark.parametrize('name', sorted(CASES))
def test_total_fp64(name):
    """compute(): all potentials summed per pair (the MD and batch path) plus
    the Ewald reciprocal and correction terms (the batched correction
    kernels where the case has scaled pairs), against the sum of the
    reference parts."""
    print('case', name, flush=True)
    c = CASES[name]
    n = len(c['pos'])
    eng = _engine(c)
    eng.ensure_nlist()
    eng.compute(build=False)
    e, g, v = eng.results(0)
    e_ref = sum(float(REF['%s/pot%d/e' % (name, ip)]) for ip in range(len(c['pots'])))
    g_ref = sum(REF['%s/pot%d/gpos' % (name, ip)] for ip in range(len(c['pots'])))
    v_ref = sum(REF['%s/pot%d/vtens' % (name, ip)] for ip in range(len(c['pots'])))
    if c['ewald'] is not None:
        for part in ('reci', 'corr'):
            e_ref += float(REF[name + '/%s/e' % part])
            g_ref = g_ref + REF[name + '/%s/gpos' % part]
            v_ref = v_ref + REF[name + '/%s/vtens' % part]
    scale_e = sum(abs(float(REF['%s/pot%d/e' % (name, ip)])) for ip in range(len(c['pots'])))
    if c['ewald'] is not None:
        scale_e += abs(float(REF[name + '/reci/e'])) + abs(float(REF[name + '/corr/e']))
    assert abs(e - e_ref) <= TOL*scale_e, (e, e_ref)
    assert rel(g, g_ref) <= TOL
    assert rel(v, v_ref) <= TOL
