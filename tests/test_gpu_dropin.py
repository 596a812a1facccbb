"""The drop-in yaff.pes interface on the GPU, used the way Yaff's own tests
use it (yaff/pes/test/test_pair_pot.py:73-111, test_ewald.py:121-145), checked
against the oracle: same rows, energies/gradients/virials to 1e-10.

Also: a batch of systems gives each system the result it has alone
(bit-identical with the same launch parameters; with other launch
parameters, which change summation orders, within 1e-11 in FP64 and within
the mixed-precision contract -- energy 1e-6, RMS gradient 1e-5 -- in mixed
precision).
"""
import os
import sys

import numpy as np
import pytest

from conftest import gpu_available

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]

TOL = 1e-10
ANG = 1.889726133921252


def _water32():
    import cases
    import yaff_amd as ya
    f = cases.fixture_arrays('water32')
    return ya.System(f['numbers'], f['pos'], ffatypes=['O', 'H', 'H']*32, bonds=f['bonds'], rvecs=f['rvecs'],
                     charges=f['charges'])


def _oracle_total(system, parts_spec, stab_by_part, alpha=None, gcut=None, diel=1.0, corr_stab=None):
    from oracle import oracle as orc
    O = orc.Oracle()
    cell = O.cell(system.cell.rvecs)
    rc = max(p.rcut for p in parts_spec)
    rows = O.nlist_build(system.pos, cell, rc, orc.rmax_of(cell, rc))
    out = []
    for p, st in zip(parts_spec, stab_by_part):
        out.append(O.pair_compute(rows, st, p, system.natom))
    if alpha is not None:
        out.append(O.ewald_reci(system.pos, system.charges, cell, alpha, orc.gmax_of(cell, gcut), gcut, diel))
        out.append(O.ewald_corr(system.pos, system.charges, cell, alpha, corr_stab, diel))
        e_neut = orc.ewald_neutralizing(system.charges, alpha, system.cell.volume, diel)
        out.append((e_neut, np.zeros((system.natom, 3)), -e_neut*np.eye(3)))
    return rows, out


def test_forcefield_water32_matches_oracle():
    import yaff_amd as ya
    from oracle import oracle as orc
    print('case water32 drop-in', flush=True)
    system = _water32()
    n = system.natom
    rng = np.random.default_rng(3)
    sig = rng.uniform(2.5, 3.5, n)*ANG
    eps = rng.uniform(1e-4, 5e-4, n)
    pauli = rng.integers(0, 2, n).astype(np.int32)
    nlist = ya.NeighborList(system)
    sc_vdw = ya.Scalings(system, 0.0, 0.0, 1.0)
    sc_ei = ya.Scalings(system, 0.0, 0.0, 0.5)
    alpha, gcut = 0.2, 0.4
    lj = ya.PairPotLJ(sig, eps, 9.0*ANG, ya.Switch3(4.0*ANG))
    mm3 = ya.PairPotMM3(sig*0.5, eps, pauli, 8.0*ANG, ya.Hammer(1.0*ANG))
    ex = ya.PairPotExpRep(np.array([0, 1, 1]*32), np.array([[4.1, 2.2], [2.2, 1.3]]),
                          np.array([[4.5, 3.9], [3.9, 3.1]])/ANG, 7.5*ANG)
    ei = ya.PairPotEI(system.charges, alpha, 9.0*ANG)
    parts = [ya.ForcePartPair(system, nlist, sc_vdw, lj), ya.ForcePartPair(system, nlist, sc_vdw, mm3),
             ya.ForcePartPair(system, nlist, sc_vdw, ex), ya.ForcePartPair(system, nlist, sc_ei, ei),
             ya.ForcePartEwaldReciprocal(system, alpha, gcut), ya.ForcePartEwaldCorrection(system, alpha, sc_ei),
             ya.ForcePartEwaldNeutralizing(system, alpha)]
    ff = ya.ForceField(system, parts, nlist)
    gpos = np.zeros((n, 3))
    vtens = np.zeros((3, 3))
    e = ff.compute(gpos, vtens)

    spec = [orc.Pot(orc.LJ, 9.0*ANG, orc.TR_SWITCH3, 4.0*ANG, p0=sig, p1=eps),
            orc.Pot(orc.MM3, 8.0*ANG, orc.TR_HAMMER, 1.0*ANG, p0=sig*0.5, p1=eps, onlypauli=pauli),
            orc.Pot(orc.EXPREP, 7.5*ANG, ffatype_ids=np.array([0, 1, 1]*32), amp_cross=ex.amp_cross,
                    b_cross=ex.b_cross),
            orc.Pot(orc.EI, 9.0*ANG, p0=system.charges, alpha=alpha)]
    rows, ref = _oracle_total(system, spec, [sc_vdw.stab]*3 + [sc_ei.stab], alpha, gcut, corr_stab=sc_ei.stab)
    e_ref = sum(r[0] for r in ref)
    g_ref = sum(r[1] for r in ref)
    v_ref = sum(r[2] for r in ref)
    assert abs(e - e_ref) <= TOL*abs(e_ref), (e, e_ref)
    assert np.abs(gpos - g_ref).max() <= TOL*np.abs(g_ref).max()
    assert np.abs(vtens - v_ref).max() <= TOL*np.abs(v_ref).max()
    # per-part caches (ff.py:147-157) and Yaff's rows
    for part, r in zip(parts[:6], ref):
        assert abs(part.energy - r[0]) <= TOL*max(abs(r[0]), 1e-300)
        assert np.abs(part.gpos - r[1]).max() <= TOL*max(np.abs(r[1]).max(), 1e-300)
    assert nlist.neighs.tobytes() == rows.tobytes()
    # low-level PairPot.compute on Yaff rows
    g2 = np.zeros((n, 3))
    v2 = np.zeros((3, 3))
    e2 = ei.compute(nlist.neighs, sc_ei.stab, g2, v2, nlist.nneigh)
    assert abs(e2 - ref[3][0]) <= TOL*abs(ref[3][0])
    assert np.abs(g2 - ref[3][1]).max() <= TOL*np.abs(ref[3][1]).max()
    assert np.abs(v2 - ref[3][2]).max() <= TOL*np.abs(ref[3][2]).max()
    # every potential and truncation through the row entry (gradients summed
    # per atom in row order: bitwise the same on a second call), also with
    # the virial alone and with neither derivative
    for ip, (pot, st) in enumerate(((lj, sc_vdw.stab), (mm3, sc_vdw.stab), (ex, sc_vdw.stab), (ei, sc_ei.stab))):
        ga, va = np.zeros((n, 3)), np.zeros((3, 3))
        gb, vb = np.zeros((n, 3)), np.zeros((3, 3))
        ea = pot.compute(nlist.neighs, st, ga, va, nlist.nneigh)
        eb = pot.compute(nlist.neighs, st, gb, vb, nlist.nneigh)
        assert ea == eb and np.array_equal(ga, gb) and np.array_equal(va, vb)
        assert abs(ea - ref[ip][0]) <= TOL*abs(ref[ip][0]), ip
        assert np.abs(ga - ref[ip][1]).max() <= TOL*np.abs(ref[ip][1]).max(), ip
        assert np.abs(va - ref[ip][2]).max() <= TOL*np.abs(ref[ip][2]).max(), ip
        vc = np.zeros((3, 3))
        ec = pot.compute(nlist.neighs, st, None, vc, nlist.nneigh)
        assert abs(ec - ea) <= 1e-14*abs(ea) and np.array_equal(vc, va)
        ed = pot.compute(nlist.neighs, st, None, None, nlist.nneigh)
        assert abs(ed - ea) <= TOL*abs(ea)
    # a second evaluation after moving the atoms
    pos = system.pos + rng.normal(0, 0.01, system.pos.shape)
    ff.update_pos(pos)
    gpos[:] = 0
    e3 = ff.compute(gpos)
    rows3, ref3 = _oracle_total(system, spec, [sc_vdw.stab]*3 + [sc_ei.stab], alpha, gcut, corr_stab=sc_ei.stab)
    assert abs(e3 - sum(r[0] for r in ref3)) <= TOL*abs(e3)


def test_skin_recompute_rows():
    """nlist_recompute semantics (test_nlist.py:266-319): below the skin the
    rows are kept and recomputed, above it the list is rebuilt."""
    import yaff_amd as ya
    from oracle import oracle as orc
    print('case water32 skin', flush=True)
    system = _water32()
    nlist = ya.NeighborList(system, skin=2.0*ANG)
    nlist.request_rcut(6.0*ANG)
    nlist.update()
    rows0 = nlist.neighs.copy()
    O = orc.Oracle()
    cell = O.cell(system.cell.rvecs)
    pos_old = system.pos.copy()
    rng = np.random.default_rng(1)
    system.pos[:] = pos_old + rng.uniform(-0.1, 0.1, pos_old.shape)*0.45*ANG/np.sqrt(3)/(2*(nlist.rmax.max() + 1))
    nlist.update()
    ref = O.nlist_recompute(system.pos, pos_old, cell, rows0)
    assert nlist.neighs.tobytes() == ref.tobytes()


def test_batch_equals_alone():
    import torch
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import Engine
    print('case batch of 6 vs alone', flush=True)
    systems = workloads.batch_of_systems(6, (150, 900), (80, 100), seed=4)
    for precision in ('fp64', 'mixed'):
        tun = dict(pair_lanes=32, apl=4, skb=64, fkb=64)
        big = Engine([nbsystem(s) for s in systems], precision=precision)
        big.set_tuning(**tun)
        big.ensure_nlist()
        big.compute(build=False)
        res = big.results()
        for i, s in enumerate(systems):
            one = Engine([nbsystem(s)], precision=precision)
            if one.has_ewald:
                one.set_tuning(**tun)
            else:
                one.set_tuning(pair_lanes=32)
            one.ensure_nlist()
            one.compute(build=False)
            e, g, v = one.results(0)
            eb, gb, vb = res[i]
            assert e == eb and (g == gb).all() and (v == vb).all(), (precision, i)
            for lanes, other in ((16, dict(apl=2, skb=256, fkb=32, fapl=2)), (64, dict(apl=8, fkb=128, fapl=1))):
                one.set_tuning(pair_lanes=lanes)
                if one.has_ewald:
                    one.set_tuning(**other)
                one.compute(build=False)
                e, g, v = one.results(0)
                de = abs(e - eb)/abs(eb)
                dg = np.abs(g - gb).max()/np.abs(gb).max()
                rms = np.sqrt(((g - gb)**2).sum(axis=1).mean()/(gb**2).sum(axis=1).mean())
                print(precision, i, s.spec.kind, s.natom, 'launch parameters %s: rel dE %.2e, max rel dG %.2e, '
                      'rms rel dG %.2e' % (other, de, dg, rms), flush=True)
                if precision == 'fp64':
                    # summation order only
                    assert de <= 1e-11 and dg <= 1e-11, (i, other, de, dg)
                else:
                    # FP32 partial structure factors summed over other atom chunks: the mixed contract
                    assert de <= 1e-6 and rms <= 1e-5, (i, other, de, rms)
