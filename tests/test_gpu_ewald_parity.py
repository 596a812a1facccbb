"""Ewald atom groups, the batched Ewald correction and the device Verlet step
against the oracle.

* ForcePartEwaldReciprocal with nlow/nhigh (ewald.c:61-95, 104-107) through
  the drop-in, alone and next to a neighbour list whose own nlow/nhigh
  differ: the force pass with atom groups (k_ewald_force_groups) and the
  grouped k sums, water32 at the reference's (15,33), (0,12), (81,96)
  (yaff/pes/test/test_ewald.py:288-300) and rocksalt at (10,40), (40,40),
  (0,50); energy, gradient and virial to 1e-10.
* ForcePartEwaldCorrection (ewald.c:278-349) with scalings, nlow/nhigh and
  a dielectric, through the drop-in, and the correction inside a batch of
  the engine (bonded water32 next to rocksalt and an LJ fluid).
* One device velocity-Verlet step, then two joined ones, against
  yaff/sampling/verlet.py:339-367 restated in numpy on oracle forces, for a
  bonded water32 (scaled pairs, so the correction's gradient is part of the
  step) and a rocksalt box in one batch.
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


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max()/max(np.abs(b).max(), 1e-300)


def _water32():
    import cases
    import yaff_amd as ya
    f = cases.fixture_arrays('water32')
    return ya.System(f['numbers'], f['pos'], ffatypes=['O', 'H', 'H']*32, bonds=f['bonds'], rvecs=f['rvecs'],
                     charges=f['charges'])


def _rocksalt_system(reps=(3, 3, 2), seed=4, angles=(88.0, 92.0, 90.5)):
    from yaff_amd import workloads
    return workloads.rocksalt(reps, seed=seed, angles=angles)


def _reci_ref(system, alpha, gcut, nlow, nhigh, diel=1.0):
    from oracle import oracle as orc
    O = orc.Oracle()
    cell = O.cell(system.cell.rvecs)
    return O.ewald_reci(system.pos, system.charges, cell, alpha, orc.gmax_of(cell, gcut), gcut, diel, nlow, nhigh)


def _check(got, ref):
    e, g, v = got
    er, gr, vr = ref
    assert abs(e - er) <= TOL*abs(er), (e, er)
    assert _rel(g, gr) <= TOL
    assert _rel(v, vr) <= TOL


@pytest.mark.parametrize('nlow,nhigh', [(15, 33), (0, 12), (81, 96), (10, 40), (40, 40), (0, 50)])
def test_reci_groups_water32_dropin(nlow, nhigh):
    import yaff_amd as ya
    print('case water32 reci nlow=%d nhigh=%d' % (nlow, nhigh), flush=True)
    system = _water32()
    alpha, gcut = 0.35, 0.4
    part = ya.ForcePartEwaldReciprocal(system, alpha, gcut, nlow=nlow, nhigh=nhigh)
    ff = ya.ForceField(system, [part])
    g = np.zeros((system.natom, 3))
    v = np.zeros((3, 3))
    e = ff.compute(g, v)
    _check((e, g, v), _reci_ref(system, alpha, gcut, nlow, nhigh))


def test_reci_groups_next_to_other_nlist_groups():
    """The reciprocal part's groups are its own: a neighbour list (and pair
    part) with nlow=5, nhigh=70 next to a reciprocal part with (15, 33) and a
    dielectric."""
    import yaff_amd as ya
    from oracle import oracle as orc
    print('case water32 reci (15,33) with nlist (5,70)', flush=True)
    system = _water32()
    alpha, gcut, diel = 0.3, 0.45, 1.3
    nlist = ya.NeighborList(system, nlow=5, nhigh=70)
    sc = ya.Scalings(system, 0.0, 0.0, 0.5)
    ei = ya.PairPotEI(system.charges, alpha, 9.0*ANG)
    pair = ya.ForcePartPair(system, nlist, sc, ei)
    reci = ya.ForcePartEwaldReciprocal(system, alpha, gcut, dielectric=diel, nlow=15, nhigh=33)
    ff = ya.ForceField(system, [pair, reci], nlist)
    ff.compute(np.zeros((system.natom, 3)), np.zeros((3, 3)))
    _check((reci.energy, reci.gpos, reci.vtens), _reci_ref(system, alpha, gcut, 15, 33, diel))
    O = orc.Oracle()
    cell = O.cell(system.cell.rvecs)
    rows = O.nlist_build(system.pos, cell, 9.0*ANG, orc.rmax_of(cell, 9.0*ANG), 5, 70)
    assert nlist.neighs.tobytes() == rows.tobytes()
    ref = O.pair_compute(rows, sc.stab, orc.Pot(orc.EI, 9.0*ANG, p0=system.charges, alpha=alpha), system.natom)
    _check((pair.energy, pair.gpos, pair.vtens), ref)


@pytest.mark.parametrize('nlow,nhigh', [(10, 40), (40, 40), (0, 50), (72, 144)])
def test_reci_groups_rocksalt(nlow, nhigh):
    import yaff_amd as ya
    print('case rocksalt144 reci nlow=%d nhigh=%d' % (nlow, nhigh), flush=True)
    s = _rocksalt_system()
    assert s.natom == 144
    sp = s.spec
    part = ya.ForcePartEwaldReciprocal(s, sp.alpha, sp.gcut, nlow=nlow, nhigh=nhigh)
    g = np.zeros((s.natom, 3))
    v = np.zeros((3, 3))
    e = ya.ForceField(s, [part]).compute(g, v)
    _check((e, g, v), _reci_ref(s, sp.alpha, sp.gcut, nlow, nhigh))


@pytest.mark.parametrize('nlow,nhigh,diel', [(0, -1, 1.0), (15, 33, 1.0), (0, 12, 1.4), (81, 96, 1.0)])
def test_corr_dropin_water32(nlow, nhigh, diel):
    import yaff_amd as ya
    from oracle import oracle as orc
    print('case water32 corr nlow=%d nhigh=%d diel=%.1f' % (nlow, nhigh, diel), flush=True)
    system = _water32()
    alpha = 0.35
    sc = ya.Scalings(system, 0.0, 0.25, 0.5)
    part = ya.ForcePartEwaldCorrection(system, alpha, sc, dielectric=diel, nlow=nlow, nhigh=nhigh)
    g = np.zeros((system.natom, 3))
    v = np.zeros((3, 3))
    e = ya.ForceField(system, [part]).compute(g, v)
    O = orc.Oracle()
    ref = O.ewald_corr(system.pos, system.charges, O.cell(system.cell.rvecs), alpha, sc.stab, diel, nlow, nhigh)
    _check((e, g, v), ref)
    # a second call (cached tables) after the atoms moved
    system.pos[:] += np.random.default_rng(1).normal(0, 0.02, system.pos.shape)
    ff = ya.ForceField(system, [part])
    ff.update_pos(system.pos)
    g[:] = 0
    v[:] = 0
    e = ff.compute(g, v)
    ref = O.ewald_corr(system.pos, system.charges, O.cell(system.cell.rvecs), alpha, sc.stab, diel, nlow, nhigh)
    _check((e, g, v), ref)


# ---------------------------------------------------------------------------
# the engine: correction inside a batch, and the Verlet step

def _water32_nb(alpha=0.35, gcut=0.4):
    import yaff_amd as ya
    from yaff_amd import _lib
    from yaff_amd.engine import NBSystem, PotSpec, EwaldSpec
    from yaff_amd.units import amu
    s = _water32()
    n = s.natom
    rng = np.random.default_rng(6)
    sig = rng.uniform(2.6, 3.2, n)*ANG
    eps = rng.uniform(1e-4, 4e-4, n)
    st_lj = ya.Scalings(s, 0.0, 0.0, 1.0).stab
    st_ei = ya.Scalings(s, 0.0, 0.0, 0.5).stab
    pots = [PotSpec(_lib.POT_LJ, 8.0*ANG, _lib.TR_SWITCH3, 3.0*ANG, sig, eps, stab=st_lj),
            PotSpec(_lib.POT_EI, 8.0*ANG, p0=s.charges, alpha=alpha, stab=st_ei)]
    s.masses = np.where(s.numbers == 8, 15.999, 1.008)*amu
    nb = NBSystem(s.pos, s.cell, pots, stab=st_ei, ewald=EwaldSpec(alpha, gcut, correction=True),
                  charges=s.charges, masses=s.masses)
    return s, nb


def _orc_pot(ps):
    from oracle import oracle as orc
    from yaff_amd import _lib
    kind = {_lib.POT_LJ: orc.LJ, _lib.POT_MM3: orc.MM3, _lib.POT_EXPREP: orc.EXPREP, _lib.POT_EI: orc.EI}[ps.kind]
    if ps.kind == _lib.POT_EXPREP:
        return orc.Pot(kind, ps.rcut, ps.trunc, ps.trunc_par, ffatype_ids=np.asarray(ps.p0).astype(np.int64),
                       amp_cross=ps.amp_cross, b_cross=ps.b_cross)
    if ps.kind == _lib.POT_MM3:
        return orc.Pot(kind, ps.rcut, ps.trunc, ps.trunc_par, p0=ps.p0, p1=ps.p1,
                       onlypauli=np.asarray(ps.p2).astype(np.int32))
    return orc.Pot(kind, ps.rcut, ps.trunc, ps.trunc_par, p0=ps.p0, p1=ps.p1, alpha=ps.alpha,
                   dielectric=ps.dielectric)


def _oracle_force_field(nb, pos):
    """(E, gpos, vtens) of an engine NBSystem at positions pos by the
    oracle: its pair potentials (each with its scaling rows) on Yaff's rows,
    the Ewald reciprocal term and, when requested, the correction."""
    from oracle import oracle as orc
    O = orc.Oracle()
    cell = O.cell(nb.cell.rvecs)
    rows = O.nlist_build(pos, cell, nb.rc, orc.rmax_of(cell, nb.rc))
    parts = [O.pair_compute(rows, nb.stab if p.stab is None else p.stab, _orc_pot(p), nb.natom) for p in nb.pots]
    ew = nb.ewald
    if ew is not None:
        parts.append(O.ewald_reci(pos, nb.charges, cell, ew.alpha, orc.gmax_of(cell, ew.gcut), ew.gcut,
                                  ew.dielectric, ew.nlow, ew.nhigh))
        if ew.correction:
            parts.append(O.ewald_corr(pos, nb.charges, cell, ew.alpha, nb.stab, ew.dielectric, ew.nlow, ew.nhigh))
    return sum(p[0] for p in parts), sum(p[1] for p in parts), sum(p[2] for p in parts)


def _batch():
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem
    w, nbw = _water32_nb()
    r = _rocksalt_system()
    lj = workloads.lj_fluid((2, 2, 3), seed=3, angles=(85.0, 95.0, 90.0))
    return [w, r, lj], [nbw, nbsystem(r), nbsystem(lj)]


def test_engine_batch_with_correction():
    from yaff_amd.engine import Engine
    print('case batch water32 (corrected) + rocksalt144 + lj48', flush=True)
    systems, nbs = _batch()
    eng = Engine(nbs, precision='fp64')
    assert eng.has_corr and eng.hsys[1].ew_const != 0.0
    eng.ensure_nlist()
    eng.compute(build=False)
    for i, nb in enumerate(nbs):
        _check(eng.results(i), _oracle_force_field(nb, nb.pos))
    # scaled pairs with an Ewald part but no correction are refused for MD and compute
    from yaff_amd.engine import EwaldSpec
    w, nbw = _water32_nb()
    nbw.ewald = EwaldSpec(nbw.ewald.alpha, nbw.ewald.gcut)
    with pytest.raises(ValueError):
        Engine([nbw]).compute()


def test_verlet_step_matches_restated_verlet():
    """verlet.py:344-357 in numpy on oracle forces: acc = -gpos/m,
    vel += 0.5 acc dt, pos += dt vel, forces, acc = -gpos/m, vel += 0.5 acc dt;
    one device step (kick-drift, forces, kick) and then two more (the joined
    kick-kick-drift between them)."""
    import torch
    from yaff_amd.engine import Engine
    from yaff_amd.batch import maxwell_boltzmann
    from yaff_amd.units import femtosecond
    print('case Verlet water32 (corrected) + rocksalt144', flush=True)
    systems, nbs = _batch()
    systems, nbs = systems[:2], nbs[:2]
    eng = Engine(nbs, precision='fp64')
    eng.ensure_nlist()
    eng.compute(build=False)
    dts = [0.5*femtosecond, 1.0*femtosecond]
    vel = [maxwell_boltzmann(s.masses, 300.0, 10 + i) for i, s in enumerate(systems)]
    eng.set_velocities(vel)
    eng.set_timestep(dts)
    state = []
    for i, nb in enumerate(nbs):
        state.append([nb.pos.copy(), vel[i].copy(), _oracle_force_field(nb, nb.pos)[1]])

    def restated_step(i):
        pos, v, g = state[i]
        m = nbs[i].masses.reshape(-1, 1)
        acc = -g/m
        v = v + 0.5*acc*dts[i]
        pos = pos + dts[i]*v
        e, g, vt = _oracle_force_field(nbs[i], pos)
        acc = -g/m
        v = v + 0.5*acc*dts[i]
        state[i] = [pos, v, g]
        return e, vt

    for nsteps in (1, 2):
        eng.md_steps(nsteps)
        torch.cuda.synchronize()
        assert eng.nlist_status()[0] == 0
        for i in range(2):
            for _ in range(nsteps):
                e, vt = restated_step(i)
            a0, n = eng.atom0[i], systems[i].natom
            pos, v, g = state[i]
            assert _rel(eng.pos_d[a0:a0 + n].cpu().numpy(), pos) <= 1e-14
            assert _rel(eng.vel_d[a0:a0 + n].cpu().numpy(), v) <= TOL
            eg, gg, vg = eng.results(i)
            _check((eg, gg, vg), (e, g, vt))
            ekin2 = float(eng.ekin2_d[i].item())
            ref = float((nbs[i].masses.reshape(-1, 1)*v*v).sum())
            assert abs(ekin2 - ref) <= TOL*ref
