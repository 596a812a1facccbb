"""Device MD: velocity Verlet (yaff/sampling/verlet.py:339-367) on the GPU.

* A HIP graph holding one step, replayed several times with eager work and
  synchronisations in between, reproduces eager steps bit for bit (every
  kernel is deterministic: no floating-point atomics, fixed reduction order).
* NVE energy conservation of the 864-atom LJ fluid (FCC 6x6x6, L = 34.7 A,
  sigma 3.405 A, eps 3.79e-4 Ha, rcut 12 A, Switch3 4 A, dt 2 fs, 120 K),
  the criterion of yaff/pes/test/test_md.py:100-126 (std of the conserved
  energy small against the kinetic energy fluctuations).
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]


def _ionic_engine(reps=(3, 3, 2), precision='fp64'):
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import Engine
    s = workloads.rocksalt(reps, seed=3, angles=(88.0, 91.0, 93.0))
    return s, Engine([nbsystem(s)], precision=precision)


def _lj_engine():
    from yaff_amd import workloads, _lib
    from yaff_amd.engine import Engine, NBSystem, PotSpec
    s = workloads.lj_fluid((6, 6, 6), seed=0)
    sp = s.spec
    n = s.natom
    pot = PotSpec(_lib.POT_LJ, sp.rcut, _lib.TR_SWITCH3, sp.width, np.full(n, sp.sigma), np.full(n, sp.epsilon))
    return s, Engine([NBSystem(s.pos, s.cell, [pot], rc=sp.rcut, masses=s.masses)])


def test_graph_replay_matches_eager():
    import torch
    from yaff_amd.batch import maxwell_boltzmann
    from yaff_amd.units import femtosecond
    print('case rocksalt 144 atoms, graph vs eager', flush=True)
    s, eng = _ionic_engine()
    eng.ensure_nlist()
    eng.compute(build=False)
    eng.set_velocities([maxwell_boltzmann(s.masses, 300.0, 2)])
    eng.set_timestep(1.0*femtosecond)
    for _ in range(3):
        eng.md_step()
    torch.cuda.synchronize()
    assert eng.nlist_status()[0] == 0
    state = [t.clone() for t in (eng.pos_d, eng.vel_d, eng.gpos_d)]
    eager = []
    for _ in range(4):
        eng.md_step()
        eager.append((eng.pos_d.clone(), eng.vel_d.clone(), eng.total_d.clone()))
    torch.cuda.synchronize()
    for t, v in zip((eng.pos_d, eng.vel_d, eng.gpos_d), state):
        t.copy_(v)
    g = eng.capture_md(1)
    for k in range(4):
        g.replay()
        torch.cuda.synchronize()
        scratch = torch.randn(1000, device='cuda').sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(eng.pos_d, eager[k][0]), k
        assert torch.equal(eng.vel_d, eager[k][1]), k
        assert torch.equal(eng.total_d, eager[k][2]), k
        del scratch
    assert eng.nlist_status()[0] == 0


def test_nve_energy_conservation_lj864():
    import torch
    from yaff_amd.batch import maxwell_boltzmann
    from yaff_amd.units import femtosecond
    print('case LJ 864 NVE 100 steps', flush=True)
    s, eng = _lj_engine()
    eng.ensure_nlist()
    eng.compute(build=False)
    eng.set_velocities([maxwell_boltzmann(s.masses, 120.0, 0)])
    eng.set_timestep(2.0*femtosecond)
    g = eng.capture_md(1)
    econs, ekin = [], []
    for _ in range(100):
        g.replay()
        tot = eng.total_d.cpu().numpy()[0]
        e2 = eng.ekin2_d.cpu().numpy()[0]
        econs.append(tot[0] + 0.5*e2)
        ekin.append(0.5*e2)
    assert eng.nlist_status()[0] == 0
    econs = np.array(econs)
    ekin = np.array(ekin)
    print('std econs', econs.std(), 'std ekin', ekin.std(), flush=True)
    assert econs.std() < 0.05*ekin.std()


def test_skin_md_matches_rebuild_every_step():
    """A Verlet skin (rebuild only after a displacement of skin/2, decided on
    the device) gives the trajectory of rebuilding every step: the pair set
    within rcut is the same, only summation orders differ."""
    import torch
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem, maxwell_boltzmann
    from yaff_amd.engine import Engine
    from yaff_amd.units import femtosecond, angstrom
    print('case rocksalt 432 atoms, skin 0.15 A vs 0, 60 steps', flush=True)
    s = workloads.rocksalt((3, 3, 6), seed=5, angles=(87.0, 92.0, 95.0))
    vel = maxwell_boltzmann(s.masses, 300.0, 4)
    engs = []
    for skin in (0.0, 0.15*angstrom):
        e = Engine([nbsystem(s, skin)], precision='fp64')
        e.ensure_nlist()
        e.compute(build=False)
        e.set_velocities([vel])
        e.set_timestep(1.0*femtosecond)
        engs.append(e)
    rebuilds = 0
    for step in range(60):
        for e in engs:
            e.md_step()
        rebuilds += int(engs[1].rebuild_d[0].item())
    assert engs[1].nlist_status()[0] == 0 and engs[0].nlist_status()[0] == 0
    p0, p1 = engs[0].pos_d.cpu().numpy(), engs[1].pos_d.cpu().numpy()
    e0, e1 = engs[0].total_d.cpu().numpy()[0, 0], engs[1].total_d.cpu().numpy()[0, 0]
    print('rebuilds with skin: %d of 60; max |dpos| %.2e, rel dE %.2e' % (rebuilds, np.abs(p0 - p1).max(),
                                                                      abs(e0 - e1)/abs(e0)), flush=True)
    assert 1 <= rebuilds < 60
    assert np.abs(p0 - p1).max() < 1e-9
    assert abs(e0 - e1) <= 1e-10*abs(e0)


def test_skin_graph_replay_matches_eager():
    import torch
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem, maxwell_boltzmann
    from yaff_amd.engine import Engine
    from yaff_amd.units import femtosecond, angstrom
    print('case LJ 256 + rocksalt 256, skin 0.8 A, graph vs eager', flush=True)
    systems = [workloads.lj_fluid((4, 4, 4), seed=1, angles=(85.0, 95.0, 90.0)),
               workloads.rocksalt((4, 4, 2), seed=2, angles=(91.0, 88.0, 93.0))]
    vel = [maxwell_boltzmann(x.masses, 300.0, 7 + i) for i, x in enumerate(systems)]

    def make():
        e = Engine([nbsystem(x, 0.8*angstrom) for x in systems], precision='mixed')
        e.ensure_nlist()
        e.compute(build=False)
        e.set_velocities(vel)
        e.set_timestep([2.0*femtosecond, 1.0*femtosecond])
        return e
    ea, eb = make(), make()
    for _ in range(12):
        ea.md_step()
    g = eb.capture_md(3)
    for _ in range(4):
        g.replay()
        torch.cuda.synchronize()
    assert ea.nlist_status()[0] == 0 and eb.nlist_status()[0] == 0
    assert torch.equal(ea.pos_d, eb.pos_d) and torch.equal(ea.vel_d, eb.vel_d)
    assert torch.equal(ea.total_d, eb.total_d)


@pytest.mark.parametrize('kind', ['ionic', 'batch'])
def test_joined_steps_match_single_steps(kind):
    """md_steps(n) joins the kick ending each step with the kick and drift
    starting the next in one kernel: positions, velocities, gradients,
    energy totals and kinetic energies bit for bit those of n md_step calls,
    eagerly and replayed from a graph."""
    import torch
    from yaff_amd.batch import maxwell_boltzmann
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import Engine
    from yaff_amd.units import femtosecond
    print('case joined steps, %s' % kind, flush=True)
    if kind == 'ionic':
        s, eng = _ionic_engine()
        systems = [s]
    else:
        systems = workloads.batch_of_systems(5, (150, 700), (80.0, 100.0), seed=6)
        eng = Engine([nbsystem(x) for x in systems], precision='mixed')
    eng.ensure_nlist()
    eng.compute(build=False)
    eng.set_velocities([maxwell_boltzmann(x.masses, 300.0, 2 + i) for i, x in enumerate(systems)])
    eng.set_timestep(1.0*femtosecond)
    eng.md_step()
    torch.cuda.synchronize()
    assert eng.nlist_status()[0] == 0
    bufs = (eng.pos_d, eng.vel_d, eng.gpos_d, eng.total_d, eng.ekin2_d)
    state = [t.clone() for t in bufs]

    def restore():
        torch.cuda.synchronize()
        for t, v in zip(bufs, state):
            t.copy_(v)
        torch.cuda.synchronize()

    for _ in range(4):
        eng.md_step()
    ref = [t.clone() for t in bufs]
    restore()
    eng.md_steps(4)
    torch.cuda.synchronize()
    for i, (t, r) in enumerate(zip(bufs, ref)):
        assert torch.equal(t, r), i
    restore()
    g = eng.capture_md(4)
    restore()
    g.replay()
    torch.cuda.synchronize()
    for i, (t, r) in enumerate(zip(bufs, ref)):
        assert torch.equal(t, r), ('graph', i)
    assert eng.nlist_status()[0] == 0
