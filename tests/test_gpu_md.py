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
    from bench import rocksalt_system
    from yaff_amd.engine import Engine
    s, nb = rocksalt_system(reps, seed=3, angles=(88.0, 91.0, 93.0))
    return s, Engine([nb], precision=precision)


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
    from bench import maxwell_boltzmann
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
    from bench import maxwell_boltzmann
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
