"""The persistent slot engine on the GPU: systems admitted, retired and
re-admitted in place (free ranges reused, capacities grown) give every
member the bits of a one-system Engine with the same launch parameters, for
evaluations and for MD steps, FP64 and mixed."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]


def _systems():
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem
    systems = workloads.batch_of_systems(7, (150, 1500), (80.0, 100.0), seed=12)
    # the last one larger than all others together: the atom, k and spart spaces must grow
    systems.append(workloads.rocksalt(workloads.natom_to_reps(6000, 8), seed=3, angles=(85.0, 92.0, 97.0)))
    return systems, [nbsystem(s) for s in systems]


def _alone(nb, precision, vel=None, dts=()):
    from yaff_amd import serve
    from yaff_amd.engine import Engine
    eng = Engine([nb], precision=precision, **serve.SERVER_TUNING)
    if vel is not None:
        eng.set_velocities([vel])
    eng.ensure_nlist()
    eng.compute(build=False)
    for dt in dts:
        eng.set_timestep(dt)
        eng.md_step()
    assert eng.nlist_status()[0] == 0
    e, g, v = eng.results(0)
    a0 = eng.atom0[0]
    return e, g, v, eng.pos_d[a0:a0 + nb.natom].cpu().numpy(), eng.vel_d[a0:a0 + nb.natom].cpu().numpy()


def _check(eng, slot, ref):
    e, g, v = eng.results(slot)
    a0, n = eng.atom0[slot], eng.systems[slot].natom
    assert e == ref[0] and np.array_equal(g, ref[1]) and np.array_equal(v, ref[2]), slot
    assert np.array_equal(eng.pos_d[a0:a0 + n].cpu().numpy(), ref[3])
    assert np.array_equal(eng.vel_d[a0:a0 + n].cpu().numpy(), ref[4])


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_admit_retire_grow_matches_alone(precision):
    import torch
    from yaff_amd import serve
    from yaff_amd.slots import SlotEngine
    print('case slot engine admit/retire/grow %s' % precision, flush=True)
    nbs = _systems()[1]
    t = serve.SERVER_TUNING
    eng = SlotEngine(6, precision, pair_lanes=t['pair_lanes'], apl=t['apl'], skb=t['skb'], fkb=t['fkb'],
                     fapl=t['fapl'], atom_cap=256)

    def evaluate(members):
        eng.commit()
        eng.dt_d.zero_()
        eng.ensure_nlist()
        eng.compute(build=False)
        torch.cuda.synchronize()
        for slot, k in members.items():
            _check(eng, slot, refs[k])
    refs = {k: _alone(nb, precision) for k, nb in enumerate(nbs)}
    members = {eng.admit(nbs[k]): k for k in range(4)}
    evaluate(members)
    # retire two in the middle, re-admit into their ranges and grow for the large one
    for slot in [s for s, k in members.items() if k in (1, 2)]:
        eng.retire(slot)
        del members[slot]
    for k in (4, 7, 5):
        members[eng.admit(nbs[k])] = k
    assert eng.sp['atoms'].cap > 256 and eng.atom0[[s for s, k in members.items() if k == 4][0]] < eng.npad
    evaluate(members)
    # everything out, then one back: a batch of one in the grown engine
    for slot in list(members):
        eng.retire(slot)
    members = {eng.admit(nbs[6]): 6}
    evaluate(members)
    assert eng.npad == 256*((nbs[6].natom + 255)//256)


def test_md_steps_in_slots_match_alone():
    import torch
    from yaff_amd import serve
    from yaff_amd.batch import maxwell_boltzmann, timestep
    from yaff_amd.slots import SlotEngine
    print('case slot engine md', flush=True)
    systems, nbs = _systems()
    systems, nbs = systems[:5], nbs[:5]
    t = serve.SERVER_TUNING
    eng = SlotEngine(4, 'fp64', pair_lanes=t['pair_lanes'], apl=t['apl'], skb=t['skb'], fkb=t['fkb'],
                     fapl=t['fapl'], atom_cap=256)
    vels = [maxwell_boltzmann(nb.masses, 300.0, k) for k, nb in enumerate(nbs)]
    dts = [timestep(s) for s in systems]

    def run(members, nstep):
        eng.commit()
        eng.dt_d.zero_()
        eng.ensure_nlist()
        eng.compute(build=False)
        for _ in range(nstep):
            dt = np.zeros(eng.nsys)
            for slot, k in members.items():
                dt[slot] = dts[k]
            eng.dt_d.copy_(torch.from_numpy(dt))
            eng.md_step()
        assert eng.nlist_status()[0] == 0
    members = {eng.admit(nbs[k], vel=vels[k]): k for k in range(3)}
    run(members, 4)
    for slot, k in members.items():
        _check(eng, slot, _alone(nbs[k], 'fp64', vels[k], [dts[k]]*4))
    slot0 = [s for s, k in members.items() if k == 0][0]
    eng.retire(slot0)
    del members[slot0]
    # the members that stay continue from their state; the new one starts
    for k in (3, 4):
        members[eng.admit(nbs[k], vel=vels[k])] = k
    state = {s: (eng.pos_d[eng.atom0[s]:eng.atom0[s] + nbs[k].natom].cpu().numpy(),
                 eng.vel_d[eng.atom0[s]:eng.atom0[s] + nbs[k].natom].cpu().numpy()) for s, k in members.items()
             if k in (1, 2)}
    assert len(state) == 2
    run(members, 3)
    for slot, k in members.items():
        if k in (3, 4):
            _check(eng, slot, _alone(nbs[k], 'fp64', vels[k], [dts[k]]*3))
        else:
            # continued members: 4 steps, a force evaluation at the same positions, 3 more steps
            nb = nbs[k]
            from yaff_amd.engine import NBSystem
            cont = NBSystem(state[slot][0], nb.cell, nb.pots, rc=nb.rc, rmax=nb.rmax, stab=nb.stab,
                            ewald=nb.ewald, charges=nb.charges, masses=nb.masses)
            ref = _alone(cont, 'fp64', state[slot][1], [dts[k]]*3)
            _check(eng, slot, ref)
