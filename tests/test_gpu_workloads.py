"""The batched and served workloads at their stated sizes.

* Batched: the 64 systems of the batched benchmark (200-6000 atoms, LJ
  fluid / rocksalt alternating, triclinic angles 80-100 degrees, seed 0) in
  one engine, FP64 and mixed: every system's energy, gradient and virial
  bit-identical to the same system in an engine of its own (same launch
  parameters); the largest LJ and the largest ionic system against the
  oracle (FP64 to 1e-10, mixed within its contract); then 200 graph-replayed
  mixed-precision MD steps with no capacity overflow, finite state and each
  system's momentum conserved.
* Served: a stream with single evaluations and 20-step MD segments of
  100-50000 atoms, among them a rocksalt request of more than 45,000 atoms:
  each result bit-identical served in the mix and served alone, FP64 and
  mixed.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]

TOL = 1e-10
TUNING = dict(pair_lanes=16, apl=4, skb=64, fkb=128, fapl=2)


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max()/np.abs(b).max()


def _oracle(s):
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
        parts.append(O.ewald_corr(s.pos, s.charges, cell, sp.alpha, None))
    return sum(p[0] for p in parts), sum(p[1] for p in parts), sum(p[2] for p in parts)


def _engine(systems, precision):
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import Engine
    eng = Engine([nbsystem(s) for s in systems], precision=precision)
    if eng.has_ewald:
        eng.set_tuning(**TUNING)
    else:
        eng.set_tuning(pair_lanes=TUNING['pair_lanes'])
    eng.ensure_nlist()
    eng.compute(build=False)
    return eng


@pytest.fixture(scope='module')
def batch64():
    from yaff_amd import workloads
    systems = workloads.batch_of_systems(64, (200, 6000), (80.0, 100.0), seed=0)
    assert len(systems) == 64 and 200 <= min(s.natom for s in systems) and max(s.natom for s in systems) <= 6100
    return systems


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_batch64_each_system_as_alone_and_oracle(batch64, precision):
    print('case batch64 %s: %d atoms' % (precision, sum(s.natom for s in batch64)), flush=True)
    big = _engine(batch64, precision)
    assert big.nlist_status()[0] == 0
    res = big.results()
    del big
    for i, s in enumerate(batch64):
        e, g, v = _engine([s], precision).results(0)
        eb, gb, vb = res[i]
        assert e == eb and np.array_equal(g, gb) and np.array_equal(v, vb), i
    for kind in ('lj', 'rocksalt'):
        i = max((k for k, s in enumerate(batch64) if (s.spec.kind == 'lj') == (kind == 'lj')),
                key=lambda k: batch64[k].natom)
        s = batch64[i]
        er, gr, vr = _oracle(s)
        e, g, v = res[i]
        if precision == 'fp64':
            assert abs(e - er) <= TOL*abs(er) and _rel(g, gr) <= TOL and _rel(v, vr) <= TOL, i
        else:
            rms = np.sqrt(((g - gr)**2).sum(axis=1).mean()/(gr**2).sum(axis=1).mean())
            print('%s %d atoms mixed vs oracle: rel dE %.2e rms dG %.2e' % (kind, s.natom, abs(e - er)/abs(er), rms))
            assert abs(e - er) <= 1e-6*abs(er) and rms <= 1e-5 and _rel(v, vr) <= 1e-6


def test_batch64_200_graph_steps(batch64):
    import torch
    from yaff_amd.batch import BatchMD
    print('case batch64 200 mixed graph steps', flush=True)
    md = BatchMD(batch64, precision='mixed', tuning=TUNING)
    md.warm(2)
    eng = md.engine
    masses = eng.masses_d.cpu().numpy()

    def momenta():
        p = eng.vel_d.cpu().numpy()*masses.reshape(-1, 1)
        return np.array([p[a0:a0 + s.natom].sum(axis=0) for a0, s in zip(eng.atom0, batch64)]), \
            np.array([np.abs(p[a0:a0 + s.natom]).sum() for a0, s in zip(eng.atom0, batch64)])
    p0, scale = momenta()
    md.capture(10)
    etot, ekin = [], []
    for _ in range(20):   # 200 steps
        md.graph.replay()
        torch.cuda.synchronize()
        ek = 0.5*eng.ekin2_d.cpu().numpy()
        ekin.append(ek)
        etot.append(eng.total_d[:, 0].cpu().numpy() + ek)
    assert eng.nlist_status()[0] == 0
    assert torch.isfinite(eng.pos_d).all() and torch.isfinite(eng.vel_d).all()
    assert torch.isfinite(eng.total_d).all()
    p1, _ = momenta()
    drift = np.abs(p1 - p0).max(axis=1)/scale
    etot, ekin = np.array(etot), np.array(ekin)
    ratio = etot.std(axis=0)/ekin.std(axis=0)
    print('momentum drift (relative to sum |p|): max %.2e; conserved-energy std / kinetic std: median %.3f '
          'max %.3f' % (drift.max(), np.median(ratio), ratio.max()), flush=True)
    # mixed precision: the reciprocal forces cancel to the phase rounding only
    assert (drift <= 1e-5).all()
    assert np.median(ratio) < 0.2 and ratio.max() < 1.0


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_served_stream_with_large_request(precision):
    from yaff_amd import serve, workloads
    print('case served stream with a 45k+ atom request, %s' % precision, flush=True)

    def stream():
        rng = np.random.default_rng(21)
        plan = [('rocksalt', 47000, 'md'), ('lj', 120, 'eval'), ('rocksalt', 900, 'md'), ('lj', 5000, 'eval'),
                ('rocksalt', 20000, 'eval'), ('lj', 2500, 'md')]
        reqs = []
        for k, (kind, n, what) in enumerate(plan):
            angles = tuple(rng.uniform(80.0, 100.0, 3))
            seed = int(rng.integers(0, 2**31 - 1))
            if kind == 'lj':
                s = workloads.lj_fluid(workloads.natom_to_reps(n, 4), seed=seed, angles=angles)
            else:
                s = workloads.rocksalt(workloads.natom_to_reps(n, 8), seed=seed, angles=angles)
            reqs.append(serve.Request(k, s, what, 20 if what == 'md' else 0, arrival=0.001*k))
        return reqs
    mix = stream()
    assert max(r.natom for r in mix) >= 45000
    srv = serve.Server(precision=precision, max_quantum=4)
    done = srv.run(mix, realtime=False)
    assert len(done) == len(mix) and srv.stats['forms'] > 1
    for r in mix:
        alone = stream()[r.rid]
        serve.Server(precision=precision).run([alone], realtime=False)
        assert r.energy == alone.energy and np.array_equal(r.gpos, alone.gpos), r.rid
        assert np.array_equal(r.vtens, alone.vtens), r.rid
        if r.kind == 'md':
            assert np.array_equal(r.pos, alone.pos) and np.array_equal(r.vel_out, alone.vel_out), r.rid
