"""Serving loop on the GPU: continuous batching gives every request the
result it gets alone (bit-identical: fixed launch parameters, deterministic
per-system kernels), an MD segment served in a mix equals the same segment
run eagerly on a one-system engine, admission keeps the estimated device
memory of the batch within the budget, and mixed-precision results agree
with FP64 to the stated tolerance (RMS force error < 1e-5)."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]


def _requests(seed=5, n=10, rate=None):
    from yaff_amd import serve
    rs = serve.poisson_requests(n, 1000.0 if rate is None else rate, natom_range=(100, 2500), nsteps=5,
                                seed=seed)
    return rs


def _same(a, b):
    assert a.energy == b.energy, (a.rid, a.energy, b.energy)
    assert (a.gpos == b.gpos).all() and (a.vtens == b.vtens).all(), a.rid
    if a.kind == 'md':
        assert (a.pos == b.pos).all() and (a.vel_out == b.vel_out).all(), a.rid


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_alone_equals_mix(precision):
    from yaff_amd import serve
    print('case serve mix vs alone', precision, flush=True)
    mix = _requests()
    srv = serve.Server(precision=precision, max_quantum=3)
    done = srv.run(mix, realtime=False)
    assert len(done) == len(mix)
    assert srv.stats['forms'] > 1            # the batch was re-formed as requests finished
    for r in mix:
        alone = _requests()[r.rid]
        serve.Server(precision=precision).run([alone], realtime=False)
        _same(r, alone)


def test_md_segment_matches_eager_engine():
    from yaff_amd import serve
    from yaff_amd.engine import Engine
    print('case serve md vs eager', flush=True)
    rs = [r for r in _requests(seed=7, n=12) if r.kind == 'md'][:2]
    others = [r for r in _requests(seed=8, n=6)]
    for k, r in enumerate(others):
        r.rid = 100 + k
    serve.Server(precision='fp64', max_quantum=2).run(rs + others, realtime=False)
    for r in rs:
        eng = Engine([r.nb], precision='fp64', **serve.SERVER_TUNING)
        eng.set_velocities([r.vel])
        eng.set_timestep(r.dt)
        eng.compute(build=True)
        for _ in range(r.nsteps):
            eng.md_step()
        assert eng.nlist_status()[0] == 0
        e, g, v = eng.results(0)
        a0 = eng.atom0[0]
        assert e == r.energy and (g == r.gpos).all() and (v == r.vtens).all()
        assert (eng.pos_d[a0:a0 + r.natom].cpu().numpy() == r.pos).all()
        assert (eng.vel_d[a0:a0 + r.natom].cpu().numpy() == r.vel_out).all()


def test_admission_budget_and_mixed_accuracy():
    from yaff_amd import serve
    print('case serve budget + mixed accuracy', flush=True)
    rs64 = _requests(seed=9, n=8)
    sizes = [serve.estimate_bytes([r.nb]) for r in rs64]
    budget = int(max(sizes)*1.5)
    srv = serve.Server(precision='fp64', hbm_budget=budget)
    srv.run(rs64, realtime=False)
    assert 0 < srv.peak_bytes <= budget
    rsm = _requests(seed=9, n=8)
    serve.Server(precision='mixed').run(rsm, realtime=False)
    for a, b in zip(rs64, rsm):
        if a.kind != 'eval':
            continue
        rms = np.sqrt(((a.gpos - b.gpos)**2).sum(axis=1).mean()/(a.gpos**2).sum(axis=1).mean())
        print(a.rid, a.natom, 'mixed rms force err %.2e energy err %.2e' % (rms, abs(a.energy - b.energy)/abs(a.energy)))
        assert rms < 1e-5
        assert abs(a.energy - b.energy) <= 1e-6*abs(a.energy)
    # the server's persistent engine holds, after serving one request, memory
    # within a factor 2 of the estimate for that request
    r = _requests(seed=9, n=8)[1]
    srv = serve.Server(precision='fp64', reserve_atoms=256)
    srv.run([r], realtime=False)
    used = srv.engine.device_bytes()
    est = serve.estimate_bytes([r.nb])
    print('estimate %d engine tables %d' % (est, used))
    assert 0.5*used <= est <= 2.0*used


def test_overflow_restarts_requests_with_same_bits(monkeypatch):
    """A neighbour-list overflow (row capacities forced far too small from
    the host) is caught after its quantum: capacities grow and the affected
    requests run again from their inputs, with the bits of a run without
    the overflow."""
    from yaff_amd import serve, slots
    print('case serve overflow restart', flush=True)
    ref = _requests(seed=13, n=6)
    serve.Server(precision='fp64').run(ref, realtime=False)
    monkeypatch.setattr(slots, '_nlist_kcap', lambda s: 8)
    rs = _requests(seed=13, n=6)
    srv = serve.Server(precision='fp64', max_quantum=2)
    done = srv.run(rs, realtime=False)
    assert len(done) == len(rs) and srv.stats['regrow'] >= 1
    for a, b in zip(rs, ref):
        _same(a, b)
