"""The launch-parameter autotuner on a small mixed batch (LJ fluids and
rocksalt, triclinic): it picks every parameter from its candidate set, the
tuned evaluation gives the untuned one's energies, gradients and virials up
to summation order (launch parameters change only how sums are grouped:
1e-11 relative in FP64), and a second engine of the same kind gets the same
parameters back from the cache without timing anything."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]


def _engine(systems):
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import Engine
    e = Engine([nbsystem(s) for s in systems], precision='fp64')
    e.ensure_nlist()
    return e


def test_autotune_candidates_results_and_cache(tmp_path):
    from yaff_amd import autotune, workloads
    print('case autotune, 4 systems', flush=True)
    systems = workloads.batch_of_systems(4, (150, 700), (80.0, 100.0), seed=3)
    eng = _engine(systems)
    eng.compute(build=False)
    ref = eng.results()
    cache = str(tmp_path / 'autotune.json')
    best = autotune.tune(eng, cache=cache)
    allowed = {name: values for name, values, _ in autotune.CANDIDATES if values is not None}
    for name, values in allowed.items():
        assert best[name] in values, (name, best[name])
    assert best['skb'] >= 16 and best['fkb'] >= 16
    eng.compute(build=False)
    res = eng.results()
    for i, ((e0, g0, v0), (e1, g1, v1)) in enumerate(zip(ref, res)):
        assert abs(e1 - e0) <= 1e-11*abs(e0), (i, e0, e1)
        assert np.abs(g1 - g0).max() <= 1e-11*np.abs(g0).max(), i
        assert np.abs(v1 - v0).max() <= 1e-11*np.abs(v0).max(), i
    # the same kind of batch: parameters from the cache
    eng2 = _engine(systems)
    assert autotune.cache_key(eng2) == autotune.cache_key(eng)
    assert autotune.tune(eng2, cache=cache) == best
    assert {k: v for k, v in eng2.tuning().items() if k in best} == best
