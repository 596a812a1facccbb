"""Host-side logic without a GPU: the Ewald k-vector set and its launch
tables, scaling tables, the drop-in classes' bookkeeping, workload
generators and the serving loop's admission."""
import itertools
import math
import os
import sys

import numpy as np
import pytest


def _brute_kvectors(rvecs, alpha, gcut):
    """The k set of compute_ewald_reci (ewald.c:41-58) by plain loops: g2 >= 0,
    half space, |k| <= 2 pi gcut, visited g2, g1, g0 ascending."""
    gvecs = np.linalg.inv(rvecs).T
    gsp = 1.0/np.linalg.norm(rvecs, axis=1)   # cell.c: gspacing_i = 1/|r_i|
    gmax = np.ceil(gcut/gsp - 0.5).astype(int)
    out = []
    vol = abs(np.linalg.det(rvecs))
    for g2 in range(0, gmax[2] + 1):
        for g1 in range(-gmax[1], gmax[1] + 1):
            for g0 in range(-gmax[0], gmax[0] + 1):
                if g2 == 0 and (g1 < 0 or (g1 == 0 and g0 <= 0)):
                    continue
                k = 2*math.pi*(g0*gvecs[0] + g1*gvecs[1] + g2*gvecs[2])
                ksq = (k**2).sum()
                if ksq > (2*math.pi*gcut)**2:
                    continue
                c = 4*math.pi/vol*math.exp(-ksq/(4*alpha*alpha))/ksq
                out.append((g0, g1, g2, k[0], k[1], k[2], c))
    return gmax, np.array(out)


@pytest.mark.parametrize('angles', [(90.0, 90.0, 90.0), (84.0, 97.0, 101.0)])
def test_kvectors_brute_force(angles):
    from yaff_amd import workloads
    from yaff_amd.engine import ewald_kvectors
    s = workloads.rocksalt((2, 3, 2), angles=angles)
    sp = s.spec
    gmax, g, kv = ewald_kvectors(s.cell, sp.alpha, sp.gcut)
    gmax_b, ref = _brute_kvectors(s.cell.rvecs, sp.alpha, sp.gcut)
    assert (gmax == gmax_b).all()
    assert (g == ref[:, :3].astype(int)).all()
    scale = np.abs(ref[:, 3:6]).max()
    assert np.abs(kv[:, :3] - ref[:, 3:6]).max() <= 1e-13*scale
    assert np.allclose(kv[:, 3], ref[:, 6], rtol=1e-12, atol=0)


def test_ewald_local_plan_covers_every_k_and_atom():
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import _ewald_local_plan, ewald_kvectors_cached
    s = workloads.rocksalt((3, 2, 2), angles=(88.0, 93.0, 95.0))
    nb = nbsystem(s)
    nb.ewald.nlow, nb.ewald.nhigh = 5, s.natom - 9
    g = ewald_kvectors_cached(nb)[1]
    for apl, skb, fkb in ((2, 16, 32), (4, 64, 128), (8, 512, 16)):
        p = _ewald_local_plan(nb, apl, skb, fkb)
        nk = p['nk']
        assert nk == len(g)
        # rows tile the k vectors: consecutive g0 at fixed (g1, g2), starting at krow_k
        covered = np.zeros(nk, int)
        for (g0, g1, g2, ln), k in zip(p['krow'], p['krow_k']):
            assert ln >= 1
            covered[k:k + ln] += 1
            gg = g[k:k + ln]
            assert (gg[:, 0] == g0 + np.arange(ln)).all() and (gg[:, 1] == g1).all() and (gg[:, 2] == g2).all()
        assert (covered == 1).all()
        assert (np.diff(p['krow_k']) == p['krow'][:-1, 3]).all()
        # row blocks tile the rows in order; a block exceeds its k limit only when it is one row
        for blk, kmax in ((p['skblk'], skb), (p['fblk'], fkb)):
            assert blk[0, 0] == 0 and blk[-1, 1] == p['nrow'] and (blk[1:, 0] == blk[:-1, 1]).all()
            for r0, r1 in blk:
                assert r1 > r0
                nkb = p['krow'][r0:r1, 3].sum()
                assert nkb <= kmax or r1 == r0 + 1
            # greedy: a block and the next row together exceed the limit
            for (r0, r1), (q0, _) in zip(blk[:-1], blk[1:]):
                assert p['krow'][r0:r1, 3].sum() + p['krow'][q0, 3] > kmax
        # chunks tile the atoms, with the nlow/nhigh groups
        cov = np.zeros(s.natom, int)
        for (_, a, ln, grp) in p['chunks']:
            assert 1 <= ln <= 64*apl
            cov[a:a + ln] += 1
            want = 1 if a < 5 else (2 if a >= s.natom - 9 else 0)
            assert grp == want and (a + ln <= 5 or a >= 5) and (a + ln <= s.natom - 9 or a >= s.natom - 9)
        assert (cov == 1).all()
        # force records: rows start at multiples of 4 records, each k at its row's start plus its place
        assert (p['krow_f'] % 4 == 0).all() and p['krow_f'][0] == 0
        assert (np.diff(p['krow_f']) == (p['krow'][:-1, 3] + 3)//4*4).all()
        assert p['nrec'] == p['krow_f'][-1] + (p['krow'][-1, 3] + 3)//4*4
        for (g0, g1, g2, ln), k, f in zip(p['krow'], p['krow_k'], p['krow_f']):
            assert (p['kpos'][k:k + ln] == f + np.arange(ln)).all()
        assert (p['kpos'][nk:] == -1).all() and len(set(p['kpos'][:nk])) == nk
        # work items: every (structure-factor block, chunk) pair once
        from yaff_amd.engine import work_table
        wk = work_table(p, 3, 10, 20)
        assert (wk[:, 0] == 3).all() and (wk[:, 3] == 0).all()
        pairs = [(int(w[1]) - 10, int(w[2]) - 20) for w in wk]
        assert sorted(pairs) == sorted(itertools.product(range(len(p['skblk'])), range(len(p['chunks']))))


def test_scalings_match_oracle_table():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    import cases
    from oracle.oracle import scaling_table
    import yaff_amd as ya
    f = cases.fixture_arrays('water32')
    system = ya.System(f['numbers'], f['pos'], bonds=f['bonds'], rvecs=f['rvecs'])
    for sc in [(0.0, 0.0, 0.5), (0.2, 0.4, 0.8, 0.9), (1.0, 1.0, 1.0)]:
        mine = ya.Scalings(system, *sc).stab
        ref = scaling_table(system.natom, f['bonds'], *sc)
        assert mine.tobytes() == np.asarray(ref, dtype=mine.dtype).tobytes()
    with pytest.raises(ValueError):
        ya.Scalings(system, 1.5)


def test_neighborlist_rmax_and_arguments():
    import yaff_amd as ya
    from yaff_amd import workloads
    s = workloads.lj_fluid((2, 2, 2), angles=(80.0, 95.0, 100.0))
    nl = ya.NeighborList(s, skin=1.0)
    nl.request_rcut(10.0)
    nl.request_rcut(7.0)
    assert nl.rcut == 10.0
    assert (nl.rmax == np.ceil(11.0/s.cell.rspacings - 0.5).astype(int)).all()
    with pytest.raises(ValueError):
        ya.NeighborList(s, skin=-1)
    with pytest.raises(ValueError):
        ya.NeighborList(s, nlow=-1)
    with pytest.raises(ValueError):
        ya.NeighborList(s, nlow=10, nhigh=5)


def test_workloads():
    from yaff_amd import workloads
    from yaff_amd.units import angstrom
    s = workloads.rocksalt((12, 12, 6))
    assert s.natom == 6912 and abs(s.charges.sum()) == 0
    l = workloads.lj_fluid()
    assert l.natom == 864 and abs(l.cell.volume - (34.7*angstrom)**3) < 1e-6*l.cell.volume
    for n, per in ((200, 4), (6000, 8), (50000, 8), (100, 4)):
        r = workloads.natom_to_reps(n, per)
        assert abs(np.prod(r)*per - n) <= 0.25*n
    b = workloads.batch_of_systems(8, (200, 6000), (80.0, 100.0), seed=3)
    assert [x.spec.kind for x in b] == ['lj', 'ionic']*4
    for x in b:
        ang = np.degrees(np.arccos([np.dot(x.cell.rvecs[1], x.cell.rvecs[2]),
                                    np.dot(x.cell.rvecs[0], x.cell.rvecs[2]),
                                    np.dot(x.cell.rvecs[0], x.cell.rvecs[1])]
                                   / np.array([np.linalg.norm(x.cell.rvecs[i])*np.linalg.norm(x.cell.rvecs[j])
                                               for i, j in ((1, 2), (0, 2), (0, 1))])))
        assert ((ang > 79.999) & (ang < 100.001)).all()


def test_poisson_stream_and_admission():
    from yaff_amd import serve
    rs = serve.poisson_requests(400, 50.0, natom_range=(100, 800), seed=2)
    gaps = np.diff([0.0] + [r.arrival for r in rs])
    assert abs(gaps.mean()*50.0 - 1.0) < 0.15
    assert 0.35 < np.mean([r.kind == 'md' for r in rs]) < 0.65
    assert all(r.iters_total == (21 if r.kind == 'md' else 1) for r in rs)
    # admission: FIFO while the estimated batch footprint fits the budget
    few = rs[:6]
    one = max(serve.estimate_bytes([r.nb]) for r in few)
    srv = serve.Server(precision='fp64', hbm_budget=int(2.2*one), max_systems=64)
    srv.waiting.extend(few)
    srv._admit(0.0)
    assert 1 <= len(srv.active) < len(few)
    assert serve.estimate_bytes([r.nb for r in srv.active]) <= srv.hbm_budget
    assert [r.rid for r in srv.active] == [r.rid for r in few[:len(srv.active)]]
    # the estimate grows with every member
    sizes = [serve.estimate_bytes([r.nb for r in few[:k]]) for k in range(1, 7)]
    assert all(b > a for a, b in zip(sizes, sizes[1:]))
    # a request that can never fit is refused loudly
    srv = serve.Server(precision='fp64', hbm_budget=1000)
    srv.waiting.append(rs[0])
    with pytest.raises(RuntimeError):
        srv._admit(0.0)


def test_bench_reads_committed_traffic_summary():
    """bench.py takes each kernel's HBM bytes per launch from the committed
    rocprofv3 summary: the instance the launch parameters select when the
    summary has it, else an FP64 instance of other launch parameters, named
    in the result; the dominant kernel and the build are present."""
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    tun = {'pair_lanes': 32, 'apl': 4, 'skb': 16, 'fapl': 1, 'fkb': 128}
    assert bench.fp64_instance('ewald_force', tun) == 'k_ewald_force<0, 1>'
    assert bench.fp64_instance('pair', dict(tun, pair_lanes=16)) == 'k_pair_form<0, 2, 16>'
    assert bench.fp64_instance('nlist_bin_order', tun) == 'k_bin_order'
    assert bench.fp64_instance('no_such_kernel', tun) is None
    for k in ('ewald_force', 'ewald_sk', 'pair', 'nlist_build_binned', 'verlet_kick_kick_drift'):
        tr = bench.profiled_traffic(k, tun)
        assert tr is not None, k
        rd, wr, name = tr
        assert rd > 0 and wr > 0 and name.startswith(bench.ROCPROF_NAMES[k]), (k, tr)
    # an instance the summary lacks falls back to one of the same kernel, named
    rd, wr, name = bench.profiled_traffic('pair', dict(tun, pair_lanes=64))
    assert name.startswith('k_pair_form<0,')
    assert bench.profiled_traffic('nlist_skin_check', tun) is None
    # the build writes its 8-byte entries: at least 8 bytes per half pair of the 6912-atom workload, twice over
    assert bench.profiled_traffic('nlist_build_binned', tun)[1] >= 8*2*860000


def test_corr_plan_tables():
    """The host tables of ynb_ewald_corr: each atom's row references, summed
    in their order with the rows' delta * g (here from numpy), give the
    oracle's correction gradient; the self term is the oracle's energy with
    no scaled pairs; row blocks cover every system (one at least)."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden'))
    import cases
    from oracle import oracle as orc
    from yaff_amd.engine import corr_plan
    from yaff_amd import Scalings, System
    from math import erf
    f = cases.fixture_arrays('water32')
    s = System(f['numbers'], f['pos'], bonds=f['bonds'], rvecs=f['rvecs'], charges=f['charges'])
    st = Scalings(s, 0.0, 0.25, 0.5).stab
    alpha, nlow, nhigh = 0.35, 15, 81
    npad = 512
    entries = [None, dict(atom0=256, natom=s.natom, alpha=alpha, dielectric=1.0, nlow=nlow, nhigh=nhigh, stab=st,
                          charges=s.charges)]
    p = corr_plan(entries, npad)
    cs = p['csys']
    assert list(cs['nblk']) == [1, 1] and p['nblk_tot'] == 2 and list(p['blk_sys']) == [0, 1]
    assert cs[1]['nrow'] == len(st) == p['nrow_tot'] and cs[0]['self_energy'] == 0.0
    O = orc.Oracle()
    cell = O.cell(s.cell.rvecs)
    e_self = O.ewald_corr(s.pos, s.charges, cell, alpha, None, 1.0, nlow, nhigh)[0]
    assert abs(cs[1]['self_energy'] - e_self) <= 1e-14*abs(e_self)
    # rows as the kernel forms them, gradients summed through the references
    rowg = np.zeros((len(st), 3))
    for r, w in enumerate(st):
        a, b = int(w['a']), int(w['b'])
        if (a < nlow and b < nlow) or (a >= nhigh and b >= nhigh):
            continue
        dl = O.mic(s.pos[b] - s.pos[a], cell)
        d = np.sqrt((dl*dl).sum())
        x = alpha*d
        pot = erf(x)/d
        fac = (1 - w['scale'])*s.charges[b]*s.charges[a]
        rowg[r] = dl*(-fac*(2/np.sqrt(np.pi)*alpha*np.exp(-x*x) - pot)/d/d)
    g = np.zeros((npad, 3))
    for atom in range(npad):
        for ref in p['atom_ref'][p['atom_ref0'][atom]:p['atom_ref0'][atom + 1]]:
            g[atom] += -rowg[ref >> 1] if ref & 1 else rowg[ref >> 1]
    ref = O.ewald_corr(s.pos, s.charges, cell, alpha, st, 1.0, nlow, nhigh)[1]
    assert not g[:256].any() and not g[256 + s.natom:].any()
    assert np.abs(g[256:256 + s.natom] - ref).max() <= 1e-13*np.abs(ref).max()
    # each atom lists its rows in ascending row order
    for atom in range(npad):
        refs = p['atom_ref'][p['atom_ref0'][atom]:p['atom_ref0'][atom + 1]]
        assert list(refs >> 1) == sorted(refs >> 1)


def test_row_refs_order():
    """_lib.row_refs: every (row, side) once, each atom's references in
    ascending row order with b before a on a self-image row."""
    from yaff_amd import _lib
    rng = np.random.default_rng(4)
    n = 300
    rows = np.zeros(n, _lib.ROW_DTYPE)
    rows['a'] = rng.integers(0, 20, n)
    rows['b'] = np.where(rng.uniform(size=n) < 0.1, rows['a'], rng.integers(0, 20, n))
    ref0, ref = _lib.row_refs(rows, 25)
    assert ref0[0] == 0 and ref0[-1] == 2*n and sorted(ref) == list(range(2*n))
    for atom in range(25):
        mine = ref[ref0[atom]:ref0[atom + 1]]
        want = []
        for r in range(n):
            if rows['b'][r] == atom:
                want.append(2*r)
            if rows['a'][r] == atom:
                want.append(2*r + 1)
        assert list(mine) == want


def test_scalings_check_mic_and_bond_paths():
    """Scalings.check_mic (yaff/pes/scaling.py:108-170): the reference's
    MIL-53 unit cell (yaff/test/common.py:308) is too small for its 1-3/1-4
    exclusions (AssertionError, test_scaling.py:102-105), water32 and quartz
    2x2x2 are fine; bond_paths on caffeine gives the paths of
    test_scaling.py:78-99."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden'))
    import cases
    from yaff_amd import Scalings, System
    from yaff_amd.scaling import bond_paths

    def fixture(name):
        f = cases.fixture_arrays(name)
        return System(f['numbers'], f['pos'], bonds=f['bonds'], rvecs=f.get('rvecs'))
    with pytest.raises(AssertionError):
        Scalings(fixture('mil53'))
    Scalings(fixture('mil53'), 1.0, 1.0, 1.0, 1.0)   # nothing scaled: nothing to check
    Scalings(fixture('water32'), 0.0, 0.0, 0.5)
    Scalings(fixture('quartz222'), 0.0, 0.0, 0.0)
    caf = fixture('caffeine')
    assert set(bond_paths(caf, 2, 8, 3)) == {(2, 7, 6, 8), (2, 9, 4, 8)}
    assert set(bond_paths(caf, 13, 5, 5)) == {(13, 4, 8, 6, 7, 5), (13, 4, 9, 2, 7, 5)}
    assert set(bond_paths(caf, 18, 19, 2)) == {(18, 12, 19)}


def test_failure_detection():
    """The reference's guards, on the host before any device work:
    ForcePart.compute refuses NaN energies, gradients and virials with
    ValueError (yaff/pes/ff.py:158-167); NeighborList.update refuses more
    than 10 atoms per bohr^3 (yaff/pes/nlist.py:160-162) and a negative skin
    (nlist.py:93-94); nlow/nhigh are checked (ff.py:1265-1277)."""
    from yaff_amd import System, NeighborList
    from yaff_amd.ff import ForcePart, check_nlow_nhigh
    rng = np.random.default_rng(0)
    s = System(np.ones(100, int), rng.uniform(0, 2.0, (100, 3)), rvecs=np.eye(3)*2.0)

    class Fake(ForcePart):
        def __init__(self, system, what):
            ForcePart.__init__(self, 'fake', system)
            self.what = what

        def _internal_compute(self, gpos, vtens):
            if self.what == 'gpos' and gpos is not None:
                gpos[3, 1] = np.nan
            if self.what == 'vtens' and vtens is not None:
                vtens[2, 2] = np.nan
            return np.nan if self.what == 'energy' else 1.0
    for what in ('energy', 'gpos', 'vtens'):
        with pytest.raises(ValueError):
            Fake(s, what).compute(np.zeros((100, 3)), np.zeros((3, 3)))
    assert Fake(s, 'gpos').compute() == 1.0    # no gradient asked: nothing to check
    nl = NeighborList(s)
    nl.request_rcut(1.0)
    with pytest.raises(ValueError, match='density'):
        nl.update()
    with pytest.raises(ValueError):
        NeighborList(s, skin=-1.0)
    with pytest.raises(ValueError):
        check_nlow_nhigh(s, -1, -1)
    with pytest.raises(ValueError):
        check_nlow_nhigh(s, 50, 10)
    assert check_nlow_nhigh(s, 5, -1) == (5, 100)


def test_ranges_first_fit():
    """The slot engine's range allocator: first fit, aligned, freed ranges
    reused, the high-water mark and the capacity a growth needs."""
    from yaff_amd.slots import Ranges
    r = Ranges(1024, 256)
    a = r.alloc(300)            # 512 units
    b = r.alloc(1)              # 256
    assert (a, b) == (0, 512) and r.high == 768
    assert r.alloc(300) is None            # 512 more do not fit in 1024
    assert r.need(300) == 1280
    r.free(a)
    assert r.high == 768
    c = r.alloc(256)            # first fit: the freed range's start
    d = r.alloc(200)
    assert (c, d) == (0, 256)
    assert r.alloc(256) == 768 and r.high == 1024
    with pytest.raises(KeyError):
        r.free(5)
    r2 = Ranges(10)
    assert [r2.alloc(3), r2.alloc(3), r2.alloc(3), r2.alloc(3)] == [0, 3, 6, None]
    r2.free(3)
    assert r2.alloc(4) is None and r2.alloc(2) == 3 and r2.alloc(1) == 5


def test_system_entry_offsets_and_forms():
    """Per-system planning shared by the batch and slot engines: the pair
    form and its record, LJ sqrt(epsilon), scaling and cross offsets local to
    the system until shifted, and the Ewald self term of a correction
    without scaled pairs."""
    import yaff_amd as ya
    from yaff_amd import _lib, workloads
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import system_entry, shift_offsets, corr_self_energy
    lib = _lib.load()
    lj = nbsystem(workloads.lj_fluid((3, 3, 3), seed=1))
    e = system_entry(lib, lj, 512)
    assert e.h.atom0 == 512 and e.h.natom == lj.natom and e.h.pair_form == _lib.FORM_LJ and e.h.nbin > 0
    np.testing.assert_array_equal(e.rec[:, 1], np.sqrt(lj.pots[0].p1))
    np.testing.assert_array_equal(e.par[0, 2], np.sqrt(lj.pots[0].p1))
    salt = nbsystem(workloads.rocksalt((2, 2, 2), seed=1))
    e = system_entry(lib, salt, 0)
    assert e.h.pair_form == _lib.FORM_IONIC and e.h.pot[0].cross0 == 0 and len(e.cross) == 1
    q = salt.charges
    assert e.h.ew_const == corr_self_energy(salt.ewald.alpha, q, 0, salt.natom)/salt.ewald.dielectric
    np.testing.assert_array_equal(e.rec[:, 2], q/np.sqrt(salt.pots[2].dielectric))
    shift_offsets(e.h, 7, 12)
    assert e.h.pot[0].cross0 == 12 and e.h.pot[0].stab0 == 7 and e.h.pot[1].stab0 == 7
    # a system with scaling rows takes the generic form
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden'))
    import cases
    f = cases.fixture_arrays('water32')
    w = ya.System(f['numbers'], f['pos'], ffatypes=['O', 'H', 'H']*32, bonds=f['bonds'], rvecs=f['rvecs'],
                  charges=f['charges'])
    sc = ya.Scalings(w, 0.0, 0.0, 0.5)
    from yaff_amd.engine import NBSystem, PotSpec
    nb = NBSystem(w.pos, w.cell, [PotSpec(_lib.POT_EI, 9.0, p0=w.charges, alpha=0.2),
                                  PotSpec(_lib.POT_LJ, 9.0, p0=np.full(96, 5.0), p1=np.full(96, 1e-4))],
                  stab=sc.stab)
    e = system_entry(lib, nb, 0)
    assert e.h.pair_form == _lib.FORM_GENERIC and e.h.any_stab == 1
    assert (e.h.pot[0].stab0, e.h.pot[0].nstab, e.h.pot[1].stab0) == (0, len(sc.stab), len(sc.stab))
    assert len(e.stabs) == 2


def test_stager_segments(monkeypatch):
    """The slot engine's staged table updates: copies and fills resolved by
    attribute into one buffer (segment table, then data), split into
    segments of at most 64 KiB, applied here by a CPU stand-in of
    ynb_scatter (copy or fill each segment); overlapping updates refused."""
    import ctypes
    import torch
    from yaff_amd import _lib, slots
    monkeypatch.setattr(slots, '_pinned', lambda n: torch.empty(n, dtype=torch.uint8))

    class Lib(object):
        def ynb_scatter(self, segs, nseg, src, stream):
            tab = np.frombuffer(ctypes.string_at(segs, nseg*_lib.SEGMENT_DTYPE.itemsize), _lib.SEGMENT_DTYPE)
            assert (tab['nbytes'] <= slots.SEG_BYTES).all() and (tab['nbytes'] % tab['elem'] == 0).all()
            for s in tab:
                n, el = int(s['nbytes']), int(s['elem'])
                if s['src'] < 0:
                    ctypes.memmove(int(s['dst']), int(s['fill']).to_bytes(8, 'little')[:el]*(n//el), n)
                else:
                    ctypes.memmove(int(s['dst']), src + int(s['src']), n)
            return 0

    class Eng(object):
        pass
    e = Eng()
    e.lib, e.device, e.stream = Lib(), 'cpu', (lambda: None)
    e._stage_d = torch.zeros(16, dtype=torch.uint8)
    e.a = torch.zeros((100, 3), dtype=torch.float64)
    e.b = torch.full((50,), 7, dtype=torch.int32)
    e.c = torch.zeros(80, dtype=torch.uint8)
    e.d = torch.zeros(50000, dtype=torch.float64)
    st = slots._Stager()
    x = np.random.default_rng(0).normal(size=(10, 3))
    big = np.random.default_rng(1).normal(size=40000)
    st.put('a', 30, x)
    st.fill('a', 60, 9, 1e150)
    st.fill('b', 5, 10, -1)
    st.put('b', 20, np.arange(7, dtype=np.int32))
    st.put('c', 8, np.arange(16, dtype=np.uint8))
    st.put('d', 3, big)
    st.flush(e)
    a, b, c, d = e.a.numpy(), e.b.numpy(), e.c.numpy(), e.d.numpy()
    assert np.array_equal(a[10:20], x) and (a[20:23] == 1e150).all() and not a[:10].any() and not a[23:].any()
    assert (b[5:15] == -1).all() and (b[:5] == 7).all() and (b[15:20] == 7).all() and (b[27:] == 7).all()
    assert (b[20:27] == np.arange(7)).all()
    assert (c[8:24] == np.arange(16)).all() and not c[:8].any() and not c[24:].any()
    assert np.array_equal(d[3:40003], big) and d[0] == 0 and d[40003] == 0
    assert st.ops == []
    st.put('b', 0, np.arange(3, dtype=np.int32))
    st.fill('b', 2, 4, 5)
    with pytest.raises(ValueError):
        st.flush(e)
    s2 = slots._Stager()   # bytes into a byte table go as 8-byte elements
    s2.put('c', 0, np.arange(3, dtype=np.uint8))
    with pytest.raises(ValueError):
        s2.flush(e)
