#!/usr/bin/env python
"""Benchmark of the MI355X nonbonded engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the headline value): the 6912-atom rocksalt system (12x12x6
conventional cells, a = 5.64 A, +-0.05 A jitter, seed 0) with ExpRep + LJ +
real-space EI (rcut 11 A, alpha 0.28/A) and the Ewald reciprocal term
(gcut 0.32/A), in FP64.  One step is one velocity-Verlet MD step at
dt = 1 fs done on the device: kick/drift, neighbour-list rebuild (skin 0),
all pair potentials, Ewald reciprocal, kick.  Inputs are resident in HBM; the
timed region holds K steps between two synchronisations.

value = atom-force-evals/s = natom * K / elapsed.

Also reported on the same line: per-kernel durations from HIP events recorded
on the launch stream over the timed region, the roofline of the dominant
kernel, the mixed-precision rate on the same system with its energy / force /
virial error against the FP64 result, and the CPU baseline (the oracle's FP64
restatement of Yaff's C path on one host core, bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

# Flop-equivalents per unit of work (add/sub/mul 1, FMA 2, div/sqrt/exp/erfc/
# sin/cos 1), read off the reference C expressions.
FLOP_PAIR_TAIL = 57          # pair_pot.c:141-171 (scaling, gpos, vtens, energy)
FLOP_POT = {'lj': 19, 'ei': 18, 'exprep': 5, 'mm3': 22}
FLOP_TRUNC = 8               # chain rule pair_pot.c:134-138
FLOP_SWITCH3 = (1, 10)       # flat zone, switching zone
FLOP_EWALD_SK = 11           # ewald.c:62-76 per (k, atom): k.r 5, q cos/q sin 2, sums 2, sin, cos
FLOP_EWALD_FORCE = 9         # ewald.c:104-111 per (k, atom)


from yaff_amd.batch import maxwell_boltzmann  # noqa: E402


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-mixed', action='store_true')
    ap.add_argument('--cpu-evals', type=int, default=5)
    ap.add_argument('--no-batched', action='store_true')
    ap.add_argument('--batch-steps', type=int, default=200)
    ap.add_argument('--batch-systems', type=int, default=64)
    ap.add_argument('--no-served', action='store_true')
    ap.add_argument('--no-config0', action='store_true', help='skip the 864-atom LJ fluid line')
    ap.add_argument('--no-overlap', action='store_true',
                    help='MD steps on one stream (the Ewald chain after the pair pass): kernel traces whose averages '
                         'are each kernel alone on the chip')
    ap.add_argument('--no-cpu-aggregate', action='store_true',
                    help='skip the all-core oracle runs on the batched and served workloads')
    ap.add_argument('--no-precision', action='store_true',
                    help='skip the 55,296-atom triclinic FP64-versus-mixed comparison')
    ap.add_argument('--serve-requests', type=int, default=150)
    ap.add_argument('--serve-rates', type=str, default='25,100,400')
    ap.add_argument('--tuning', type=str, default=None,
                    help='JSON launch parameters for the single-system engines (skips the autotuner, e.g. '
                         'for profiler runs whose kernel averages must not mix parameter sets)')
    ap.add_argument('--copy-probe-mb', type=int, default=0,
                    help='after the timed region, run an elementwise FP64 copy of this many MiB (a kernel of known '
                         'HBM traffic for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)')
    return ap.parse_args()


def rocksalt_system(reps=(12, 12, 6), seed=0, angles=(90.0, 90.0, 90.0)):
    """A workloads.rocksalt system and its engine input (ExpRep + LJ + EI,
    Ewald reciprocal and correction)."""
    from yaff_amd import workloads
    from yaff_amd.batch import nbsystem
    s = workloads.rocksalt(reps, seed=seed, angles=angles)
    return s, nbsystem(s)


def pair_flops(engine, s_idx=0):
    """Algorithmic flop-equivalents of one pair launch: half pairs within each
    potential's cutoff times its per-pair count."""
    import torch
    eng = engine
    sys_ = eng.systems[s_idx]
    n = sys_.natom
    cnt = eng.cnt_d[:n].cpu().numpy()
    half_pairs = int(cnt.sum())//2
    total = 0
    names = {0: 'lj', 1: 'mm3', 2: 'exprep', 3: 'ei'}
    for p in sys_.pots:
        f = FLOP_POT[names[p.kind]] + FLOP_PAIR_TAIL
        if p.trunc:
            f += FLOP_TRUNC + FLOP_SWITCH3[1]
        total += half_pairs*f
    return total, half_pairs


def run_steps(eng, K):
    import torch
    for _ in range(K):
        eng.md_step()


def timed(eng, K, profile=True):
    """K eager MD steps issued as the graphs of timed_graph hold them (K/S
    sequences of S joined steps); per-kernel HIP events when profile, with
    the Ewald chain run after the pair pass on one stream then, so that each
    kernel's duration is its own (the graphs overlap the two chains)."""
    import torch
    from yaff_amd import _lib
    torch.cuda.synchronize()
    if profile:
        _lib.profile_enable(True)
        eng.overlap = False   # kernel durations measured with each kernel alone on the chip
    t0 = time.perf_counter()
    S = graph_steps(K)
    for _ in range(K//S):
        eng.md_steps(S)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = _lib.profile_read() if profile else []
    if profile:
        _lib.profile_enable(False)
        del eng.overlap   # back to the class setting
    return t1 - t0, prof


def graph_steps(K, most=10):
    """Steps per captured graph: the largest divisor of K up to `most`."""
    return max(d for d in range(1, min(K, most) + 1) if K % d == 0)


def timed_graph(eng, K, W):
    """K MD steps as K/S replays of a HIP graph holding S steps
    (S = graph_steps(K)), after at least W warm-up steps."""
    import torch
    S = graph_steps(K)
    g = eng.capture_md(S)
    for _ in range(max(1, -(-W//S))):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K//S):
        g.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    flags, _ = eng.nlist_status()
    if flags:
        raise RuntimeError('neighbour list overflowed during graph replays')
    return t1 - t0


def kernel_table(prof, K):
    tab = {}
    for name, ms in prof:
        t = tab.setdefault(name, [0, 0.0])
        t[0] += 1
        t[1] += ms
    return {k: {'launches': v[0], 'avg_ms': v[1]/v[0], 'ms_per_step': v[1]/K} for k, v in tab.items()}


def cpu_baseline(s, evals):
    """The oracle (FP64 restatement of Yaff's C path, one host core) on the same
    system: neighbour list, three pair potentials, the Ewald reciprocal and
    correction terms."""
    from oracle import oracle as orc
    O = orc.Oracle()
    sp = s.spec
    n = s.natom
    cell = O.cell(s.cell.rvecs)
    rmax = orc.rmax_of(cell, sp.rcut)
    gmax = orc.gmax_of(cell, sp.gcut)
    pots = [orc.Pot(orc.EXPREP, sp.rcut, ffatype_ids=s.ffatype_ids, amp_cross=np.full((2, 2), sp.exprep_a),
                    b_cross=np.full((2, 2), sp.exprep_b)),
            orc.Pot(orc.LJ, sp.rcut, p0=np.full(n, sp.lj_sigma), p1=np.full(n, sp.lj_epsilon)),
            orc.Pot(orc.EI, sp.rcut, p0=s.charges, alpha=sp.alpha)]
    t0 = time.perf_counter()
    res = None
    for _ in range(evals):
        rows = O.nlist_build(s.pos, cell, sp.rcut, rmax)
        e = 0.0
        g = np.zeros((n, 3))
        v = np.zeros((3, 3))
        for p in pots:
            pe, pg, pv = O.pair_compute(rows, None, p, n)
            e += pe
            g += pg
            v += pv
        pe, pg, pv = O.ewald_reci(s.pos, s.charges, cell, sp.alpha, gmax, sp.gcut)
        ce, cg, cv = O.ewald_corr(s.pos, s.charges, cell, sp.alpha, None)
        res = (e + pe + ce, g + pg + cg, v + pv + cv)
    t = time.perf_counter() - t0
    return n*evals/t, t, res


# bench kernel names -> rocprofv3 kernel names (the FP64 instance's template
# arguments follow from the launch parameters)
ROCPROF_NAMES = {'nlist_bin': 'k_bin', 'nlist_bin_order': 'k_bin_order', 'nlist_build_binned': 'k_build_binned',
                 'pair': 'k_pair_form', 'ewald_prep': 'k_ewald_prep', 'ewald_sk': 'k_ewald_sk',
                 'ewald_kreduce': 'k_ewald_kreduce', 'ewald_force': 'k_ewald_force', 'verlet_kick_drift': 'k_kick_drift',
                 'verlet_kick': 'k_kick', 'verlet_kick_kick_drift': 'k_kick_kick_drift', 'total': 'k_total'}


def fp64_instance(kernel, tuning):
    """rocprofv3 name of the FP64 instance that the launch parameters select
    (None where the name has no template arguments or they are not known)."""
    base = ROCPROF_NAMES.get(kernel)
    if base is None:
        return None
    if kernel == 'pair':
        return '%s<0, 2, %d>' % (base, tuning['pair_lanes'])
    if kernel == 'ewald_sk':
        return '%s<0, %d>' % (base, tuning['apl'])
    if kernel == 'ewald_force':
        return '%s<0, %d>' % (base, tuning['fapl'])
    if kernel == 'nlist_build_binned':
        return '%s<false, 160>' % base
    return base


def profiled_traffic(kernel, tuning=None):
    """HBM bytes per launch (read, written, rocprofv3 name) of `kernel`
    (bench name, e.g. 'ewald_force') from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of the 6912-atom rocksalt step (counters
    scaled by a copy kernel of known size): the instance the launch
    parameters select when the summary has it, else its FP64 instance of
    other launch parameters (the name says which); None when absent."""
    path = os.path.join(ROOT, 'profiles', 'rocksalt6912_hbm_traffic.md')
    base = ROCPROF_NAMES.get(kernel)
    if base is None or not os.path.exists(path):
        return None
    text = open(path).read()
    if 'Machine-readable' not in text:
        return None
    block = text.split('Machine-readable', 1)[1].split('```')[1]
    rows = {}
    for line in block.strip().splitlines():
        name, rd, wr = line.rsplit(' ', 2)
        rows[name] = (float(rd), float(wr), name)
    want = fp64_instance(kernel, tuning) if tuning else None
    if want in rows:
        return rows[want]
    for name, row in rows.items():
        if name.split('<')[0] == base and ('<' not in name or name.split('<')[1][0] in '0f'):
            return row
    return None


def precision_large(steps, warmup):
    """The precision workload: rocksalt 24x24x12 conventional cells (55,296
    atoms), triclinic with angles drawn uniformly in 85-95 degrees (seed 2),
    ExpRep + LJ + EI + Ewald as the headline system.  One FP64 and one mixed
    evaluation (errors of the mixed one: energy, RMS and max force, virial;
    the FP64 GPU evaluation is the reference, itself checked against the
    oracle to 1e-10 by the test suite), then MD steps of each from a graph."""
    import torch
    from yaff_amd.engine import Engine
    from yaff_amd.units import femtosecond
    from yaff_amd import autotune
    rng = np.random.default_rng(2)
    angles = tuple(float(a) for a in rng.uniform(85.0, 95.0, 3))
    s, nb = rocksalt_system(reps=(24, 24, 12), seed=2, angles=angles)
    n = s.natom
    res, out = {}, {'workload': 'rocksalt 24x24x12 cells (%d atoms), triclinic angles %s deg, ExpRep+LJ+EI rcut 11 A '
                                '+ Ewald reci gcut 0.32/A' % (n, ', '.join('%.2f' % a for a in angles)),
                    'natom': n}
    for prec in ('fp64', 'mixed'):
        eng = Engine([nb], precision=prec)
        eng.ensure_nlist()
        tun = autotune.tune(eng, cache=None)
        eng.compute(build=False)
        res[prec] = eng.results(0)
        eng.set_velocities([maxwell_boltzmann(s.masses, 300.0, 1)])
        eng.set_timestep(1.0*femtosecond)
        run_steps(eng, warmup)
        torch.cuda.synchronize()
        flags, _ = eng.nlist_status()
        if flags:
            eng.ensure_nlist()
        dt = timed_graph(eng, steps, warmup)
        out[prec] = {'value': n*steps/dt, 'ms_per_step': dt/steps*1e3, 'steps': steps, 'tuning': tun}
        if prec == 'fp64':
            out['kvectors'] = int((eng.ewald_h['kvec'][:, 3] != 0).sum())
        del eng
        torch.cuda.empty_cache()
    e64, g64, v64 = res['fp64']
    em, gm, vm = res['mixed']
    fn = np.sqrt((g64**2).sum(axis=1))
    out['mixed_error'] = {
        'energy_rel_err': float(abs(em - e64)/abs(e64)),
        'force_rms_rel_err': float(np.sqrt(((gm - g64)**2).sum(axis=1).mean())/np.sqrt((fn**2).mean())),
        'force_max_abs_err': float(np.abs(gm - g64).max()),
        'force_max_rel_err': float(np.sqrt(((gm - g64)**2).sum(axis=1)).max()/fn.max()),
        'virial_rel_err': float(np.abs(vm - v64).max()/np.abs(v64).max()),
        'reference': 'FP64 GPU evaluation of the same positions'}
    # both against the oracle: molmod/yaff's own C kernels on the same system
    # (tests/golden/rocksalt55296_ref.npz: energy, virial and the gradient of
    # 8192 seeded atoms)
    fx = os.path.join(ROOT, 'tests', 'golden', 'rocksalt55296_ref.npz')
    if os.path.exists(fx):
        ref = np.load(fx)
        idx, gref = ref['sample'], ref['gpos_sample']
        gn = np.sqrt((gref**2).sum(axis=1).mean())
        for prec, (e, g, v) in res.items():
            d = g[idx] - gref
            out['vs_oracle_' + prec] = {
                'energy_rel_err': float(abs(e - float(ref['e']))/abs(float(ref['e']))),
                'force_rms_rel_err': float(np.sqrt((d**2).sum(axis=1).mean())/gn),
                'force_max_rel_err': float(np.sqrt((d**2).sum(axis=1)).max()/np.sqrt((gref**2).sum(axis=1)).max()),
                'virial_rel_err': float(np.abs(v - ref['vtens']).max()/np.abs(ref['vtens']).max()),
                'reference': 'molmod/yaff C kernels compiled from the reference sources (fixture '
                             'tests/golden/rocksalt55296_ref.npz; forces on its 8192 sampled atoms)'}
    return out


def config0(steps, warmup, cpu_evals=20):
    """The 864-atom LJ fluid (6x6x6 FCC cells, L = 34.7 A, jitter seed 0),
    FP64: `steps` velocity-Verlet NVE steps at dt = 2 fs from Maxwell-Boltzmann
    velocities at 120 K (seed 1), replayed from a HIP graph; the conserved
    energy's drift over the run; the oracle's FP64 rate on the same system
    (neighbour list and LJ with Switch3, one host core)."""
    import torch
    from yaff_amd import workloads, autotune
    from yaff_amd.batch import nbsystem
    from yaff_amd.engine import Engine
    from yaff_amd.units import femtosecond
    s = workloads.lj_fluid((6, 6, 6), seed=0)
    n = s.natom
    eng = Engine([nbsystem(s)], precision='fp64')
    eng.ensure_nlist()
    tun = autotune.tune(eng, cache=None)
    eng.compute(build=False)
    eng.set_velocities([maxwell_boltzmann(s.masses, 120.0, 1)])
    eng.set_timestep(2.0*femtosecond)
    torch.cuda.synchronize()

    def conserved():
        return float(eng.total_d[0, 0].item()) + 0.5*float(eng.ekin2_d[0].item())
    run_steps(eng, max(warmup, 1))
    torch.cuda.synchronize()
    e0 = conserved()
    dt = timed_graph(eng, steps, 0)
    e1 = conserved()
    ekin = 0.5*float(eng.ekin2_d[0].item())
    out = {'workload': 'LJ fluid 864 atoms (6x6x6 FCC cells, L=34.7 A), sigma 3.405 A, eps 3.79e-4 Ha, rcut 12 A, '
                       'Switch3 4 A, FP64, NVE dt 2 fs, 120 K', 'natom': n, 'steps': steps,
           'value': n*steps/dt, 'unit': 'atom-force-evals/s', 'ms_per_step': dt/steps*1e3, 'tuning': tun,
           'conserved_energy_drift_over_kinetic': abs(e1 - e0)/ekin}
    try:
        from oracle import oracle as orc
        O = orc.Oracle()
        sp = s.spec
        cell = O.cell(s.cell.rvecs)
        pot = orc.Pot(orc.LJ, sp.rcut, orc.TR_SWITCH3, sp.width, p0=np.full(n, sp.sigma), p1=np.full(n, sp.epsilon))
        t0 = time.perf_counter()
        for _ in range(cpu_evals):
            rows = O.nlist_build(s.pos, cell, sp.rcut, orc.rmax_of(cell, sp.rcut))
            O.pair_compute(rows, None, pot, n)
        t = time.perf_counter() - t0
        out['cpu_oracle'] = {'value': n*cpu_evals/t, 'cores': 1, 'kind': 'port',
                             'sample': '%d FP64 evaluations (neighbour list + LJ), %.2f s' % (cpu_evals, t)}
    except OSError as exc:
        out['cpu_oracle'] = {'value': None, 'sample': 'unavailable: %s' % exc}
    del eng
    torch.cuda.empty_cache()
    return out


def _oracle_eval(s):
    """One FP64 oracle evaluation of a workload system (neighbour list, its
    pair potentials, and for rocksalt the Ewald reciprocal and correction):
    (atoms, seconds).  Runs in a worker process."""
    from oracle import oracle as orc
    O = orc.Oracle()
    sp = s.spec
    n = s.natom
    cell = O.cell(s.cell.rvecs)
    t0 = time.perf_counter()
    rows = O.nlist_build(s.pos, cell, sp.rcut, orc.rmax_of(cell, sp.rcut))
    if sp.kind == 'lj':
        O.pair_compute(rows, None, orc.Pot(orc.LJ, sp.rcut, orc.TR_SWITCH3, sp.width, p0=np.full(n, sp.sigma),
                                           p1=np.full(n, sp.epsilon)), n)
    else:
        for p in (orc.Pot(orc.EXPREP, sp.rcut, ffatype_ids=s.ffatype_ids, amp_cross=np.full((2, 2), sp.exprep_a),
                          b_cross=np.full((2, 2), sp.exprep_b)),
                  orc.Pot(orc.LJ, sp.rcut, p0=np.full(n, sp.lj_sigma), p1=np.full(n, sp.lj_epsilon)),
                  orc.Pot(orc.EI, sp.rcut, p0=s.charges, alpha=sp.alpha)):
            O.pair_compute(rows, None, p, n)
        O.ewald_reci(s.pos, s.charges, cell, sp.alpha, orc.gmax_of(cell, sp.gcut), sp.gcut)
        O.ewald_corr(s.pos, s.charges, cell, sp.alpha, None)
    return n, time.perf_counter() - t0


def _oracle_warm(_):
    from oracle import oracle as orc
    orc.Oracle()
    return 0


def cpu_aggregate(systems, label):
    """The oracle on every host core at once: one worker process per core
    (spawned: the workers touch no GPU), each system evaluated once; the
    aggregate rate is the atoms evaluated over the wall time."""
    import multiprocessing as mp
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)          # the host share of one GPU
    ctx = mp.get_context('spawn')
    with ctx.Pool(cores) as pool:
        pool.map(_oracle_warm, range(cores))                # workers started, oracle loaded
        t0 = time.perf_counter()
        res = pool.map(_oracle_eval, systems, chunksize=1)
        wall = time.perf_counter() - t0
    natom = sum(r[0] for r in res)
    return {'value': natom/wall, 'unit': 'atom-force-evals/s', 'cores': cores, 'kind': 'port',
            'sample': '%s: %d systems, %d atoms, one FP64 evaluation each, %.1f s wall (%.1f s of one-core time)'
                      % (label, len(systems), natom, wall, sum(r[1] for r in res)),
            'one_core_value': natom/sum(r[1] for r in res)}


def progress(msg):
    print('bench: %s' % msg, file=sys.stderr, flush=True)


def batched(nsys, steps, warmup):
    """64 independent systems (200-6000 atoms, even index LJ fluid, odd index
    rocksalt, triclinic angles 80-100 degrees, seed 0) in one set of launches:
    autotuned launch parameters, MD steps replayed from a HIP graph, mixed
    precision.  Compared with the same systems evaluated one at a time
    through the FP64 drop-in yaff.pes interface (host arrays in and out, the
    list rebuilt every call), with the FP64 batch, and with the mixed-vs-FP64
    force error of every system."""
    import torch
    from yaff_amd import workloads, autotune
    from yaff_amd.batch import BatchMD, forcefield, nbsystem
    from yaff_amd.engine import Engine
    systems = workloads.batch_of_systems(nsys, (200, 6000), (80.0, 100.0), seed=0)
    natom = sum(s.natom for s in systems)
    out = {'workload': '%d systems, %d atoms (LJ fluid / rocksalt alternating, 200-6000 atoms, angles 80-100 deg)'
                       % (nsys, natom), 'natom': natom, 'steps': steps}

    # FP64 drop-in, one system per ForceField.compute call
    ffs = [forcefield(s) for s in systems]
    for ff, s in zip(ffs, systems):
        ff.compute(np.zeros((s.natom, 3)), np.zeros((3, 3)))
    reps = 3
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for ff, s in zip(ffs, systems):
            ff.update_pos(s.pos)
            ff.compute(np.zeros((s.natom, 3)), np.zeros((3, 3)))
    t_drop = time.perf_counter() - t0
    out['dropin_fp64'] = {'value': natom*reps/t_drop, 'ms_per_pass': t_drop/reps*1e3,
                          'note': 'ForceField.compute(gpos, vtens) of each system in turn, FP64, host buffers'}
    del ffs

    # mixed vs FP64 forces of the same systems, batched
    nbs = [nbsystem(s) for s in systems]
    res = {}
    for prec in ('fp64', 'mixed'):
        e = Engine(nbs, precision=prec)
        e.ensure_nlist()
        e.compute(build=False)
        res[prec] = e.results()
        del e
    rms, emax, vmax = [], [], []
    for (e64, g64, v64), (em, gm, vm) in zip(res['fp64'], res['mixed']):
        rms.append(np.sqrt(((gm - g64)**2).sum(axis=1).mean()/(g64**2).sum(axis=1).mean()))
        emax.append(abs(em - e64)/abs(e64))
        vmax.append(np.abs(vm - v64).max()/np.abs(v64).max())
    out['mixed_error'] = {'force_rms_rel_err_max': float(max(rms)), 'energy_rel_err_max': float(max(emax)),
                          'virial_rel_err_max': float(max(vmax))}

    for prec in ('mixed', 'fp64'):
        md = BatchMD(systems, precision=prec)
        t0 = time.perf_counter()
        tun = autotune.tune(md.engine, cache=None)
        t_tune = time.perf_counter() - t0
        md.warm(warmup)
        nst = steps if prec == 'mixed' else max(steps//4, 1)
        md.capture(graph_steps(nst))
        dt = md.run(nst)
        # per-kernel durations of a few eager steps of the same batch (one
        # stream: each kernel alone on the chip)
        from yaff_amd import _lib
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        md.engine.overlap = False
        md.engine.md_steps(3)
        kt = kernel_table(_lib.profile_read(), 3)
        del md.engine.overlap
        _lib.profile_enable(False)
        progress('batched %s %.3g atom-evals/s' % (prec, natom*nst/dt))
        out['batched_' + prec] = {'value': natom*nst/dt, 'ms_per_step': dt/nst*1e3, 'steps': nst,
                                  'tuning': tun, 'tune_s': t_tune,
                                  'kernels_ms_per_step': {k: v['ms_per_step'] for k, v in kt.items()}}
        del md
        torch.cuda.empty_cache()
    # the same MD with a Verlet skin of 1 A: lists rebuilt (on the device) only after a displacement of 0.5 A
    from yaff_amd.units import angstrom
    md = BatchMD(systems, precision='mixed', skin=1.0*angstrom)
    md.engine.set_tuning(**out['batched_mixed']['tuning'])
    md.warm(warmup)
    md.capture(graph_steps(steps))
    dt = md.run(steps)
    from yaff_amd import _lib
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    md.engine.overlap = False
    md.engine.md_steps(10)
    kt = kernel_table(_lib.profile_read(), 10)
    del md.engine.overlap
    _lib.profile_enable(False)
    out['batched_mixed_skin1A'] = {'value': natom*steps/dt, 'ms_per_step': dt/steps*1e3, 'steps': steps,
                                   'skin_angstrom': 1.0,
                                   'kernels_ms_per_step': {k: v['ms_per_step'] for k, v in kt.items()},
                                   'note': 'same systems and tuning; Verlet skin, per-system rebuild decided on '
                                           'the device inside the graph'}
    progress('batched mixed skin 1A %.3g atom-evals/s' % (natom*steps/dt))
    del md
    torch.cuda.empty_cache()
    out['mixed_batched_vs_dropin_fp64'] = out['batched_mixed']['value']/out['dropin_fp64']['value']
    out['mixed_batched_skin_vs_dropin_fp64'] = out['batched_mixed_skin1A']['value']/out['dropin_fp64']['value']
    out['mixed_batched_vs_fp64_batched'] = out['batched_mixed']['value']/out['batched_fp64']['value']
    return out


def served(nreq, rates):
    """Poisson request streams (single evaluations and 20-step MD segments,
    half and half; systems log-uniform in 100-50000 atoms, LJ fluid /
    rocksalt alternating, triclinic angles 80-100 degrees) served by the
    continuous-batching server, FP64 and mixed, at each arrival rate, and at
    saturation (every request waiting at the start): p50/p99 latency (arrival
    to results on the host), throughput, the host's share of the span spent
    placing requests (admission, retirement and table uploads, overlapped
    with the quantum in flight) and the device's busy share (HIP-event time
    of the quanta)."""
    from yaff_amd import serve
    out = {'workload': '%d requests per rate, log-uniform 100-50000 atoms, 50%% eval / 50%% 20-step MD' % nreq}
    warm = serve.poisson_requests(6, 1000.0, natom_range=(100, 3000), nsteps=2, seed=99)
    serve.Server(precision='mixed').run(warm, realtime=False)
    for prec in ('mixed', 'fp64'):
        res = []
        srv = serve.Server(precision=prec)   # one persistent engine for every rate
        for rate in rates + [None]:
            reqs = serve.poisson_requests(nreq, 1000.0 if rate is None else rate, seed=11)
            srv.stats.clear()
            done = srv.run(reqs, realtime=rate is not None)
            summ = serve.latency_summary(done)
            summ.update({'rate_per_s': 'saturation' if rate is None else rate, 'commits': srv.stats['forms'],
                         'iterations': srv.stats['iterations'], 'form_s': srv.stats['form_s'],
                         'form_frac': srv.stats['form_s']/summ['span_s'], 'wait_s': srv.stats['wait_s'],
                         'device_busy_frac': srv.stats['device_s']/summ['span_s'],
                         'regrow': srv.stats['regrow'], 'peak_est_bytes': srv.peak_bytes,
                         'engine_bytes': srv.engine.device_bytes()})
            res.append(summ)
            progress('served %s rate %s: p50 %.1f ms p99 %.1f ms, %.3g atom-evals/s, host placing %.1f%% and '
                     'device busy %.1f%% of the span' % (prec, summ['rate_per_s'], summ['p50_ms'], summ['p99_ms'],
                                                        summ['throughput'], 100*summ['form_frac'],
                                                        100*summ['device_busy_frac']))
        out[prec] = res
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures the throughput of one MI355X; no multi-GPU path is built (a batch or a '
              'large system would shard by systems or by atom blocks, see the design notes); run with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)
    import torch
    from yaff_amd.engine import Engine
    from yaff_amd.units import femtosecond
    K, W = args.steps, args.warmup
    if args.no_overlap:
        Engine.overlap = False

    s, nb = rocksalt_system()
    n = s.natom
    from yaff_amd import autotune
    fixed = json.loads(args.tuning) if args.tuning else None

    def tune(e):
        if fixed:
            e.set_tuning(**fixed)
            return e.tuning()
        return autotune.tune(e, cache=None)

    eng = Engine([nb], precision='fp64')
    eng.ensure_nlist()
    tuning = tune(eng)
    eng.compute(build=False)
    eng.set_velocities([maxwell_boltzmann(s.masses, 300.0, 1)])
    eng.set_timestep(1.0*femtosecond)
    run_steps(eng, W)
    torch.cuda.synchronize()
    flags, maxcnt = eng.nlist_status()
    if flags:
        eng.ensure_nlist()
    dt_eager, prof = timed(eng, K)
    flags, _ = eng.nlist_status()
    if flags:
        raise RuntimeError('neighbour list overflowed during the timed steps')
    kt = kernel_table(prof, K)
    dt = timed_graph(eng, K, W)
    value = n*K/dt

    # roofline of the dominant kernel
    nk_real = int((eng.ewald_h['kvec'][:, 3] != 0).sum())
    pflops, half_pairs = pair_flops(eng)
    algo = {'ewald_sk': nk_real*n*FLOP_EWALD_SK, 'ewald_force': nk_real*n*FLOP_EWALD_FORCE, 'pair': pflops}
    dom = max(kt, key=lambda k: kt[k]['ms_per_step'])
    for k, v in kt.items():
        if k in algo:
            v['achieved_tflops'] = algo[k]/(v['avg_ms']*1e-3)/1e12
            v['frac_fp64_peak'] = v['achieved_tflops']/FP64_PEAK_TFLOPS
        # achieved HBM rate: the profiled bytes per launch over this run's
        # average launch duration
        tr = profiled_traffic(k, tuning)
        if tr:
            v['hbm_bytes_per_launch'] = tr[0] + tr[1]
            v['hbm_gbs'] = (tr[0] + tr[1])/(v['avg_ms']*1e-3)/1e9
            v['frac_hbm_peak'] = v['hbm_gbs']/HBM_PEAK_GBS
            v['hbm_profiled_as'] = tr[2]
    if dom in algo:
        ach = algo[dom]/(kt[dom]['avg_ms']*1e-3)/1e12
        tr = profiled_traffic(dom, tuning)
        roofline = {'kernel': dom, 'bound': 'valu', 'achieved': ach, 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                    'frac': ach/FP64_PEAK_TFLOPS, 'traffic': (tr[0] + tr[1]) if tr else None,
                    'traffic_source': ('profiles/rocksalt6912_hbm_traffic.md, %s (rocprofv3 FETCH_SIZE + '
                                       'WRITE_SIZE per launch, scaled by a 512 MiB copy kernel)' % tr[2])
                    if tr else None,
                    'algorithmic_flop_per_launch': algo[dom]}
    else:
        roofline = {'kernel': dom, 'bound': 'hbm', 'achieved': None, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                    'frac': None, 'traffic': None}

    out = {'metric': 'atom-force-evals/s (FP64 vs mixed @ RMS force err<1e-5) + p99 served latency',
           'value': value, 'unit': 'atom-force-evals/s', 'n_gpus': 1, 'steps': K, 'warmup': W,
           'ms_per_step': dt/K*1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f64', 'data': 'synthetic',
           'config': {'workload': 'rocksalt 6912 atoms (12x12x6 cells), ExpRep+LJ+EI rcut 11 A + Ewald reci '
                                  'gcut 0.32/A, FP64, velocity-Verlet MD step with nlist rebuild',
                      'natom': n, 'half_pairs': half_pairs, 'kvectors': nk_real, 'parallelism': 'single GPU',
                      'tuning': tuning},
           'roofline': roofline, 'kernels': kt,
           'timing': 'value: K MD steps as K/S replays of a HIP graph holding S = %d steps; kernels: HIP events '
                     'around each launch of the same K steps issued eagerly on the launch stream' % graph_steps(K),
           'eager': {'value': n*K/dt_eager, 'ms_per_step': dt_eager/K*1e3}}

    # mixed precision on the same system: rate and error against FP64
    if not args.no_mixed:
        eng.upload_positions()
        eng.ensure_nlist()
        eng.compute(build=False)
        e64, g64, v64 = eng.results(0)
        engm = Engine([nb], precision='mixed')
        engm.ensure_nlist()
        tuning_m = tune(engm)
        engm.compute(build=False)
        em, gm, vm = engm.results(0)
        engm.set_velocities([maxwell_boltzmann(s.masses, 300.0, 1)])
        engm.set_timestep(1.0*femtosecond)
        run_steps(engm, W)
        dtm_eager, profm = timed(engm, K)
        dtm = timed_graph(engm, K, W)
        rms = np.sqrt(((gm - g64)**2).sum(axis=1).mean())/np.sqrt((g64**2).sum(axis=1).mean())
        out['mixed'] = {'value': n*K/dtm, 'ms_per_step': dtm/K*1e3,
                        'energy_rel_err': abs(em - e64)/abs(e64),
                        'force_rms_rel_err': float(rms),
                        'force_max_abs_err': float(np.abs(gm - g64).max()),
                        'virial_rel_err': float(np.abs(vm - v64).max()/np.abs(v64).max()),
                        'tuning': tuning_m,
                        'kernels': kernel_table(profm, K)}
        # mixed kernels keep most arithmetic in FP64 (FP32 for exp/erfc and
        # the phase seeds); without packed FP32 the FP32 vector peak equals
        # the FP64 one, 78.6 TFLOP/s
        for k, v in out['mixed']['kernels'].items():
            if k in algo:
                v['achieved_tflops'] = algo[k]/(v['avg_ms']*1e-3)/1e12
                v['frac_vector_peak'] = v['achieved_tflops']/FP64_PEAK_TFLOPS

    if args.copy_probe_mb:
        src = torch.ones(args.copy_probe_mb*(1 << 20)//8, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        out['copy_probe'] = {'bytes_read': src.numel()*8, 'bytes_written': dst.numel()*8,
                             'kernel': 'torch elementwise mul (out=)'}
        del src, dst
    progress('single-system FP64/mixed done')
    if not args.no_config0:
        out['config0'] = config0(100, W)
        progress('config 0 (864-atom LJ, 100 NVE steps) done: %.3g atom-evals/s' % out['config0']['value'])
    if not args.no_precision:
        out['precision'] = precision_large(max(K//2, 1), 2)
        progress('precision (55,296 atoms) done: mixed RMS force error %.2e'
                 % out['precision']['mixed_error']['force_rms_rel_err'])
    if not args.no_batched:
        out['batched'] = batched(args.batch_systems, args.batch_steps, 2)
        progress('batched done')

    if not args.no_served:
        out['served'] = served(args.serve_requests, [float(x) for x in args.serve_rates.split(',')])
        progress('served done')

    if not args.no_cpu_baseline:
        try:
            rate, t, res = cpu_baseline(s, args.cpu_evals)
            out['cpu_baseline'] = {'value': rate, 'unit': 'atom-force-evals/s', 'cores': 1, 'kind': 'port',
                                   'sample': '%d full FP64 evaluations (nlist + ExpRep/LJ/EI + Ewald reci) of the '
                                             '6912-atom system, %.1f s' % (args.cpu_evals, t)}
            eng.upload_positions()
            eng.ensure_nlist()
            eng.compute(build=False)
            e, g, v = eng.results(0)
            out['cpu_baseline']['gpu_vs_oracle_energy_rel_err'] = abs(e - res[0])/abs(res[0])
            out['cpu_baseline']['gpu_vs_oracle_force_max_rel_err'] = float(np.abs(g - res[1]).max()/np.abs(res[1]).max())
        except OSError as exc:
            out['cpu_baseline'] = {'value': None, 'unit': 'atom-force-evals/s', 'cores': 1, 'kind': 'port',
                                   'sample': 'unavailable: %s' % exc}
        if not args.no_cpu_aggregate and out['cpu_baseline']['value'] is not None:
            # all host cores on the batched workload and on a bounded sample of the served stream
            from yaff_amd import workloads, serve
            agg = {'batched': cpu_aggregate(workloads.batch_of_systems(args.batch_systems, (200, 6000),
                                                                      (80.0, 100.0), seed=0),
                                            'the batched workload (200-6000 atoms)')}
            stream = [r.system for r in serve.poisson_requests(args.serve_requests, 400.0, seed=11)
                      if r.natom <= 6000][:64]
            agg['served'] = cpu_aggregate(stream, 'served-stream requests of at most 6000 atoms')
            out['cpu_baseline']['all_cores'] = agg
            progress('cpu aggregate: batched %.3g, served sample %.3g atom-evals/s on %d cores'
                     % (agg['batched']['value'], agg['served']['value'], agg['batched']['cores']))
    print(json.dumps(out))


if __name__ == '__main__':
    main()
