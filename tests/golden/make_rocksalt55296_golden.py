"""Writes tests/golden/rocksalt55296_ref.npz: molmod/yaff's own C kernels
(oracle/_ref, built by oracle/Makefile from the reference sources) on the
precision workload -- rocksalt 24x24x12 conventional cells (55,296 atoms),
triclinic with angles drawn uniformly in 85-95 degrees from seed 2, jitter
seed 2; ExpRep + LJ + EI real space (rcut 11 A), Ewald reciprocal (gcut
0.32/A, alpha 0.28/A) and its correction (the self term).  Run where the
reference tree and oracle/_ref exist (minutes of CPU):

    make -C oracle && python tests/golden/make_rocksalt55296_golden.py

Stored (small enough to commit): the row count and the SHA-256 of the
72-byte rows in nlist_build_low's order; the total energy and virial and
those of each part; the gradient of 8192 seeded atoms; and 8 projections
sum_i w_k,i . g_i of the full gradient on seeded random weights, which
compare the whole gradient to rounding without storing it.
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

NSAMPLE = 8192
NPROJ = 8


def workload():
    from yaff_amd import workloads
    rng = np.random.default_rng(2)
    angles = tuple(float(a) for a in rng.uniform(85.0, 95.0, 3))
    return workloads.rocksalt((24, 24, 12), seed=2, angles=angles), angles


def sample_and_weights(natom):
    rng = np.random.default_rng(55296)
    idx = np.sort(rng.choice(natom, NSAMPLE, replace=False))
    w = rng.normal(0, 1, (NPROJ, natom, 3))
    return idx, w


def main():
    from oracle import oracle as orc
    R = orc.Oracle(reference=True)
    s, angles = workload()
    sp = s.spec
    n = s.natom
    cell = R.cell(s.cell.rvecs)
    t0 = time.time()
    rows = R.nlist_build(s.pos, cell, sp.rcut, orc.rmax_of(cell, sp.rcut), chunk=1 << 22)
    print('rows', len(rows), '%.0f s' % (time.time() - t0), flush=True)
    out = {'natom': np.array(n), 'angles': np.array(angles), 'nrows': np.array(len(rows)),
           'rows_sha256': np.array(hashlib.sha256(rows.tobytes()).hexdigest())}
    pots = [('exprep', orc.Pot(orc.EXPREP, sp.rcut, ffatype_ids=s.ffatype_ids,
                                amp_cross=np.full((2, 2), sp.exprep_a), b_cross=np.full((2, 2), sp.exprep_b))),
            ('lj', orc.Pot(orc.LJ, sp.rcut, p0=np.full(n, sp.lj_sigma), p1=np.full(n, sp.lj_epsilon))),
            ('ei', orc.Pot(orc.EI, sp.rcut, p0=s.charges, alpha=sp.alpha))]
    parts = []
    for name, p in pots:
        parts.append((name, R.pair_compute(rows, None, p, n)))
    del rows
    print('pairs %.0f s' % (time.time() - t0), flush=True)
    parts.append(('reci', R.ewald_reci(s.pos, s.charges, cell, sp.alpha, orc.gmax_of(cell, sp.gcut), sp.gcut)))
    print('reci %.0f s' % (time.time() - t0), flush=True)
    parts.append(('corr', R.ewald_corr(s.pos, s.charges, cell, sp.alpha, None)))
    e = sum(p[1][0] for p in parts)
    g = sum(p[1][1] for p in parts)
    v = sum(p[1][2] for p in parts)
    for name, (pe, pg, pv) in parts:
        out[name + '/e'] = np.array(pe)
        out[name + '/vtens'] = pv
    idx, w = sample_and_weights(n)
    out.update({'e': np.array(e), 'vtens': v, 'sample': idx, 'gpos_sample': g[idx],
                'gpos_proj': np.einsum('kij,ij->k', w, g), 'gpos_sumsq': np.array((g*g).sum())})
    np.savez_compressed(os.path.join(HERE, 'rocksalt55296_ref.npz'), **out)
    print('written, %.0f s' % (time.time() - t0))


if __name__ == '__main__':
    main()
