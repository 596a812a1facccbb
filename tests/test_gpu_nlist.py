"""Neighbour-list build paths beyond the defaults.

* nlow/nhigh groups (nlist.c:57: a pair is kept when its larger index is
  at least nlow and its smaller one below nhigh): rows bit-identical to the
  oracle's on the 864-atom LJ fluid and a 144-atom triclinic rocksalt.
* A row capacity far below the longest row: the build flags the overflow
  (status bit 1), counts every row in full, writes no entry past a row's
  capacity (each row holds exactly the first kcap entries of the row that
  a large capacity gives, so the next row was not touched), and after
  ensure_nlist grows the capacity the rows equal the oracle's.
"""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]


def _system(name):
    from yaff_amd import workloads
    if name == 'lj864':
        return workloads.lj_fluid()
    return workloads.rocksalt((3, 3, 2), seed=4, angles=(87.0, 92.0, 94.0))


def _engine(s, kcap=None, nlow=0, nhigh=-1):
    from yaff_amd import _lib
    from yaff_amd.engine import Engine, NBSystem, PotSpec
    n = s.natom
    rc = s.spec.rcut
    pot = PotSpec(_lib.POT_LJ, rc, p0=np.full(n, 3.0*1.889726133921252), p1=np.full(n, 1e-4))
    return Engine([NBSystem(s.pos, s.cell, [pot], rc=rc, nlow=nlow, nhigh=nhigh)], kcap=kcap)


def _oracle_rows(s, nlow=0, nhigh=-1):
    from oracle import oracle as orc
    O = orc.Oracle()
    cell = O.cell(s.cell.rvecs)
    return O.nlist_build(s.pos, cell, s.spec.rcut, orc.rmax_of(cell, s.spec.rcut), nlow, nhigh)


@pytest.mark.parametrize('name,nlow,nhigh', [('lj864', 300, 700), ('lj864', 0, 500), ('lj864', 400, -1),
                                             ('rocksalt144', 50, 120)])
def test_groups_rows_bit_exact(name, nlow, nhigh):
    print('case %s nlow %d nhigh %d' % (name, nlow, nhigh), flush=True)
    s = _system(name)
    eng = _engine(s, nlow=nlow, nhigh=nhigh)
    eng.ensure_nlist()
    got = eng.rows(0)
    ref = _oracle_rows(s, nlow, nhigh)
    assert len(ref) > 0 and len(got) == len(ref) and got.tobytes() == ref.tobytes()


def test_row_capacity_overflow():
    print('case lj864 kcap 40', flush=True)
    s = _system('lj864')
    n = s.natom
    big = _engine(s)
    big.ensure_nlist()
    kb = big.kcap
    cnt_big = big.cnt_d[:n].cpu().numpy()
    j_big = big.nbr_j_d.view(-1, kb)[:n].cpu().numpy()
    i_big = big.nbr_img_d.view(-1, kb)[:n].cpu().numpy()
    K = 40
    assert cnt_big.min() > K
    small = _engine(s, kcap=K)
    small.build_nlist()
    flags, longest = small.nlist_status()
    assert flags & 1 and longest == cnt_big.max()
    assert np.array_equal(small.cnt_d[:n].cpu().numpy(), cnt_big)
    assert np.array_equal(small.nbr_j_d.view(-1, K)[:n].cpu().numpy(), j_big[:, :K])
    assert np.array_equal(small.nbr_img_d.view(-1, K)[:n].cpu().numpy(), i_big[:, :K])
    small.ensure_nlist()
    assert small.kcap >= cnt_big.max()
    ref = _oracle_rows(s)
    assert small.rows(0).tobytes() == ref.tobytes()
