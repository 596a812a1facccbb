"""ynb_pair_rows_static, the entry point that allocates nothing, against
ynb_pair_rows (the same kernels on caller-given work buffers), through the
C ABI: two calls in a row on one device, the second with a smaller input,
each bitwise equal to the allocating entry point; sizes beyond the static
array are refused before any launch."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import gpu_available

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'golden'))

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason='needs a HIP device')]

ANG = 1.889726133921252


def _rows_and_params(natom_cells):
    """Yaff rows of an LJ fluid with Switch3 (oracle rows), per-atom sigma,
    epsilon with a spread."""
    from oracle import oracle as orc
    from yaff_amd import workloads
    s = workloads.lj_fluid(natom_cells, seed=5)
    O = orc.Oracle()
    cell = O.cell(s.cell.rvecs)
    rc = s.spec.rcut
    rows = O.nlist_build(s.pos, cell, rc, orc.rmax_of(cell, rc))
    rng = np.random.default_rng(8)
    sig = rng.uniform(3.2, 3.6, s.natom)*ANG
    eps = rng.uniform(2e-4, 5e-4, s.natom)
    return s, rows, sig, eps


def _call(static, rows, sig, eps, natom, rc, width):
    import torch
    from yaff_amd import _lib
    from yaff_amd.engine import Engine
    L = _lib.load()
    dev = torch.device('cuda')

    def up(a, dt=np.float64):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=dt).copy()).to(dev)
    rows_d = torch.from_numpy(np.frombuffer(rows.tobytes(), np.uint8).copy()).to(dev)
    stab_d = torch.zeros(32, dtype=torch.uint8, device=dev)
    p0, p1 = up(sig), up(eps)
    pot = _lib.YnbPot(_lib.POT_LJ, _lib.TR_SWITCH3, rc, width, 0.0, 1.0, 0, 0, 0, 0)
    ref0, ref = _lib.row_refs(rows, natom)
    ref0_d, ref_d = up(ref0, np.int32), up(ref, np.int32)
    g = torch.zeros((natom, 3), dtype=torch.float64, device=dev)
    out = torch.zeros(10, dtype=torch.float64, device=dev)
    if static:
        _lib.check(L.ynb_pair_rows_static(rows_d.data_ptr(), len(rows), stab_d.data_ptr(), 0, pot, p0.data_ptr(),
                                          p1.data_ptr(), None, None, 1, natom, ref0_d.data_ptr(), ref_d.data_ptr(),
                                          g.data_ptr(), out.data_ptr(), Engine.stream()))
    else:
        rowh = torch.empty(3*len(rows), dtype=torch.float64, device=dev)
        partial = torch.empty(8*((len(rows) + 255)//256), dtype=torch.float64, device=dev)
        _lib.check(L.ynb_pair_rows(rows_d.data_ptr(), len(rows), stab_d.data_ptr(), 0, pot, p0.data_ptr(),
                                   p1.data_ptr(), None, None, 1, natom, ref0_d.data_ptr(), ref_d.data_ptr(),
                                   g.data_ptr(), rowh.data_ptr(), partial.data_ptr(), out.data_ptr(),
                                   Engine.stream()))
    torch.cuda.synchronize()
    return g.cpu().numpy(), out.cpu().numpy()


def test_static_entry_matches_allocating_entry():
    from oracle import oracle as orc
    from yaff_amd import _lib
    for cells, nkeep in (((4, 4, 4), None), ((3, 3, 2), 5000)):
        print('case no-alloc pair rows, lj cells %s, rows kept %s' % (cells, nkeep), flush=True)
        s, rows, sig, eps = _rows_and_params(cells)
        if nkeep is not None:
            rows = rows[:nkeep].copy()
        assert 0 < len(rows) <= _lib.PAIR_ROWS_STATIC_MAX_ROWS
        rc, width = s.spec.rcut, s.spec.width
        g0, o0 = _call(False, rows, sig, eps, s.natom, rc, width)
        g1, o1 = _call(True, rows, sig, eps, s.natom, rc, width)
        assert np.array_equal(g0, g1) and np.array_equal(o0, o1)
        # and the oracle (pair_pot_compute restated) on the same rows
        e, g, v = orc.Oracle().pair_compute(rows, None, orc.Pot(orc.LJ, rc, orc.TR_SWITCH3, width, p0=sig, p1=eps),
                                            s.natom)
        assert abs(o1[0] - e) <= 1e-10*abs(e)
        assert np.abs(g1 - g).max() <= 1e-10*np.abs(g).max()
        assert np.abs(o1[1:].reshape(3, 3) - v).max() <= 1e-10*np.abs(v).max()


def test_static_entry_refuses_oversize():
    from yaff_amd import _lib
    L = _lib.load()
    pot = _lib.YnbPot(_lib.POT_LJ, 0, 10.0, 0.0, 0.0, 1.0, 0, 0, 0, 0)
    dummy = C.c_void_p(16)   # never dereferenced: the size check comes first
    rc = L.ynb_pair_rows_static(dummy, _lib.PAIR_ROWS_STATIC_MAX_ROWS + 1, None, 0, pot, dummy, dummy, None, None,
                                0, 10, None, None, None, dummy, None)
    assert rc == -4 and b'static work array' in L.ynb_last_error()
