"""The C ABI without a GPU: libyaff_nb.so loads, exports every function
include/yaff_nb.h declares, and the ctypes structures of yaff_amd._lib have
the header's layout (sizes and field offsets read from a C probe compiled
against the header with gcc).  Host-only entry points (version, the cell-bin
planner, workspace sizes) are called; nothing touches a device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'yaff_nb.h')


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'^[A-Za-z_][\w \*]*?\b(ynb_\w+)\s*\(', text, flags=re.M)))


@pytest.fixture(scope='module')
def lib():
    from yaff_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from yaff_amd import build
        build.build()
    return _lib.load()


def test_exports_every_declared_function(lib):
    from yaff_amd import _lib
    names = _declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == names


STRUCTS = {
    'ynb_pot': 'YnbPot', 'ynb_sys': 'YnbSys', 'ynb_batch': 'YnbBatch', 'ynb_nlist': 'YnbNlist',
    'ynb_pairpar': 'YnbPairpar', 'ynb_ewald': 'YnbEwald', 'ynb_corr_sys': 'YnbCorrSys', 'ynb_corr': 'YnbCorr',
}


def _c_layout(tmp_path):
    from yaff_amd import _lib
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "yaff_nb.h"', 'int main(void) {']
    for cname, pyname in STRUCTS.items():
        st = getattr(_lib, pyname)
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in st._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ['printf("ynb_neigh_row sizeof %zu\\n", sizeof(ynb_neigh_row));',
              'printf("ynb_scaling_row sizeof %zu\\n", sizeof(ynb_scaling_row));', 'return 0; }']
    src = tmp_path / 'probe.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'probe'
    subprocess.check_call(['gcc', '-std=c99', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)])
    out = {}
    for line in subprocess.check_output([str(exe)]).decode().splitlines():
        a, b, c = line.split()
        out[(a, b)] = int(c)
    return out


def test_struct_layout_matches_header(tmp_path):
    from yaff_amd import _lib
    lay = _c_layout(tmp_path)
    for cname, pyname in STRUCTS.items():
        st = getattr(_lib, pyname)
        assert C.sizeof(st) == lay[(cname, 'sizeof')], cname
        for f in st._fields_:
            assert getattr(st, f[0]).offset == lay[(cname, f[0])], (cname, f[0])
    assert _lib.ROW_DTYPE.itemsize == lay[('ynb_neigh_row', 'sizeof')] == 72
    assert _lib.SCALING_DTYPE.itemsize == lay[('ynb_scaling_row', 'sizeof')] == 32


def test_host_entry_points(lib):
    from yaff_amd import _lib, workloads
    assert lib.ynb_version().decode().startswith('yaff_nb')
    s = workloads.rocksalt((3, 3, 2), angles=(85.0, 92.0, 97.0))
    h = _lib.YnbSys()
    cell = s.cell
    for k in range(9):
        h.rvecs[k] = cell.rvecs_full.ravel()[k]
        h.gvecs[k] = cell.gvecs_full.ravel()[k]
    h.rc = s.spec.rcut
    h.volume = cell.volume
    h.nvec = 3
    h.natom = s.natom
    h.npot = 1
    rmax = np.ceil(h.rc/cell.rspacings - 0.5).astype(int)
    for k in range(3):
        h.rmax[k] = int(rmax[k])
    h.nhigh = s.natom
    _lib.check(lib.ynb_sys_plan(C.byref(h)))
    grid = np.array(h.grid[:])
    assert h.nbin == grid.prod() and (grid >= 1).all()
    # every bin is at least rc/range wide along each reciprocal direction
    width = cell.rspacings/grid
    assert (width*np.array(h.range[:]) >= h.rc*(1 - 1e-12)).all()
    # invalid arguments come back as an error code with a message, not a crash
    h.nvec = 7
    assert lib.ynb_sys_plan(C.byref(h)) != 0
    assert lib.ynb_last_error()
    assert lib.ynb_nlist_workspace_bytes(64*10, 8, 32) > 0
    assert lib.ynb_rows_workspace_bytes(100, 1000) >= lib.ynb_rows_workspace_bytes(100, 0)
