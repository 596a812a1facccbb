"""ctypes binding of libyaff_nb.so (include/yaff_nb.h).

The structures below mirror the header field by field.  Loading fails loudly
when the library has not been built: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libyaff_nb.so')

MAX_POT = 4
POT_LJ, POT_MM3, POT_EXPREP, POT_EI = 0, 1, 2, 3
TR_NONE, TR_SWITCH3, TR_HAMMER = 0, 1, 2
FP64, MIXED = 0, 1
BATCH_ALLPAIRS, BATCH_EXACT_BUILD = 1, 2
BATCH_FORM_GENERIC, BATCH_FORM_LJ, BATCH_FORM_IONIC = 4, 8, 16
BATCH_WIDE_SEARCH = 32
BUILD_NB_SMALL = 160   # YNB_BUILD_NB_SMALL
NLIST_PREBINNED = 1
EWALD_PREPARED = 1
FORM_GENERIC, FORM_LJ, FORM_IONIC = 0, 1, 2
PAIR_ROWS_STATIC_MAX_ROWS = 262144

ROW_DTYPE = np.dtype([('a', np.int64), ('b', np.int64), ('d', np.float64), ('dx', np.float64),
                      ('dy', np.float64), ('dz', np.float64), ('r0', np.int64), ('r1', np.int64),
                      ('r2', np.int64)])
SCALING_DTYPE = np.dtype([('a', np.int64), ('b', np.int64), ('scale', np.float64), ('nbond', np.int64)])
# ynb_segment (include/yaff_nb.h)
SEGMENT_DTYPE = np.dtype([('dst', np.uint64), ('src', np.int64), ('nbytes', np.int64), ('fill', np.uint64),
                          ('elem', np.int32), ('pad', np.int32)])


class YnbPot(C.Structure):
    _fields_ = [('kind', C.c_int32), ('trunc', C.c_int32), ('rcut', C.c_double), ('trunc_par', C.c_double),
                ('alpha', C.c_double), ('dielectric', C.c_double), ('nffatype', C.c_int32),
                ('cross0', C.c_int32), ('stab0', C.c_int32), ('nstab', C.c_int32)]


class YnbSys(C.Structure):
    _fields_ = [('rvecs', C.c_double * 9), ('gvecs', C.c_double * 9), ('rc', C.c_double),
                ('volume', C.c_double), ('nvec', C.c_int32), ('natom', C.c_int32), ('atom0', C.c_int32),
                ('npot', C.c_int32), ('rmax', C.c_int32 * 3), ('nlow', C.c_int32), ('nhigh', C.c_int32),
                ('grid', C.c_int32 * 3), ('range', C.c_int32 * 3), ('bin0', C.c_int32), ('nbin', C.c_int32),
                ('any_stab', C.c_int32), ('box_safe', C.c_int32), ('pot', YnbPot * MAX_POT),
                ('alpha', C.c_double), ('dielectric', C.c_double), ('k0', C.c_int32), ('nk', C.c_int32),
                ('frb0', C.c_int32), ('nfrb', C.c_int32), ('nslot', C.c_int32), ('ew_nlow', C.c_int32),
                ('chunk0', C.c_int32), ('nchunk', C.c_int32), ('spart0', C.c_int64), ('nchunk_low', C.c_int32),
                ('nchunk_high', C.c_int32), ('pair_form', C.c_int32), ('ew_nhigh', C.c_int32),
                ('ew_const', C.c_double)]


class YnbBatch(C.Structure):
    _fields_ = [('sys', C.c_void_p), ('wave_sys', C.c_void_p), ('bin_sys', C.c_void_p), ('nsys', C.c_int32),
                ('nbin', C.c_int32), ('npad', C.c_int64), ('max_bin_atoms', C.c_int32), ('flags', C.c_int32)]


class YnbNlist(C.Structure):
    _fields_ = [('j', C.c_void_p), ('img', C.c_void_p), ('cnt', C.c_void_p), ('status', C.c_void_p),
                ('rebuild', C.c_void_p), ('kcap', C.c_int32), ('flags', C.c_int32)]


class YnbPairpar(C.Structure):
    _fields_ = [('par', C.c_void_p), ('cross', C.c_void_p), ('stab', C.c_void_p), ('rec', C.c_void_p),
                ('ctr', C.c_void_p), ('lanes', C.c_int32), ('pad_', C.c_int32)]


class YnbEwald(C.Structure):
    _fields_ = [('kvec', C.c_void_p), ('krow', C.c_void_p), ('krow_k', C.c_void_p), ('skblk', C.c_void_p),
                ('fblk', C.c_void_p), ('chunk', C.c_void_p), ('work', C.c_void_p), ('kblk_sys', C.c_void_p),
                ('kpos', C.c_void_p), ('krow_f', C.c_void_p), ('fctr', C.c_void_p),
                ('nwork', C.c_int32), ('nk_tot', C.c_int32),
                ('apl', C.c_int32), ('fslots', C.c_int32),
                ('groups', C.c_int32), ('fapl', C.c_int32), ('flags', C.c_int32), ('pad_', C.c_int32),
                ('frac', C.c_void_p), ('step', C.c_void_p),
                ('spart', C.c_void_p), ('coef', C.c_void_p), ('frec', C.c_void_p), ('kpart', C.c_void_p),
                ('gpart', C.c_void_p)]


class YnbCorrSys(C.Structure):
    _fields_ = [('alpha', C.c_double), ('dielectric', C.c_double), ('self_energy', C.c_double),
                ('nlow', C.c_int32), ('nhigh', C.c_int32), ('row0', C.c_int32), ('nrow', C.c_int32),
                ('blk0', C.c_int32), ('nblk', C.c_int32)]


class YnbCorr(C.Structure):
    _fields_ = [('csys', C.c_void_p), ('stab', C.c_void_p), ('blk_sys', C.c_void_p), ('atom_ref0', C.c_void_p),
                ('atom_ref', C.c_void_p), ('rowg', C.c_void_p), ('partial', C.c_void_p), ('ctr', C.c_void_p),
                ('nrow_tot', C.c_int32), ('nblk_tot', C.c_int32)]


EXPORTS = ['ynb_version', 'ynb_last_error', 'ynb_sys_plan', 'ynb_build_search_bins', 'ynb_nlist_workspace_bytes',
           'ynb_rows_workspace_bytes', 'ynb_nlist_build', 'ynb_nlist_rows', 'ynb_nlist_recompute',
           'ynb_pair_compute', 'ynb_pair_rows', 'ynb_pair_rows_static', 'ynb_ewald_reci', 'ynb_ewald_corr', 'ynb_verlet_kick_drift', 'ynb_verlet_kick',
           'ynb_verlet_kick_kick_drift',
           'ynb_total', 'ynb_profile_enable', 'ynb_profile_count', 'ynb_profile_read',
           'ynb_nlist_skin_check', 'ynb_ewald_kvectors', 'ynb_row_blocks', 'ynb_scatter',
           'ynb_ewald_records', 'ynb_ewald_work', 'ynb_ewald_work_device']


class YnbError(RuntimeError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Loads libyaff_nb.so once; raises OSError when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError('libyaff_nb.so is not built (%s); run python -m yaff_amd.build' % path)
    L = C.CDLL(path)
    vp, i32, i64, sz, P = C.c_void_p, C.c_int32, C.c_int64, C.c_size_t, C.POINTER
    L.ynb_version.restype = C.c_char_p
    L.ynb_last_error.restype = C.c_char_p
    L.ynb_sys_plan.argtypes = [P(YnbSys)]
    L.ynb_build_search_bins.argtypes = [P(YnbSys)]
    L.ynb_ewald_kvectors.argtypes = [vp, C.c_double, C.c_double, C.c_double, vp, i64, vp, vp, vp, vp, P(i64),
                                     P(i64)]
    L.ynb_row_blocks.argtypes = [vp, i64, i32, vp, P(i64)]
    L.ynb_ewald_records.argtypes = [vp, i64, i64, vp, vp, P(i64)]
    L.ynb_ewald_work.argtypes = [i64, i64, i32, i32, i32, vp]
    L.ynb_ewald_work_device.argtypes = [vp, i64, i64, i32, i32, i32, vp]
    L.ynb_scatter.argtypes = [vp, i32, vp, vp]
    L.ynb_nlist_workspace_bytes.restype = sz
    L.ynb_nlist_workspace_bytes.argtypes = [i64, i32, i32]
    L.ynb_rows_workspace_bytes.restype = sz
    L.ynb_rows_workspace_bytes.argtypes = [i64, i64]
    L.ynb_nlist_build.argtypes = [P(YnbBatch), vp, vp, sz, P(YnbNlist), vp]
    L.ynb_nlist_skin_check.argtypes = [P(YnbBatch), vp, vp, vp, vp, vp]
    L.ynb_nlist_rows.argtypes = [P(YnbBatch), i32, vp, P(YnbNlist), vp, sz, vp, i64, P(i64), vp]
    L.ynb_nlist_recompute.argtypes = [P(YnbSys), vp, vp, vp, i64, vp]
    L.ynb_pair_compute.argtypes = [P(YnbBatch), P(YnbPairpar), vp, P(YnbNlist), C.c_uint32, i32, vp, i32, vp,
                                   vp, vp, vp]
    L.ynb_pair_rows.argtypes = [vp, i64, vp, i64, P(YnbPot), vp, vp, vp, vp, i32, i64, vp, vp, vp, vp, vp, vp, vp]
    L.ynb_pair_rows_static.argtypes = [vp, i64, vp, i64, P(YnbPot), vp, vp, vp, vp, i32, i64, vp, vp, vp, vp, vp]
    L.ynb_ewald_reci.argtypes = [P(YnbBatch), P(YnbEwald), vp, vp, i32, vp, i32, vp, vp]
    L.ynb_ewald_corr.argtypes = [P(YnbBatch), P(YnbCorr), vp, vp, vp, i32, vp, i32, vp]
    L.ynb_verlet_kick_drift.argtypes = [P(YnbBatch), vp, vp, vp, vp, vp, P(YnbEwald), vp, vp, sz, P(YnbNlist), vp]
    L.ynb_verlet_kick.argtypes = [P(YnbBatch)] + [vp]*13
    L.ynb_verlet_kick_kick_drift.argtypes = [P(YnbBatch)] + [vp]*13 + [P(YnbEwald), vp, vp, sz, P(YnbNlist), vp]
    L.ynb_total.argtypes = [P(YnbBatch), vp, vp, vp, vp]
    L.ynb_profile_enable.argtypes = [C.c_int]
    L.ynb_profile_read.argtypes = [C.c_int, P(C.c_char_p), P(C.c_float)]
    for name in EXPORTS:
        if name not in ('ynb_version', 'ynb_last_error', 'ynb_nlist_workspace_bytes', 'ynb_rows_workspace_bytes'):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def row_refs(rows, natom):
    """(atom_ref0, atom_ref) of ynb_pair_rows for Yaff rows (fields a, b):
    each atom's references in ascending row order, 2 r where it is row r's
    b and 2 r + 1 where it is a (b first for a self-image row)."""
    n = len(rows)
    atoms = np.empty(2*n, np.int64)
    atoms[0::2] = rows['b']
    atoms[1::2] = rows['a']
    refs = np.arange(2*n, dtype=np.int64)
    refs[0::2] = 2*np.arange(n)
    refs[1::2] = 2*np.arange(n) + 1
    order = np.argsort(atoms, kind='stable')
    ref0 = np.zeros(natom + 1, np.int32)
    ref0[1:] = np.cumsum(np.bincount(atoms, minlength=natom)[:natom])
    return ref0, refs[order].astype(np.int32)


def check(rc):
    """Raises YnbError with the library's message on a nonzero return code."""
    if rc != 0:
        raise YnbError('libyaff_nb: %s (code %d)' % (load().ynb_last_error().decode(), rc))
    return rc


def profile_enable(on=True):
    check(load().ynb_profile_enable(1 if on else 0))


def profile_read():
    """[(kernel name, milliseconds)] of the launches since profile_enable(True)."""
    L = load()
    out = []
    name = C.c_char_p()
    ms = C.c_float()
    for i in range(L.ynb_profile_count()):
        check(L.ynb_profile_read(i, C.byref(name), C.byref(ms)))
        out.append((name.value.decode(), ms.value))
    return out
