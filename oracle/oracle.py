"""ctypes driver for the test oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, never by the yaff_amd package.

Two CPU implementations of Yaff's FP64 nonbonded path are exposed with the
same Python signatures:

* ``Oracle()``            -- oracle/liboracle.so, the restatement in
                             yaff_oracle.c (travels with the repository);
* ``Oracle(reference=True)`` -- oracle/_ref/libyaffref.so, molmod/yaff's own C
                             kernels compiled from their sources by
                             oracle/Makefile (present only where it was built).

Host-side pieces that Yaff computes in Python/Cython are restated here with
numpy: gvecs from the SVD of the cell (yaff/pes/ext.pyx:103-130), rmax
(yaff/pes/nlist.py:132), gmax (yaff/pes/ff.py:386), the neutralizing
background term (yaff/pes/ff.py:636-669) and the scaling table
(yaff/pes/scaling.py:62-106 with yaff/system.py:199-234).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ROW_DTYPE = np.dtype([('a', np.int64), ('b', np.int64), ('d', np.float64),
                      ('dx', np.float64), ('dy', np.float64), ('dz', np.float64),
                      ('r0', np.int64), ('r1', np.int64), ('r2', np.int64)])
SROW_DTYPE = np.dtype([('a', np.int64), ('b', np.int64), ('scale', np.float64),
                       ('nbond', np.int64)])

LJ, MM3, EXPREP, EI = 0, 1, 2, 3
TR_NONE, TR_SWITCH3, TR_HAMMER = 0, 1, 2


class CCell(C.Structure):
    _fields_ = [('rvecs', C.c_double * 9), ('gvecs', C.c_double * 9),
                ('rspacings', C.c_double * 3), ('gspacings', C.c_double * 3),
                ('volume', C.c_double), ('nvec', C.c_int)]


class CPot(C.Structure):
    _fields_ = [('kind', C.c_int), ('trunc', C.c_int), ('rcut', C.c_double),
                ('trunc_par', C.c_double), ('p0', C.c_void_p), ('p1', C.c_void_p),
                ('onlypauli', C.c_void_p), ('ffatype_ids', C.c_void_p),
                ('amp_cross', C.c_void_p), ('b_cross', C.c_void_p),
                ('nffatype', C.c_long), ('alpha', C.c_double), ('dielectric', C.c_double)]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def cell_vectors(rvecs):
    """(rvecs9, gvecs9, nvec) exactly as Cell.update_rvecs builds them
    (yaff/pes/ext.pyx:103-130)."""
    if rvecs is None or np.asarray(rvecs).size == 0:
        return np.identity(3).ravel().copy(), np.identity(3).ravel().copy(), 0
    rvecs = np.ascontiguousarray(rvecs, dtype=float)
    nvec = len(rvecs)
    Up, Sp, Vt = np.linalg.svd(rvecs, full_matrices=True)
    S = np.ones(3, float)
    S[:nvec] = Sp
    U = np.identity(3, float)
    U[:nvec, :nvec] = Up
    mod = np.dot(U*S, Vt)
    mod[:nvec] = rvecs
    gvecs = np.dot(U/S, Vt)
    return np.ascontiguousarray(mod).ravel().copy(), np.ascontiguousarray(gvecs).ravel().copy(), nvec


def rmax_of(cell, rcut_plus_skin):
    """yaff/pes/nlist.py:132"""
    rsp = np.array(cell.rspacings[:cell.nvec])
    return np.ceil(rcut_plus_skin/rsp - 0.5).astype(np.int64)


def gmax_of(cell, gcut):
    """yaff/pes/ff.py:386"""
    gsp = np.array(cell.gspacings[:cell.nvec])
    return np.ceil(gcut/gsp - 0.5).astype(np.int64)


def neighs_from_bonds(natom, bonds):
    """1-2, 1-3, 1-4 and 1-5 neighbour sets (yaff/system.py:199-234)."""
    n1 = {i: set() for i in range(natom)}
    for i0, i1 in (bonds if bonds is not None else []):
        n1[int(i0)].add(int(i1))
        n1[int(i1)].add(int(i0))
    n2 = {i: set() for i in range(natom)}
    for i0, s0 in n1.items():
        for i1 in s0:
            for i2 in n1[i1]:
                if i2 > i0 and i2 not in n1[i0]:
                    n2[i0].add(i2)
                    n2[i2].add(i0)
    n3 = {i: set() for i in range(natom)}
    for i0, s0 in n1.items():
        for i1 in s0:
            for i3 in n2[i1]:
                if i3 != i0 and i3 not in n1[i0] and i3 not in n2[i0]:
                    n3[i0].add(i3)
                    n3[i3].add(i0)
    n4 = {i: set() for i in range(natom)}
    for i0, s0 in n1.items():
        for i1 in s0:
            for i4 in n3[i1]:
                if i4 != i0 and i4 not in n1[i0] and i4 not in n2[i0] and i4 not in n3[i0]:
                    n4[i0].add(i4)
                    n4[i4].add(i0)
    return n1, n2, n3, n4


def scaling_table(natom, bonds, scale1=0.0, scale2=0.0, scale3=1.0, scale4=1.0):
    """yaff/pes/scaling.py:62-101"""
    n1, n2, n3, n4 = neighs_from_bonds(natom, bonds)
    rows = []
    for i0 in range(natom):
        for scale, nb, neighs in ((scale1, 1, n1), (scale2, 2, n2), (scale3, 3, n3), (scale4, 4, n4)):
            if scale < 1.0:
                for i1 in neighs[i0]:
                    if i0 > i1:
                        rows.append((i0, i1, scale, nb))
    rows.sort()
    return np.array(rows, dtype=SROW_DTYPE)


def ewald_neutralizing(charges, alpha, volume, dielectric=1.0, nlow=0, nhigh=None, radii=None):
    """yaff/pes/ff.py:636-669 (non-fluctuating charges); returns energy; the
    virial diagonal receives -energy."""
    q = np.asarray(charges, float)
    nhigh = len(q) if nhigh is None else nhigh
    fac = q.sum()**2/alpha**2
    fac -= q[:nlow].sum()**2/alpha**2
    fac -= q[nhigh:].sum()**2/alpha**2
    if radii is not None:
        fac -= q.sum()*np.sum(q*radii**2)
        fac += q[:nlow].sum()*np.sum(q[:nlow]*radii[:nlow]**2)
        fac += q[nhigh:].sum()*np.sum(q[nhigh:]*radii[nhigh:]**2)
    prefactor = fac*np.pi/(2.0*dielectric)
    return prefactor/volume


class Pot(object):
    """Description of one pair potential for the oracle."""

    def __init__(self, kind, rcut, trunc=TR_NONE, trunc_par=0.0, p0=None, p1=None,
                 onlypauli=None, ffatype_ids=None, amp_cross=None, b_cross=None,
                 alpha=0.0, dielectric=1.0):
        self.kind, self.rcut, self.trunc, self.trunc_par = kind, float(rcut), trunc, float(trunc_par)
        self.p0 = None if p0 is None else np.ascontiguousarray(p0, dtype=np.float64)
        self.p1 = None if p1 is None else np.ascontiguousarray(p1, dtype=np.float64)
        if kind == EI and self.p1 is None:
            self.p1 = np.zeros_like(self.p0)
        self.onlypauli = None if onlypauli is None else np.ascontiguousarray(onlypauli, dtype=np.int32)
        self.ffatype_ids = None if ffatype_ids is None else np.ascontiguousarray(ffatype_ids, dtype=np.int64)
        self.amp_cross = None if amp_cross is None else np.ascontiguousarray(amp_cross, dtype=np.float64)
        self.b_cross = None if b_cross is None else np.ascontiguousarray(b_cross, dtype=np.float64)
        self.alpha, self.dielectric = float(alpha), float(dielectric)

    def cstruct(self):
        nff = 0 if self.amp_cross is None else self.amp_cross.shape[0]
        return CPot(self.kind, self.trunc, self.rcut, self.trunc_par, _ptr(self.p0), _ptr(self.p1),
                    _ptr(self.onlypauli), _ptr(self.ffatype_ids), _ptr(self.amp_cross),
                    _ptr(self.b_cross), nff, self.alpha, self.dielectric)


class Oracle(object):
    """FP64 CPU implementation of the path; ``reference=True`` loads Yaff's own
    C compiled into oracle/_ref."""

    def __init__(self, reference=False):
        self.reference = reference
        path = os.path.join(HERE, '_ref', 'libyaffref.so') if reference else os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise OSError('oracle library %s is not built (run make -C oracle)' % path)
        self.lib = L = C.CDLL(path)
        vp, d, l = C.c_void_p, C.c_double, C.c_long
        if reference:
            L.cell_new.restype = vp
            L.cell_update.argtypes = [vp, vp, vp, C.c_int]
            L.cell_free.argtypes = [vp]
            L.nlist_build_low.restype = C.c_int
            L.nlist_build_low.argtypes = [vp, d, vp, vp, vp, vp, l, l, l, l]
            L.nlist_recompute_low.argtypes = [vp, vp, vp, vp, l]
            L.pair_pot_new.restype = vp
            L.pair_pot_free.argtypes = [vp]
            L.pair_pot_set_rcut.argtypes = [vp, d]
            L.pair_pot_set_trunc_scheme.argtypes = [vp, vp]
            L.pair_data_free.argtypes = [vp]
            L.pair_data_lj_init.argtypes = [vp, vp, vp]
            L.pair_data_mm3_init.argtypes = [vp, vp, vp, vp]
            L.pair_data_exprep_init.argtypes = [vp, l, vp, vp, vp]
            L.pair_data_ei_init.argtypes = [vp, vp, d, d, vp]
            L.switch3_new.restype = vp
            L.switch3_new.argtypes = [d]
            L.hammer_new.restype = vp
            L.hammer_new.argtypes = [d]
            L.trunc_scheme_free.argtypes = [vp]
            L.trunc_scheme_fn.restype = d
            L.trunc_scheme_fn.argtypes = [vp, d, d, vp]
            L.pair_pot_compute.restype = d
            L.pair_pot_compute.argtypes = [vp, l, vp, l, vp, vp, vp]
            L.pair_pot_tailcorr_cut.argtypes = [vp, l, l, l, vp]
            L.pair_pot_tailcorr_switch3.argtypes = [vp, l, l, l, vp]
            L.compute_ewald_reci.restype = d
            L.compute_ewald_reci.argtypes = [vp, l, l, l, vp, vp, d, vp, d, d, vp, vp, vp]
            L.compute_ewald_corr.restype = d
            L.compute_ewald_corr.argtypes = [vp, vp, vp, d, vp, l, d, vp, vp, l, l, l]
            L.cell_mic.argtypes = [vp, vp]
            L.cell_add_vec.argtypes = [vp, vp, vp]
            L.cell_to_center.argtypes = [vp, vp, vp]
        else:
            L.orc_cell_update.argtypes = [vp, vp, vp, C.c_int]
            L.orc_mic.argtypes = [vp, vp]
            L.orc_add_vec.argtypes = [vp, vp, vp]
            L.orc_to_center.argtypes = [vp, vp, vp]
            L.orc_nlist_build.restype = l
            L.orc_nlist_build.argtypes = [vp, l, vp, d, vp, l, l, vp, l]
            L.orc_nlist_recompute.argtypes = [vp, vp, vp, vp, l]
            L.orc_switch3.restype = d
            L.orc_switch3.argtypes = [d, d, d, vp]
            L.orc_hammer.restype = d
            L.orc_hammer.argtypes = [d, d, d, vp]
            L.orc_pair_compute.restype = d
            L.orc_pair_compute.argtypes = [vp, l, vp, l, vp, vp, vp]
            L.orc_ewald_reci.restype = d
            L.orc_ewald_reci.argtypes = [vp, l, l, l, vp, vp, d, vp, d, d, vp, vp, vp]
            L.orc_ewald_corr.restype = d
            L.orc_ewald_corr.argtypes = [vp, vp, vp, d, vp, l, d, vp, vp, l, l, l]

    # -- cell -------------------------------------------------------------
    def cell(self, rvecs):
        """Returns a CCell filled as cell_update does (yaff/pes/cell.c:37-79)."""
        rv, gv, nvec = cell_vectors(rvecs)
        c = CCell()
        if self.reference:
            p = self.lib.cell_new()
            self.lib.cell_update(p, _ptr(rv), _ptr(gv), nvec)
            C.memmove(C.byref(c), p, C.sizeof(CCell))
            self.lib.cell_free(p)
        else:
            self.lib.orc_cell_update(C.byref(c), _ptr(rv), _ptr(gv), nvec)
        return c

    def mic(self, delta, cell):
        delta = np.ascontiguousarray(delta, dtype=float).copy()
        (self.lib.cell_mic if self.reference else self.lib.orc_mic)(_ptr(delta), C.byref(cell))
        return delta

    def add_vec(self, delta, cell, r):
        delta = np.ascontiguousarray(delta, dtype=float).copy()
        r = np.ascontiguousarray(r, dtype=np.int64)
        (self.lib.cell_add_vec if self.reference else self.lib.orc_add_vec)(_ptr(delta), C.byref(cell), _ptr(r))
        return delta

    def to_center(self, v, cell):
        v = np.ascontiguousarray(v, dtype=float)
        out = np.zeros(3, np.int64)
        (self.lib.cell_to_center if self.reference else self.lib.orc_to_center)(_ptr(v), C.byref(cell), _ptr(out))
        return out[:cell.nvec]

    # -- neighbour list ---------------------------------------------------
    def nlist_build(self, pos, cell, rcut, rmax, nlow=0, nhigh=-1, chunk=None):
        """All rows with d < rcut, in Yaff's order.  With ``reference=True``
        the resumable protocol of yaff/pes/nlist.py:163-184 is driven (buffer
        of 10 rows grown by 3/2, or a fixed ``chunk``)."""
        pos = np.ascontiguousarray(pos, dtype=float)
        natom = len(pos)
        nhigh = natom if nhigh == -1 else nhigh
        rmax = np.ascontiguousarray(rmax, dtype=np.int64)
        if self.reference:
            status = np.array([0, 0, 0, nlow, 0, 1, 0], np.int64)
            rows = np.empty(10 if chunk is None else chunk, ROW_DTYPE)
            start = 0
            while True:
                view = rows[start:]
                done = self.lib.nlist_build_low(_ptr(pos), rcut, _ptr(rmax), C.byref(cell), _ptr(status),
                                                _ptr(view), natom, nlow, nhigh, len(view))
                if done:
                    break
                start = len(rows)
                grow = (len(rows)*3)//2 if chunk is None else len(rows) + chunk
                new = np.empty(grow, ROW_DTYPE)
                new[:start] = rows
                rows = new
            return rows[:status[6]].copy()
        cap = 1024
        while True:
            rows = np.empty(cap, ROW_DTYPE)
            n = self.lib.orc_nlist_build(_ptr(pos), natom, C.byref(cell), rcut, _ptr(rmax), nlow, nhigh,
                                         _ptr(rows), cap)
            if n <= cap:
                return rows[:n].copy()
            cap = n

    def nlist_recompute(self, pos, pos_old, cell, rows):
        pos = np.ascontiguousarray(pos, dtype=float)
        pos_old = np.ascontiguousarray(pos_old, dtype=float)
        rows = rows.copy()
        fn = self.lib.nlist_recompute_low if self.reference else self.lib.orc_nlist_recompute
        fn(_ptr(pos), _ptr(pos_old), C.byref(cell), _ptr(rows), len(rows))
        return rows

    # -- truncation -------------------------------------------------------
    def trunc(self, kind, par, d, rcut):
        g = C.c_double(0.0)
        if self.reference:
            t = self.lib.switch3_new(par) if kind == TR_SWITCH3 else self.lib.hammer_new(par)
            h = self.lib.trunc_scheme_fn(t, d, rcut, C.byref(g))
            self.lib.trunc_scheme_free(t)
        else:
            fn = self.lib.orc_switch3 if kind == TR_SWITCH3 else self.lib.orc_hammer
            h = fn(d, rcut, par, C.byref(g))
        return h, g.value

    # -- pair potentials --------------------------------------------------
    def pair_compute(self, rows, stab, pot, natom, want_gpos=True, want_vtens=True):
        """Returns (energy, gpos or None, vtens or None) for one potential."""
        rows = np.ascontiguousarray(rows, dtype=ROW_DTYPE)
        stab = np.ascontiguousarray(stab if stab is not None else np.zeros(0, SROW_DTYPE), dtype=SROW_DTYPE)
        gpos = np.zeros((natom, 3)) if want_gpos else None
        vtens = np.zeros((3, 3)) if want_vtens else None
        if self.reference:
            L = self.lib
            pp = L.pair_pot_new()
            L.pair_pot_set_rcut(pp, pot.rcut)
            tr = None
            if pot.trunc == TR_SWITCH3:
                tr = L.switch3_new(pot.trunc_par)
            elif pot.trunc == TR_HAMMER:
                tr = L.hammer_new(pot.trunc_par)
            L.pair_pot_set_trunc_scheme(pp, tr)
            if pot.kind == LJ:
                L.pair_data_lj_init(pp, _ptr(pot.p0), _ptr(pot.p1))
            elif pot.kind == MM3:
                L.pair_data_mm3_init(pp, _ptr(pot.p0), _ptr(pot.p1), _ptr(pot.onlypauli))
            elif pot.kind == EXPREP:
                L.pair_data_exprep_init(pp, pot.amp_cross.shape[0], _ptr(pot.ffatype_ids),
                                        _ptr(pot.amp_cross), _ptr(pot.b_cross))
            else:
                L.pair_data_ei_init(pp, _ptr(pot.p0), pot.alpha, pot.dielectric, _ptr(pot.p1))
            e = L.pair_pot_compute(_ptr(rows), len(rows), _ptr(stab), len(stab), pp, _ptr(gpos), _ptr(vtens))
            L.pair_data_free(pp)
            L.pair_pot_free(pp)
            if tr is not None:
                L.trunc_scheme_free(tr)
        else:
            cp = pot.cstruct()
            e = self.lib.orc_pair_compute(_ptr(rows), len(rows), _ptr(stab), len(stab), C.byref(cp),
                                          _ptr(gpos), _ptr(vtens))
        return e, gpos, vtens

    # -- Ewald --------------------------------------------------------------
    def tailcorr(self, pot, natom, nlow=0, nhigh=-1):
        """(C0, C1) of pair_pot_tailcorr_cut / _switch3 (pair_pot.c:178-254)
        for an LJ, MM3 or ExpRep potential; the compiled reference only."""
        if not self.reference:
            raise NotImplementedError('tail corrections: compiled reference only')
        if nhigh == -1:
            nhigh = natom
        L = self.lib
        pp = L.pair_pot_new()
        L.pair_pot_set_rcut(pp, pot.rcut)
        tr = L.switch3_new(pot.trunc_par) if pot.trunc == TR_SWITCH3 else None
        L.pair_pot_set_trunc_scheme(pp, tr)
        if pot.kind == LJ:
            L.pair_data_lj_init(pp, _ptr(pot.p0), _ptr(pot.p1))
        elif pot.kind == MM3:
            L.pair_data_mm3_init(pp, _ptr(pot.p0), _ptr(pot.p1), _ptr(pot.onlypauli))
        elif pot.kind == EXPREP:
            L.pair_data_exprep_init(pp, pot.amp_cross.shape[0], _ptr(pot.ffatype_ids),
                                    _ptr(pot.amp_cross), _ptr(pot.b_cross))
        else:
            raise ValueError('no tail correction for this potential')
        corrs = np.zeros(2)
        fn = L.pair_pot_tailcorr_switch3 if tr is not None else L.pair_pot_tailcorr_cut
        fn(_ptr(corrs), natom, nlow, nhigh, pp)
        L.pair_data_free(pp)
        L.pair_pot_free(pp)
        if tr is not None:
            L.trunc_scheme_free(tr)
        return float(corrs[0]), float(corrs[1])

    def ewald_reci(self, pos, charges, cell, alpha, gmax, gcut, dielectric=1.0, nlow=0, nhigh=-1,
                   want_gpos=True, want_vtens=True):
        pos = np.ascontiguousarray(pos, dtype=float)
        q = np.ascontiguousarray(charges, dtype=float)
        natom = len(pos)
        nhigh = natom if nhigh == -1 else nhigh
        gmax = np.ascontiguousarray(gmax, dtype=np.int64)
        gpos = np.zeros((natom, 3)) if want_gpos else None
        work = np.zeros(2*natom) if want_gpos else None
        vtens = np.zeros((3, 3)) if want_vtens else None
        fn = self.lib.compute_ewald_reci if self.reference else self.lib.orc_ewald_reci
        e = fn(_ptr(pos), natom, nlow, nhigh, _ptr(q), C.byref(cell), alpha, _ptr(gmax), gcut, dielectric,
               _ptr(gpos), _ptr(work), _ptr(vtens))
        return e, gpos, vtens

    def ewald_corr(self, pos, charges, cell, alpha, stab, dielectric=1.0, nlow=0, nhigh=-1,
                   want_gpos=True, want_vtens=True):
        pos = np.ascontiguousarray(pos, dtype=float)
        q = np.ascontiguousarray(charges, dtype=float)
        natom = len(pos)
        nhigh = natom if nhigh == -1 else nhigh
        stab = np.ascontiguousarray(stab if stab is not None else np.zeros(0, SROW_DTYPE), dtype=SROW_DTYPE)
        gpos = np.zeros((natom, 3)) if want_gpos else None
        vtens = np.zeros((3, 3)) if want_vtens else None
        fn = self.lib.compute_ewald_corr if self.reference else self.lib.orc_ewald_corr
        e = fn(_ptr(pos), _ptr(q), C.byref(cell), alpha, _ptr(stab), len(stab), dielectric,
               _ptr(gpos), _ptr(vtens), natom, nlow, nhigh)
        return e, gpos, vtens


def reference_available():
    return os.path.exists(os.path.join(HERE, '_ref', 'libyaffref.so'))
