"""Periodic cell, as Yaff's ``Cell`` (yaff/pes/ext.pyx:75-130, yaff/pes/cell.c).

The reciprocal vectors come from the same numpy SVD construction as
``Cell.update_rvecs`` (ext.pyx:103-130), so the gvecs handed to the device are
the very bits Yaff's C code would use.  The spacings and the volume follow
``cell_update`` (cell.c:37-79) in plain Python floating point with the same
operation order.
"""
import math

import numpy as np

__all__ = ['Cell']


class Cell(object):
    def __init__(self, rvecs):
        self.update_rvecs(rvecs)

    def update_rvecs(self, rvecs):
        """ext.pyx:103-130"""
        if rvecs is None or np.asarray(rvecs).size == 0:
            mod = np.identity(3, float)
            gvecs = mod.copy()
            nvec = 0
        else:
            rvecs = np.asarray(rvecs)
            if rvecs.ndim != 2 or not rvecs.flags['C_CONTIGUOUS'] or rvecs.shape[0] > 3 or rvecs.shape[1] != 3:
                raise TypeError('rvecs must be a C contiguous array with three columns and at most three rows.')
            rvecs = rvecs.astype(float)
            nvec = len(rvecs)
            Up, Sp, Vt = np.linalg.svd(rvecs, full_matrices=True)
            S = np.ones(3, float)
            S[:nvec] = Sp
            U = np.identity(3, float)
            U[:nvec, :nvec] = Up
            mod = np.dot(U*S, Vt)
            mod[:nvec] = rvecs
            gvecs = np.dot(U/S, Vt)
        self._rvecs_full = np.ascontiguousarray(mod, dtype=float)
        self._gvecs_full = np.ascontiguousarray(gvecs, dtype=float)
        self.nvec = nvec
        r = self._rvecs_full.ravel().tolist()
        g = self._gvecs_full.ravel().tolist()
        self._rspacings = [1.0/math.sqrt(g[3*i]*g[3*i] + g[3*i+1]*g[3*i+1] + g[3*i+2]*g[3*i+2]) for i in range(3)]
        self._gspacings = [1.0/math.sqrt(r[3*i]*r[3*i] + r[3*i+1]*r[3*i+1] + r[3*i+2]*r[3*i+2]) for i in range(3)]
        if nvec == 0:
            vol = 0.0
        elif nvec == 1:
            vol = math.sqrt(r[0]*r[0] + r[1]*r[1] + r[2]*r[2])
        elif nvec == 2:
            t = r[0]*r[3] + r[1]*r[4] + r[2]*r[5]
            t = (r[0]*r[0] + r[1]*r[1] + r[2]*r[2])*(r[3]*r[3] + r[4]*r[4] + r[5]*r[5]) - t*t
            vol = math.sqrt(t) if t > 0 else 0.0
        else:
            vol = abs(r[0]*(r[4]*r[8] - r[5]*r[7]) + r[1]*(r[5]*r[6] - r[3]*r[8]) + r[2]*(r[3]*r[7] - r[4]*r[6]))
        self.volume = vol

    # The attributes Yaff exposes (only the first nvec rows).
    @property
    def rvecs(self):
        return self._rvecs_full[:self.nvec].copy()

    @property
    def gvecs(self):
        return self._gvecs_full[:self.nvec].copy()

    @property
    def rspacings(self):
        return np.array(self._rspacings[:self.nvec])

    @property
    def gspacings(self):
        return np.array(self._gspacings[:self.nvec])

    @property
    def rvecs_full(self):
        return self._rvecs_full

    @property
    def gvecs_full(self):
        return self._gvecs_full

    def mic(self, delta):
        """In-place minimum image, cell.c:81-106."""
        d = [float(delta[0]), float(delta[1]), float(delta[2])]
        r = self._rvecs_full.ravel().tolist()
        g = self._gvecs_full.ravel().tolist()
        for i in range(self.nvec):
            x = math.ceil(g[3*i]*d[0] + g[3*i+1]*d[1] + g[3*i+2]*d[2] - 0.5)
            d[0] -= x*r[3*i]
            d[1] -= x*r[3*i+1]
            d[2] -= x*r[3*i+2]
        delta[:] = d

    def to_center(self, pos):
        """cell.c:109-121"""
        g = self._gvecs_full.ravel().tolist()
        out = np.zeros(self.nvec, int)
        for i in range(self.nvec):
            out[i] = -math.ceil(g[3*i]*pos[0] + g[3*i+1]*pos[1] + g[3*i+2]*pos[2] - 0.5)
        return out

    def add_vec(self, delta, r):
        """In-place delta += r.rvecs, cell.c:124-143."""
        d = [float(delta[0]), float(delta[1]), float(delta[2])]
        rv = self._rvecs_full.ravel().tolist()
        for i in range(self.nvec):
            d[0] += float(r[i])*rv[3*i]
            d[1] += float(r[i])*rv[3*i+1]
            d[2] += float(r[i])*rv[3*i+2]
        delta[:] = d

    def get_rcut_images(self, rcut):
        """Number of images along each axis needed for ``rcut`` (nlist.py:132)."""
        return np.ceil(rcut/self.rspacings - 0.5).astype(np.int64)
