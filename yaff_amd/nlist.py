"""NeighborList with Yaff's interface (yaff/pes/nlist.py:54-218).

The list itself lives on the device (full ELL list built by
ynb_nlist_build); ``neighs``/``nneigh`` materialise Yaff's 72-byte rows in
nlist_build_low's order on first access after a rebuild, and after a
skin-path update the stored rows are recomputed on the device exactly as
nlist_recompute_low does.
"""
import numpy as np
import torch

from . import _lib
from ._lib import ROW_DTYPE
from .device import SystemDevice

__all__ = ['NeighborList', 'neigh_dtype']

neigh_dtype = ROW_DTYPE


class NeighborList(object):
    def __init__(self, system, skin=0, nlow=0, nhigh=-1):
        if skin < 0:
            raise ValueError('The skin parameter must be positive.')
        if nhigh == -1:
            nhigh = system.natom
        self.system = system
        self.skin = skin
        self.rcut = 0.0
        self.rmax = None
        if nlow < 0:
            raise ValueError('nlow must be a positive number, received %d.' % nlow)
        self.nlow = nlow
        if nhigh < self.nlow:
            raise ValueError('nhigh must not be smaller than nlow, received %d.' % nhigh)
        self.nhigh = nhigh
        self._pos_old = None
        self.rebuild_next = False
        self._rows = None          # host rows (materialised lazily)
        self._rows_d = None        # device rows at the last rebuild
        self._rows_state = None    # 'built' or 'recomputed'
        self._rows_pos_old = None
        self.dev = SystemDevice.of(system)
        self.dev.nlist = self

    def request_rcut(self, rcut):
        """nlist.py:114-117"""
        self.rcut = max(self.rcut, rcut)
        self.update_rmax()

    def update_rmax(self):
        """nlist.py:119-142"""
        self.rmax = np.ceil((self.rcut + self.skin)/self.system.cell.rspacings - 0.5).astype(int)
        self.rebuild_next = True
        self.dev.invalidate()

    def update(self):
        """nlist.py:144-196: rebuild on the device, or keep the pairs and let
        the kernels recompute the vectors from the new positions."""
        assert self.rcut > 0
        dev = self.dev
        if self._need_rebuild():
            if self.system.cell.volume != 0:
                if self.system.natom/self.system.cell.volume > 10:
                    raise ValueError('Atom density too high')
            dev.upload_positions()
            dev.rebuild_list()
            self._rows = None
            self._rows_d = None
            self._rows_state = 'built'
            self._checkpoint()
            self.rebuild_next = False
        else:
            dev.upload_positions()
            self._rows = None
            self._rows_state = 'recomputed'

    def _checkpoint(self):
        if self.skin > 0:
            if self._pos_old is None:
                self._pos_old = self.system.pos.copy()
            else:
                self._pos_old[:] = self.system.pos

    def _need_rebuild(self):
        """nlist.py:207-218"""
        if self.skin <= 0 or self._pos_old is None or self.rebuild_next or self.dev.list_token < 0:
            return True
        disp = np.sqrt(((self.system.pos - self._pos_old)**2).sum(axis=1).max())
        disp *= 2*(self.rmax.max() + 1)
        return disp >= self.skin

    # -- rows ----------------------------------------------------------------
    def _materialise(self):
        eng = self.dev.get_engine()
        if self._rows_d is None:
            rows = eng.rows(0)
            self._rows_d = torch.from_numpy(np.frombuffer(rows.tobytes() or b'\0'*72, np.uint8).copy()).to(eng.device)
            self._rows_n = len(rows)
            self._rows_build_pos = self.system.pos.copy()
            self._rows = rows
            if self._rows_state == 'recomputed':
                self._rows_state = 'built'
            return
        if self._rows_state == 'recomputed':
            # nlist_recompute_low relative to the positions of the last rebuild
            old = torch.zeros_like(eng.pos_d)
            old[:self.system.natom] = torch.from_numpy(self._rows_build_pos)
            _lib.check(eng.lib.ynb_nlist_recompute(eng.hsys[0], eng.pos_d.data_ptr(), old.data_ptr(),
                                                   self._rows_d.data_ptr(), self._rows_n, eng.stream()))
            raw = self._rows_d.cpu().numpy().tobytes()[:self._rows_n*ROW_DTYPE.itemsize]
            self._rows = np.frombuffer(raw, ROW_DTYPE).copy()

    @property
    def neighs(self):
        if self._rows is None:
            self._materialise()
        return self._rows

    @property
    def nneigh(self):
        return len(self.neighs)

    def to_dictionary(self):
        """nlist.py:221-237"""
        out = {}
        for row in self.neighs:
            out[(row['a'], row['b'], row['r0'], row['r1'], row['r2'])] = np.array(
                [row['d'], row['dx'], row['dy'], row['dz']])
        return out
