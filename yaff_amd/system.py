"""Minimal ``System`` with the attributes the nonbonded path reads.

Mirrors the constructor of yaff.System (yaff/system.py:50-170) for the fields
used by NeighborList, the pair potentials and the Ewald parts, the bond-derived
neighbour sets (yaff/system.py:199-234) and ``supercell`` (yaff/system.py:841-935).
"""
import numpy as np

from .cell import Cell

__all__ = ['System']


class _Neighbours(dict):
    """Neighbour sets by atom index; an atom without neighbours is absent and
    reads as an empty set."""

    def __missing__(self, key):
        return frozenset()


class System(object):
    def __init__(self, numbers, pos, ffatypes=None, ffatype_ids=None, bonds=None, rvecs=None,
                 charges=None, radii=None, masses=None, **kwargs):
        self.numbers = np.asarray(numbers)
        self.pos = np.ascontiguousarray(pos, dtype=float).copy()
        if self.pos.ndim != 2 or self.pos.shape[1] != 3 or len(self.pos) != len(self.numbers):
            raise TypeError('pos must have shape (natom, 3).')
        self.ffatypes = None if ffatypes is None else np.asarray(ffatypes)
        if ffatypes is not None and ffatype_ids is None:
            # Yaff assigns ids in order of first appearance when ffatypes are per atom.
            uniq = []
            ids = []
            for t in ffatypes:
                if t not in uniq:
                    uniq.append(t)
                ids.append(uniq.index(t))
            self.ffatypes = np.array(uniq)
            ffatype_ids = np.array(ids)
        self.ffatype_ids = None if ffatype_ids is None else np.asarray(ffatype_ids, dtype=np.int64)
        self.bonds = None if bonds is None else np.asarray(bonds, dtype=np.int64).reshape(-1, 2)
        self.cell = Cell(None if rvecs is None else np.ascontiguousarray(rvecs, dtype=float))
        self.charges = None if charges is None else np.asarray(charges, dtype=float).copy()
        self.radii = None if radii is None else np.asarray(radii, dtype=float).copy()
        self.masses = None if masses is None else np.asarray(masses, dtype=float).copy()
        self._init_derived_bonds()

    @property
    def natom(self):
        return len(self.pos)

    @property
    def nffatype(self):
        return 0 if self.ffatypes is None else len(self.ffatypes)

    def get_ffatype(self, index):
        """yaff/system.py:333-345: the force-field type name of an atom."""
        if self.ffatypes is None or self.ffatype_ids is None:
            return None
        return self.ffatypes[self.ffatype_ids[index]]

    def iter_bonds(self):
        """yaff/system.py:349-354"""
        if self.bonds is not None:
            for i0, i1 in self.bonds:
                yield int(i0), int(i1)

    @classmethod
    def from_file(cls, *fns, **user_kwargs):
        """yaff/system.py:405-470 for the .chk checkpoint format (later files
        override earlier ones, keywords override files).  .h5 files need
        h5py and .xyz/.psf files molmod, neither of which this package
        uses: they raise IOError."""
        from .chk import load_chk
        allowed = ('numbers', 'pos', 'ffatypes', 'ffatype_ids', 'bonds', 'rvecs', 'charges', 'radii', 'masses')
        kwargs = {}
        for fn in fns:
            if not fn.endswith('.chk'):
                raise IOError('Can not read from file \'%s\' (only .chk is supported).' % fn)
            for key, value in load_chk(fn).items():
                if key in allowed:
                    kwargs[key] = value
        kwargs.update(user_kwargs)
        return cls(**kwargs)

    def _init_derived_bonds(self):
        """yaff/system.py:199-234: the atoms 1, 2, 3 and 4 bonds away from each
        atom (sets by atom; atoms without such neighbours give an empty set)."""
        n1 = _Neighbours()
        for i0, i1 in (self.bonds if self.bonds is not None else []):
            n1.setdefault(int(i0), set()).add(int(i1))
            n1.setdefault(int(i1), set()).add(int(i0))
        n2 = _Neighbours()
        for i0, s0 in n1.items():
            for i1 in s0:
                for i2 in n1[i1]:
                    if i2 > i0 and i2 not in n1[i0]:
                        n2.setdefault(i0, set()).add(i2)
                        n2.setdefault(i2, set()).add(i0)
        n3 = _Neighbours()
        for i0, s0 in n1.items():
            for i1 in s0:
                for i3 in n2[i1]:
                    if i3 != i0 and i3 not in n1[i0] and i3 not in n2[i0]:
                        n3.setdefault(i0, set()).add(i3)
                        n3.setdefault(i3, set()).add(i0)
        n4 = _Neighbours()
        for i0, s0 in n1.items():
            for i1 in s0:
                for i4 in n3[i1]:
                    if i4 != i0 and i4 not in n1[i0] and i4 not in n2[i0] and i4 not in n3[i0]:
                        n4.setdefault(i0, set()).add(i4)
                        n4.setdefault(i4, set()).add(i0)
        self.neighs1, self.neighs2, self.neighs3, self.neighs4 = n1, n2, n3, n4

    def supercell(self, *reps):
        """yaff/system.py:841-935"""
        if self.cell.nvec == 0:
            raise TypeError('Can not create a supercell of a non-periodic system.')
        if self.cell.nvec != len(reps):
            raise TypeError('The number of repetitions must match the number of cell vectors.')
        nrep = int(np.prod(reps))
        kw = {}
        for name in ('numbers', 'ffatype_ids', 'charges', 'radii', 'masses'):
            v = getattr(self, name)
            if v is not None:
                kw[name] = np.tile(v, nrep)
        if self.ffatypes is not None:
            kw['ffatypes_unique'] = self.ffatypes.copy()
        rvecs = self.cell.rvecs*np.array(reps)[:, None]
        pos = np.zeros((self.natom*nrep, 3))
        offsets = {}
        start = 0
        for img in np.ndindex(reps):
            pos[start:start + self.natom] = self.pos + np.dot(img, self.cell.rvecs)
            offsets[img] = start
            start += self.natom
        bonds = None
        if self.bonds is not None:
            rel = []
            for i0, i1 in self.bonds:
                frac = np.dot(self.cell.gvecs, self.pos[i0] - self.pos[i1])
                rel.append(np.ceil(frac - 0.5))
            bonds = np.zeros((len(self.bonds)*nrep, 2), int)
            k = 0
            for img0 in np.ndindex(reps):
                for ib, (i0, i1) in enumerate(self.bonds):
                    img1 = tuple(int((img0[c] + rel[ib][c]) % reps[c]) for c in range(len(reps)))
                    bonds[k] = offsets[img0] + i0, offsets[img1] + i1
                    k += 1
        ffu = kw.pop('ffatypes_unique', None)
        s = System(kw.pop('numbers'), pos, bonds=bonds, rvecs=rvecs, **kw)
        s.ffatypes = ffu
        return s
