"""Scalings of bonded pairs (yaff/pes/scaling.py:58-106) and the check that
each scaled pair is one well-defined image (yaff/pes/scaling.py:108-170)."""
import numpy as np

from ._lib import SCALING_DTYPE

__all__ = ['Scalings', 'scaling_dtype', 'bond_paths']

scaling_dtype = SCALING_DTYPE


class Scalings(object):
    def __init__(self, system, scale1=0.0, scale2=0.0, scale3=1.0, scale4=1.0):
        for name, s in (('scale1', scale1), ('scale2', scale2), ('scale3', scale3), ('scale4', scale4)):
            if s < -1 or s > 1:
                raise ValueError('%s must be in the range [-1,1].' % name)
        self.scale1, self.scale2, self.scale3, self.scale4 = scale1, scale2, scale3, scale4
        rows = []
        for i0 in sorted(set(system.neighs1) | set(system.neighs2) | set(system.neighs3) | set(system.neighs4)):
            for scale, nb, neighs in ((scale1, 1, system.neighs1), (scale2, 2, system.neighs2),
                                      (scale3, 3, system.neighs3), (scale4, 4, system.neighs4)):
                if scale < 1.0:
                    for i1 in neighs[i0]:
                        if i0 > i1:
                            rows.append((i0, i1, scale, nb))
        rows.sort()
        self.stab = np.array(rows, dtype=SCALING_DTYPE)
        self.items = []
        self.check_mic(system)

    def check_mic(self, system):
        """A pair scaled as 2, 3 or 4 bonds apart must be one relative vector:
        every bond path of that length between the two atoms, summed bond by
        bond with the minimum image convention, must give the same vector.
        In a cell with planes closer than the bond paths this fails, and the
        pair scaling would act on the wrong image: AssertionError (use a
        supercell)."""
        if system.cell.nvec == 0:
            return
        pos = system.pos
        bad = []
        for a, b, _, nbond in self.stab:
            if nbond < 2:
                continue
            sums = [self._path_vector(system, path) for path in bond_paths(system, int(a), int(b), int(nbond))]
            if len(sums) > 1 and np.abs(np.array(sums) - np.mean(sums, axis=0)).max() > 1e-10:
                bad.append((int(a), int(b)))
        if bad:
            raise AssertionError('scaled pairs %s are joined by bond paths with different relative vectors: the '
                                 'cell is too small for these exclusions; use a supercell' % bad[:10])

    @staticmethod
    def _path_vector(system, path):
        total = np.zeros(3)
        for i0, i1 in zip(path[:-1], path[1:]):
            delta = system.pos[i0] - system.pos[i1]
            system.cell.mic(delta)
            total += delta
        return total


def bond_paths(system, start, end, nbond):
    """The simple paths of nbond bonds from start to end in the bond graph
    (tuples of atom indices, no atom twice), by depth-first search."""
    out = []
    stack = [(start,)]
    while stack:
        path = stack.pop()
        if len(path) == nbond + 1:
            if path[-1] == end:
                out.append(path)
            continue
        for nxt in sorted(system.neighs1[path[-1]]):
            if nxt not in path and (len(path) < nbond or nxt == end):
                stack.append(path + (nxt,))
    return out
