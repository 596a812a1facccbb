"""Scalings of bonded pairs (yaff/pes/scaling.py:58-106)."""
import numpy as np

from ._lib import SCALING_DTYPE

__all__ = ['Scalings', 'scaling_dtype']

scaling_dtype = SCALING_DTYPE


class Scalings(object):
    def __init__(self, system, scale1=0.0, scale2=0.0, scale3=1.0, scale4=1.0):
        for name, s in (('scale1', scale1), ('scale2', scale2), ('scale3', scale3), ('scale4', scale4)):
            if s < -1 or s > 1:
                raise ValueError('%s must be in the range [-1,1].' % name)
        self.scale1, self.scale2, self.scale3, self.scale4 = scale1, scale2, scale3, scale4
        rows = []
        for i0 in range(system.natom):
            for scale, nb, neighs in ((scale1, 1, system.neighs1), (scale2, 2, system.neighs2),
                                      (scale3, 3, system.neighs3), (scale4, 4, system.neighs4)):
                if scale < 1.0:
                    for i1 in neighs[i0]:
                        if i0 > i1:
                            rows.append((i0, i1, scale, nb))
        rows.sort()
        self.stab = np.array(rows, dtype=SCALING_DTYPE)
        self.items = []
