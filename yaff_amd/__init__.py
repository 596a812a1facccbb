"""yaff_amd -- MI355X nonbonded engine behind Yaff's yaff.pes interface.

Drop-in classes (same constructors and compute protocol as yaff.pes):
NeighborList, Switch3, Hammer, PairPotLJ, PairPotMM3, PairPotExpRep,
PairPotEI, Scalings, ForceField, ForcePartPair, ForcePartEwaldReciprocal,
ForcePartEwaldCorrection, ForcePartEwaldNeutralizing, ForcePartTailCorrection,
plus System and Cell.
The arithmetic runs in hand-written HIP kernels (libyaff_nb.so); the batched
Engine underneath also serves multi-system launches, device MD and serving.
"""
from .cell import Cell
from .system import System
from .nlist import NeighborList, neigh_dtype
from .pair_pot import Truncation, Switch3, Hammer, PairPot, PairPotLJ, PairPotMM3, PairPotExpRep, PairPotEI
from .scaling import Scalings, scaling_dtype
from .ff import (ForcePart, ForceField, ForcePartPair, ForcePartEwaldReciprocal, ForcePartEwaldCorrection,
                 ForcePartEwaldNeutralizing, ForcePartTailCorrection)
from .device import set_default_precision
from . import units

__version__ = '0.1.0'
