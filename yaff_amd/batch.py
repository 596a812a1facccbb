"""Batched evaluation and MD of many independent systems in one set of launches.

Every kernel launch covers all systems of the batch (each system owns a
64-aligned atom range, its own bins, k vectors and reductions), so a step of
64 systems costs the same number of launches as a step of one.  MD steps are
captured in a HIP graph and replayed.

``nbsystem`` turns a synthetic system of yaff_amd.workloads (LJ fluid or
rocksalt, recognised by its ``spec``) into an engine NBSystem with the force
field of that family.
"""
import time

import numpy as np
import torch

from . import _lib
from .engine import Engine, NBSystem, PotSpec, EwaldSpec
from .units import femtosecond, boltzmann

__all__ = ['nbsystem', 'forcefield', 'timestep', 'maxwell_boltzmann', 'BatchMD']


def nbsystem(s, skin=0.0, correction=True):
    """Engine input for a workloads.lj_fluid or workloads.rocksalt system;
    with skin > 0 the engine's MD steps rebuild its list only when an atom
    has moved by skin/2.  Rocksalt gets the Ewald reciprocal term and, with
    ``correction``, the Ewald correction (its self term: the rocksalt
    systems have no scaled pairs), as the drop-in force field of
    ``forcefield`` has them."""
    sp = s.spec
    n = s.natom
    if sp.kind == 'lj':
        pots = [PotSpec(_lib.POT_LJ, sp.rcut, _lib.TR_SWITCH3, sp.width, np.full(n, sp.sigma),
                        np.full(n, sp.epsilon))]
        return NBSystem(s.pos, s.cell, pots, masses=s.masses, skin=skin)
    pots = [PotSpec(_lib.POT_EXPREP, sp.rcut, p0=s.ffatype_ids.astype(float),
                    amp_cross=np.full((2, 2), sp.exprep_a), b_cross=np.full((2, 2), sp.exprep_b)),
            PotSpec(_lib.POT_LJ, sp.rcut, p0=np.full(n, sp.lj_sigma), p1=np.full(n, sp.lj_epsilon)),
            PotSpec(_lib.POT_EI, sp.rcut, p0=s.charges, alpha=sp.alpha)]
    return NBSystem(s.pos, s.cell, pots, ewald=EwaldSpec(sp.alpha, sp.gcut, correction=correction),
                    charges=s.charges, masses=s.masses, skin=skin)


def forcefield(s, skin=0.0):
    """The same force field through the drop-in yaff.pes interface: a
    ForceField of ForcePartPair parts (and, for rocksalt, the Ewald
    reciprocal, correction and neutralizing parts), as yaff's
    ForceField.generate assembles them (yaff/pes/generator.py:201-240)."""
    from .nlist import NeighborList
    from .scaling import Scalings
    from .pair_pot import PairPotLJ, PairPotExpRep, PairPotEI, Switch3
    from .ff import (ForceField, ForcePartPair, ForcePartEwaldReciprocal, ForcePartEwaldCorrection,
                     ForcePartEwaldNeutralizing)
    sp = s.spec
    n = s.natom
    nlist = NeighborList(s, skin=skin)
    sc = Scalings(s, 0.0, 0.0, 1.0)
    if sp.kind == 'lj':
        pot = PairPotLJ(np.full(n, sp.sigma), np.full(n, sp.epsilon), sp.rcut, Switch3(sp.width))
        return ForceField(s, [ForcePartPair(s, nlist, sc, pot)], nlist)
    pots = [PairPotExpRep(s.ffatype_ids, np.full((2, 2), sp.exprep_a), np.full((2, 2), sp.exprep_b), sp.rcut),
            PairPotLJ(np.full(n, sp.lj_sigma), np.full(n, sp.lj_epsilon), sp.rcut),
            PairPotEI(s.charges, sp.alpha, sp.rcut)]
    parts = [ForcePartPair(s, nlist, sc, p) for p in pots]
    parts += [ForcePartEwaldReciprocal(s, sp.alpha, sp.gcut), ForcePartEwaldCorrection(s, sp.alpha, sc),
              ForcePartEwaldNeutralizing(s, sp.alpha)]
    return ForceField(s, parts, nlist)


def timestep(s):
    """2 fs for the LJ fluid, 1 fs for rocksalt."""
    return (2.0 if s.spec.kind == 'lj' else 1.0)*femtosecond


def maxwell_boltzmann(masses, temp, seed):
    """Random velocities at temp, rescaled to it (yaff/sampling/utils.py:42-75)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(0, 1, (len(masses), 3))*np.sqrt(boltzmann*temp/masses).reshape(-1, 1)
    t = (v**2*masses.reshape(-1, 1)).mean()/boltzmann
    return v*np.sqrt(temp/t)


class BatchMD(object):
    """Velocity-Verlet MD of a batch of systems, HIP-graph replayed."""

    def __init__(self, systems, precision='mixed', tuning=None, seed=0, temps=None, skin=0.0):
        self.systems = list(systems)
        self.engine = Engine([nbsystem(s, skin) for s in self.systems], precision=precision)
        if tuning:
            self.engine.set_tuning(**tuning)
        eng = self.engine
        eng.ensure_nlist()
        eng.compute(build=False)
        vel = []
        for i, s in enumerate(self.systems):
            t = (temps[i] if temps is not None else (120.0 if s.spec.kind == 'lj' else 300.0))
            vel.append(maxwell_boltzmann(s.masses, t, seed + i))
        eng.set_velocities(vel)
        eng.set_timestep([timestep(s) for s in self.systems])
        self.graph = None

    @property
    def natom(self):
        return sum(s.natom for s in self.systems)

    def warm(self, nsteps=2):
        """Eager steps that also settle the list capacities."""
        eng = self.engine
        for _ in range(nsteps):
            eng.md_step()
        flags, longest = eng.nlist_status()
        if flags:
            eng.grow(flags, longest)
            eng.ensure_nlist()
            eng.compute(build=False)
        self.graph = None

    def capture(self, steps_per_graph=1):
        self.graph = self.engine.capture_md(steps_per_graph)
        self.steps_per_graph = steps_per_graph

    def run(self, nsteps):
        """nsteps MD steps by graph replays; returns wall seconds."""
        if self.graph is None:
            self.capture(1)
        reps = nsteps//self.steps_per_graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.graph.replay()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        flags, _ = self.engine.nlist_status()
        if flags:
            raise RuntimeError('neighbour list capacity exceeded during the replayed steps (flags %d)' % flags)
        return dt
