"""Force fields from Yaff parameter files: the nonbonded generators
(yaff/pes/generator.py:84-410, 1603-1916, 2077-2246).

``ForceField.generate(system, 'pars.txt', ...)`` reads the file
(``Parameters``), lets each generator turn its section into parts of this
package -- LJ, MM3 and EXPREP into ForcePartPair objects, FIXQ into charges
and radii on the system plus the EI real-space part and, for 3-D cells, the
Ewald reciprocal, correction and neutralizing parts -- and adds tail
corrections when asked.  Sections of the valence and other generators of
Yaff are not part of this engine: by default they raise NotImplementedError;
with ``skip_other_sections=True`` they are left out (the caller adds Yaff's
own parts for them), each named in a RuntimeWarning.
"""
import warnings

import numpy as np

from .units import parse_unit

__all__ = ['FFArgs', 'Generator', 'NonbondedGenerator', 'LJGenerator', 'MM3Generator', 'ExpRepGenerator',
           'FixedChargeGenerator', 'apply_generators', 'YAFF_OTHER_PREFIXES']

# Sections Yaff's generator.py knows but this engine does not build
# (valence terms, other pair potentials, grids).
YAFF_OTHER_PREFIXES = frozenset([
    'BONDHARM', 'BONDFUES', 'MM3QUART', 'DOUBWELL', 'BONDMORSE', 'DOUBWELL2', 'POLY4', 'BONDPOLYSIX',
    'BENDAHARM', 'BENDCHARM', 'MM3BENDA', 'UBHARM', 'BENDCOS', 'BENDCLIN', 'TORSAHARM', 'TORSCHARM',
    'TORSION', 'TORSC2HARM', 'TORSCPOLYSIX', 'OOPANGLE', 'OOPMANGLE', 'OOPCOS', 'OOPMCOS', 'OOPDIST',
    'SQOOPDIST', 'IMPROPER', 'CROSS', 'CROSSCBEND', 'CROSSBONDDIHED', 'CROSSBENDDIHED', 'LJCROSS',
    'MM3CAP', 'DAMPDISP', 'D3BJ', 'QMDFFREP', 'BONDCROSS', 'BENDCROSS'])


class FFArgs(object):
    """generator.py:84-241: the arguments of the force field being built."""

    def __init__(self, rcut=18.89726133921252, tr='default', alpha_scale=3.5, gcut_scale=1.1, skin=0,
                 smooth_ei=False, reci_ei='ewald', nlow=0, nhigh=-1, tailcorrections=False):
        from .pair_pot import Switch3
        if reci_ei not in ('ignore', 'ewald', 'ewald_interaction'):
            raise ValueError("The reci_ei option must be one of 'ignore' or 'ewald' or 'ewald_interaction'.")
        if reci_ei == 'ewald_interaction':
            raise NotImplementedError('ForcePartEwaldReciprocalInteraction is not part of this engine')
        self.rcut = rcut
        self.tr = Switch3(7.558904535685008) if tr == 'default' else tr
        self.alpha_scale = alpha_scale
        self.gcut_scale = gcut_scale
        self.skin = skin
        self.smooth_ei = smooth_ei
        self.reci_ei = reci_ei
        self.parts = []
        self.nlist = None
        self.nlow = nlow
        self.nhigh = nhigh
        self.tailcorrections = tailcorrections

    def get_nlist(self, system):
        from .nlist import NeighborList
        if self.nlist is None:
            self.nlist = NeighborList(system, skin=self.skin, nlow=self.nlow, nhigh=self.nhigh)
        return self.nlist

    def get_part(self, cls):
        for part in self.parts:
            if isinstance(part, cls):
                return part

    def get_part_pair(self, pot_cls):
        from .ff import ForcePartPair
        for part in self.parts:
            if isinstance(part, ForcePartPair) and isinstance(part.pair_pot, pot_cls):
                return part

    def add_electrostatic_parts(self, system, scalings, dielectric):
        """generator.py:192-241"""
        from .ff import ForcePartPair, ForcePartEwaldReciprocal, ForcePartEwaldCorrection, ForcePartEwaldNeutralizing
        from .pair_pot import PairPotEI
        if self.get_part_pair(PairPotEI) is not None:
            return
        nlist = self.get_nlist(system)
        if system.cell.nvec == 0:
            alpha = 0.0
        elif system.cell.nvec == 3:
            alpha = self.alpha_scale/self.rcut
        else:
            raise NotImplementedError('Only zero- and three-dimensional electrostatics are supported.')
        tr = self.tr if self.smooth_ei else None
        pot = PairPotEI(system.charges, alpha, self.rcut, tr, dielectric, system.radii)
        self.parts.append(ForcePartPair(system, nlist, scalings, pot))
        if self.reci_ei == 'ewald':
            if system.cell.nvec == 3:
                self.parts.append(ForcePartEwaldReciprocal(system, alpha, self.gcut_scale*alpha, dielectric,
                                                           self.nlow, self.nhigh))
                self.parts.append(ForcePartEwaldCorrection(system, alpha, scalings, dielectric, self.nlow,
                                                           self.nhigh))
                self.parts.append(ForcePartEwaldNeutralizing(system, alpha, dielectric, self.nlow, self.nhigh))
            elif system.cell.nvec != 0:
                raise NotImplementedError('The ewald summation is only available for 3D periodic systems.')


class Generator(object):
    """generator.py:244-410: one parameter-file section to force-field parts."""
    prefix = None
    par_info = None
    suffixes = None
    allow_superposition = False

    def __call__(self, system, parsec, ff_args):
        raise NotImplementedError

    def check_suffixes(self, parsec):
        for suffix in parsec.definitions:
            if suffix not in self.suffixes:
                parsec.complain(None, 'contains a suffix (%s) that is not recognized by generator %s.'
                                % (suffix, self.prefix))

    def process_units(self, pardef):
        """UNIT lines: {parameter name: conversion to atomic units}."""
        result = {}
        expected = [name for name, dtype in self.par_info if dtype is float]
        for counter, line in pardef:
            words = line.split()
            if len(words) != 2:
                pardef.complain(counter, 'must have two arguments in UNIT suffix')
            name = words[0].upper()
            if name not in expected:
                pardef.complain(counter, 'specifies a unit for an unknown parameter. (Must be one of %s, but got %s.)'
                                % (expected, name))
            try:
                result[name] = parse_unit(words[1])
            except (NameError, ValueError, ZeroDivisionError):
                pardef.complain(counter, 'has a UNIT suffix with an unknown unit')
        if len(result) != len(expected):
            raise IOError('Not all units are specified for generator %s in file %s. Got %s, should have %s.'
                          % (self.prefix, pardef.complain.filename, list(result.keys()), expected))
        return result

    def process_pars(self, pardef, conversions, nffatype, par_info=None):
        """PARS-like lines: {ffatype key: [parameter tuples]} with the
        permutations of iter_equiv_keys_and_pars."""
        if par_info is None:
            par_info = self.par_info
            allow_superposition = self.allow_superposition
        else:
            allow_superposition = False
        par_table = {}
        for counter, line in pardef:
            words = line.split()
            num_args = nffatype + len(par_info)
            if len(words) != num_args:
                pardef.complain(counter, 'should have %s arguments' % num_args)
            key = tuple(words[:nffatype])
            pars = []
            for i, (name, dtype) in enumerate(par_info):
                word = words[i + nffatype]
                try:
                    pars.append(float(word)*conversions[name] if dtype is float else dtype(word))
                except ValueError:
                    pardef.complain(counter, 'contains a parameter that can not be converted to a number: {}'.format(word))
            pars = tuple(pars)
            current = {}
            for alt_key, alt_pars in self.iter_equiv_keys_and_pars(key, pars):
                if alt_key in current:
                    if current[alt_key] != alt_pars:
                        pardef.complain(counter, 'contains parameters that are not consistent with the permutational '
                                                 'symmetry of the atom types')
                else:
                    current[alt_key] = alt_pars
            for alt_key, alt_pars in current.items():
                lst = par_table.setdefault(alt_key, [])
                if lst and not allow_superposition:
                    pardef.complain(counter, 'contains a duplicate energy term, possibly with different parameters, '
                                             'which is not allowed for generator %s' % self.prefix)
                lst.append(alt_pars)
        return par_table

    def iter_equiv_keys_and_pars(self, key, pars):
        if len(key) == 1:
            yield key, pars
        else:
            raise NotImplementedError


class NonbondedGenerator(Generator):
    """generator.py:1603-1676: SCALE and MIX definitions."""
    mixing_rules = None

    def process_scales(self, pardef):
        result = {}
        for counter, line in pardef:
            words = line.split()
            if len(words) != 2:
                pardef.complain(counter, 'must have 2 arguments')
            try:
                num_bonds = int(words[0])
                scale = float(words[1])
            except ValueError:
                pardef.complain(counter, 'has parameters that can not be converted. The first argument must be an '
                                         'integer. The second argument must be a float')
            if num_bonds in result and result[num_bonds] != scale:
                pardef.complain(counter, 'contains a duplicate incompatible scale suffix')
            if scale < 0 or scale > 1:
                pardef.complain(counter, 'has a scale that is not in the range [0,1]')
            result[num_bonds] = scale
        if len(result) < 3 or len(result) > 4:
            pardef.complain(None, 'must contain three or four SCALE suffixes for each non-bonding term')
        if 1 not in result or 2 not in result or 3 not in result:
            pardef.complain(None, 'must contain a scale parameter for atoms separated by 1, 2 and 3 bonds, for each '
                                  'non-bonding term')
        result.setdefault(4, 1.0)
        return result

    def process_mix(self, pardef):
        result = {}
        for counter, line in pardef:
            words = line.split()
            if len(words) < 2:
                pardef.complain(counter, 'contains a mixing rule with to few arguments. At least 2 are required')
            key = words[0].upper(), words[1].upper()
            if key not in self.mixing_rules:
                pardef.complain(counter, 'contains an unknown mixing rule')
            narg, rule_id = self.mixing_rules[key]
            if len(words) != narg + 2:
                pardef.complain(counter, 'does not have the correct number of arguments. %i arguments are required'
                                % (narg + 2))
            try:
                args = tuple(float(w) for w in words[2:])
            except ValueError:
                pardef.complain(counter, 'contains parameters that could not be converted to floating point numbers')
            result[key[0]] = rule_id, args
        if len(result) != len(set(name for name, rule in self.mixing_rules)):
            pardef.complain(None, 'does not contain enough mixing rules for the generator %s' % self.prefix)
        return result

    @staticmethod
    def scalings(system, scale_table):
        from .scaling import Scalings
        return Scalings(system, scale_table[1], scale_table[2], scale_table[3], scale_table[4])

    def add_pair(self, system, ff_args, pot_cls, pot, scale_table):
        from .ff import ForcePartPair
        if ff_args.get_part_pair(pot_cls) is not None:
            raise RuntimeError('Internal inconsistency: the %s part should not be present yet.' % self.prefix)
        ff_args.parts.append(ForcePartPair(system, ff_args.get_nlist(system), self.scalings(system, scale_table), pot))

    @staticmethod
    def atom_pars(system, par_table, natom_pars):
        """Per-atom parameter columns from a one-type table (zeros for atoms
        of types without parameters)."""
        cols = [np.zeros(system.natom) for _ in range(natom_pars)]
        for i in range(system.natom):
            lst = par_table.get((system.get_ffatype(i),), [])
            if len(lst) > 2:
                raise TypeError('Superposition should not be allowed for non-covalent terms.')
            elif len(lst) == 1:
                for c, v in zip(cols, lst[0]):
                    c[i] = v
        return cols


class LJGenerator(NonbondedGenerator):
    """generator.py:1679-1713"""
    prefix = 'LJ'
    suffixes = ['UNIT', 'SCALE', 'PARS']
    par_info = [('SIGMA', float), ('EPSILON', float)]

    def __call__(self, system, parsec, ff_args):
        from .pair_pot import PairPotLJ
        self.check_suffixes(parsec)
        conv = self.process_units(parsec['UNIT'])
        par_table = self.process_pars(parsec['PARS'], conv, 1)
        scale_table = self.process_scales(parsec['SCALE'])
        sigmas, epsilons = self.atom_pars(system, par_table, 2)
        self.add_pair(system, ff_args, PairPotLJ, PairPotLJ(sigmas, epsilons, ff_args.rcut, ff_args.tr), scale_table)


class MM3Generator(NonbondedGenerator):
    """generator.py:1760-1794"""
    prefix = 'MM3'
    suffixes = ['UNIT', 'SCALE', 'PARS']
    par_info = [('SIGMA', float), ('EPSILON', float), ('ONLYPAULI', int)]

    def __call__(self, system, parsec, ff_args):
        from .pair_pot import PairPotMM3
        self.check_suffixes(parsec)
        conv = self.process_units(parsec['UNIT'])
        par_table = self.process_pars(parsec['PARS'], conv, 1)
        scale_table = self.process_scales(parsec['SCALE'])
        sigmas, epsilons, ops = self.atom_pars(system, par_table, 3)
        pot = PairPotMM3(sigmas, epsilons, ops.astype(np.int32), ff_args.rcut, ff_args.tr)
        self.add_pair(system, ff_args, PairPotMM3, pot, scale_table)


class ExpRepGenerator(NonbondedGenerator):
    """generator.py:1831-1916"""
    prefix = 'EXPREP'
    suffixes = ['UNIT', 'SCALE', 'MIX', 'PARS', 'CPARS']
    par_info = [('A', float), ('B', float)]
    mixing_rules = {('A', 'GEOMETRIC'): (0, 0), ('A', 'GEOMETRIC_COR'): (1, 1),
                    ('B', 'ARITHMETIC'): (0, 0), ('B', 'ARITHMETIC_COR'): (1, 1)}

    def iter_equiv_keys_and_pars(self, key, pars):
        yield key, pars
        yield key[::-1], pars

    def __call__(self, system, parsec, ff_args):
        from .pair_pot import PairPotExpRep
        self.check_suffixes(parsec)
        conv = self.process_units(parsec['UNIT'])
        par_table = self.process_pars(parsec['PARS'], conv, 1)
        cpar_table = self.process_pars(parsec['CPARS'], conv, 2)
        scale_table = self.process_scales(parsec['SCALE'])
        rules = self.process_mix(parsec['MIX'])
        nff = system.nffatype
        amps, bs = np.zeros(nff), np.zeros(nff)
        for i in range(nff):
            lst = par_table.get((system.ffatypes[i],), [])
            if len(lst) > 2:
                raise TypeError('Superposition should not be allowed for non-covalent terms.')
            elif len(lst) == 1:
                amps[i], bs[i] = lst[0]
        amp_cross, b_cross = np.zeros((nff, nff)), np.zeros((nff, nff))
        for i0 in range(nff):
            for i1 in range(i0 + 1):
                lst = cpar_table.get((system.ffatypes[i0], system.ffatypes[i1]), [])
                if lst:
                    amp_cross[i0, i1], b_cross[i0, i1] = lst[0]
                    amp_cross[i1, i0], b_cross[i1, i0] = lst[0]
        amp_mix, amp_args = rules['A']
        amp_coeff = amp_args[0] if amp_mix == 1 else 0.0
        b_mix, b_args = rules['B']
        b_coeff = b_args[0] if b_mix == 1 else 0.0
        pot = PairPotExpRep(system.ffatype_ids, amp_cross, b_cross, ff_args.rcut, ff_args.tr, amps, amp_mix,
                            amp_coeff, bs, b_mix, b_coeff)
        self.add_pair(system, ff_args, PairPotExpRep, pot, scale_table)


class FixedChargeGenerator(NonbondedGenerator):
    """generator.py:2077-2173"""
    prefix = 'FIXQ'
    suffixes = ['UNIT', 'SCALE', 'ATOM', 'BOND', 'DIELECTRIC']
    par_info = [('Q0', float), ('P', float), ('R', float)]

    def __call__(self, system, parsec, ff_args):
        self.check_suffixes(parsec)
        conv = self.process_units(parsec['UNIT'])
        atoms = self.process_atoms(parsec['ATOM'], conv)
        bonds = self.process_bonds(parsec['BOND'], conv)
        scale_table = self.process_scales(parsec['SCALE'])
        dielectric = self.process_dielectric(parsec['DIELECTRIC'])
        self.apply(atoms, bonds, scale_table, dielectric, system, ff_args)

    def process_atoms(self, pardef, conv):
        result = {}
        for counter, line in pardef:
            words = line.split()
            if len(words) != 3:
                pardef.complain(counter, 'should have 3 arguments')
            if words[0] in result:
                pardef.complain(counter, 'has an atom type that was already encountered earlier')
            try:
                result[words[0]] = float(words[1])*conv['Q0'], float(words[2])*conv['R']
            except ValueError:
                pardef.complain(counter, 'contains a parameter that can not be converted to a floating point number')
        return result

    def process_bonds(self, pardef, conv):
        result = {}
        for counter, line in pardef:
            words = line.split()
            if len(words) != 3:
                pardef.complain(counter, 'should have 3 arguments')
            key = tuple(words[:2])
            if key in result:
                pardef.complain(counter, 'has a combination of atom types that were already encountered earlier')
            try:
                p = float(words[2])*conv['P']
            except ValueError:
                pardef.complain(counter, 'contains a parameter that can not be converted to floating point numbers')
            result[key] = p
            result[key[::-1]] = -p
        return result

    def process_dielectric(self, pardef):
        result = None
        for counter, line in pardef:
            if result is not None:
                pardef.complain(counter, 'is redundant. The DIELECTRIC suffix may only occur once')
            words = line.split()
            if len(words) != 1:
                pardef.complain(counter, 'must have one argument')
            try:
                result = float(words[0])
            except ValueError:
                pardef.complain(counter, 'must have a floating point argument')
        return result

    def apply(self, atoms, bonds, scale_table, dielectric, system, ff_args):
        system.charges = np.zeros(system.natom)
        system.radii = np.zeros(system.natom)
        for i in range(system.natom):
            pars = atoms.get(system.get_ffatype(i))
            if pars is not None:
                system.charges[i] += pars[0]
                system.radii[i] = pars[1]
        for i0, i1 in system.iter_bonds():
            t0, t1 = system.get_ffatype(i0), system.get_ffatype(i1)
            if t0 == t1:
                continue
            p = bonds.get((t0, t1))
            if p is not None:
                system.charges[i0] += p
                system.charges[i1] -= p
        ff_args.add_electrostatic_parts(system, self.scalings(system, scale_table), dielectric)


GENERATORS = {g.prefix: g for g in (LJGenerator, MM3Generator, ExpRepGenerator, FixedChargeGenerator)}


def apply_generators(system, parameters, ff_args, skip_other_sections=False):
    """generator.py:2176-2246 for the nonbonded generators of this engine."""
    from .ff import ForcePartPair, ForcePartTailCorrection
    from .pair_pot import PairPotEI
    for prefix, section in parameters.sections.items():
        gen = GENERATORS.get(prefix)
        if gen is not None:
            gen()(system, section, ff_args)
        elif prefix in YAFF_OTHER_PREFIXES:
            if not skip_other_sections:
                raise NotImplementedError('Section %s of the parameter file is a term this engine does not build '
                                          '(pass skip_other_sections=True to leave it out)' % prefix)
            warnings.warn('Section %s of the parameter file is left out.' % prefix, RuntimeWarning)
        else:
            warnings.warn('There is no generator named %s. It will be ignored.' % prefix, RuntimeWarning)
    if ff_args.tailcorrections:
        if system.cell.nvec == 0:
            warnings.warn('Tail corrections were requested, but this makes no sense for non-periodic system. '
                          'Not adding tail corrections...', RuntimeWarning)
        elif system.cell.nvec == 3:
            for part in list(ff_args.parts):
                if isinstance(part, ForcePartPair) and not isinstance(part.pair_pot, PairPotEI):
                    ff_args.parts.append(ForcePartTailCorrection(system, part, nlow=ff_args.nlow,
                                                                 nhigh=ff_args.nhigh))
        else:
            raise ValueError('Tail corrections not available for 1-D and 2-D periodic systems')
