"""Force fields from Yaff parameter files, for the nonbonded sections that
this engine evaluates: LJ, MM3, EXPREP and FIXQ (the file format and the
force fields they produce are those of yaff/pes/generator.py:1603-1916 and
2077-2246; the keyword arguments and defaults those of FFArgs,
generator.py:84-191).

Design: every supported section is described by a ``SectionSpec`` -- which
definitions (suffixes) it may hold, the per-type parameter columns of its
PARS lines, the pair columns of its CPARS lines, its mixing rules.  One
reader (``read_section``) validates a section against its spec and returns
plain tables (unit factors, per-type rows, per-pair rows, scale factors,
mixing rules, atom charges, bond charge increments, dielectric); a builder
per prefix turns the tables into parts.  ``ForceField.generate`` runs
``apply_generators``.

Sections of Yaff's other generators (valence terms and the like) are not
built here: by default they raise NotImplementedError; with
``skip_other_sections=True`` each is left out with a RuntimeWarning, so the
caller can add Yaff's own parts for them.  Format errors raise IOError with
the file and line, as Yaff's parameter reader does.
"""
import warnings

import numpy as np

from .units import parse_unit

__all__ = ['FFArgs', 'SectionSpec', 'SPECS', 'read_section', 'Generator', 'NonbondedGenerator', 'LJGenerator',
           'MM3Generator', 'ExpRepGenerator', 'FixedChargeGenerator', 'apply_generators', 'YAFF_OTHER_PREFIXES']

# Sections Yaff's generators know that this engine does not build.
YAFF_OTHER_PREFIXES = frozenset('''
    BONDHARM BONDFUES MM3QUART DOUBWELL BONDMORSE DOUBWELL2 POLY4 BONDPOLYSIX BENDAHARM BENDCHARM MM3BENDA
    UBHARM BENDCOS BENDCLIN TORSAHARM TORSCHARM TORSION TORSC2HARM TORSCPOLYSIX OOPANGLE OOPMANGLE OOPCOS
    OOPMCOS OOPDIST SQOOPDIST IMPROPER CROSS CROSSCBEND CROSSBONDDIHED CROSSBENDDIHED LJCROSS MM3CAP
    DAMPDISP D3BJ QMDFFREP BONDCROSS BENDCROSS'''.split())

# 10 A and 4 A in bohr, the defaults of FFArgs (generator.py:90-91)
DEFAULT_RCUT = 18.89726133921252
DEFAULT_SWITCH_WIDTH = 7.558904535685008


class FFArgs(object):
    """The keyword arguments of ForceField.generate and the force field
    under construction (its parts and the shared neighbour list).

    rcut (default 10 A), tr (default Switch3 of 4 A; None: no truncation),
    alpha_scale and gcut_scale (Ewald alpha = alpha_scale / rcut and
    gcut = gcut_scale * alpha for 3-D cells), skin (neighbour list),
    smooth_ei (truncate the real-space EI term too), reci_ei ('ewald' or
    'ignore'), nlow/nhigh (atom groups), tailcorrections."""

    RECI_CHOICES = ('ignore', 'ewald', 'ewald_interaction')

    def __init__(self, rcut=DEFAULT_RCUT, tr='default', alpha_scale=3.5, gcut_scale=1.1, skin=0, smooth_ei=False,
                 reci_ei='ewald', nlow=0, nhigh=-1, tailcorrections=False):
        from .pair_pot import Switch3
        if reci_ei not in self.RECI_CHOICES:
            raise ValueError('reci_ei must be one of %s, got %r' % (', '.join(self.RECI_CHOICES), reci_ei))
        if reci_ei == 'ewald_interaction':
            raise NotImplementedError("reci_ei='ewald_interaction' (ForcePartEwaldReciprocalInteraction) is not "
                                      "provided by this engine")
        self.rcut = rcut
        self.tr = Switch3(DEFAULT_SWITCH_WIDTH) if tr == 'default' else tr
        self.alpha_scale, self.gcut_scale = alpha_scale, gcut_scale
        self.skin = skin
        self.smooth_ei = smooth_ei
        self.reci_ei = reci_ei
        self.nlow, self.nhigh = nlow, nhigh
        self.tailcorrections = tailcorrections
        self.parts = []
        self.nlist = None

    def get_nlist(self, system):
        """The one neighbour list every pair part of the force field shares."""
        if self.nlist is None:
            from .nlist import NeighborList
            self.nlist = NeighborList(system, skin=self.skin, nlow=self.nlow, nhigh=self.nhigh)
        return self.nlist

    def get_part(self, cls):
        return next((p for p in self.parts if isinstance(p, cls)), None)

    def get_part_pair(self, pot_cls):
        from .ff import ForcePartPair
        return next((p for p in self.parts if isinstance(p, ForcePartPair) and isinstance(p.pair_pot, pot_cls)),
                    None)

    def add_pair_part(self, system, pair_pot, scalings):
        from .ff import ForcePartPair
        if self.get_part_pair(type(pair_pot)) is not None:
            raise RuntimeError('a %s pair part is already in the force field' % pair_pot.name)
        self.parts.append(ForcePartPair(system, self.get_nlist(system), scalings, pair_pot))

    def add_electrostatic_parts(self, system, scalings, dielectric):
        """Real-space EI for the system's charges (and radii), plus, for a
        3-D cell with reci_ei='ewald', the Ewald reciprocal, correction and
        neutralizing parts; alpha = alpha_scale / rcut (0 without a cell).
        Nothing happens when an EI part is there already."""
        from .ff import ForcePartEwaldReciprocal, ForcePartEwaldCorrection, ForcePartEwaldNeutralizing
        from .pair_pot import PairPotEI
        if self.get_part_pair(PairPotEI) is not None:
            return
        nvec = system.cell.nvec
        if nvec not in (0, 3):
            raise NotImplementedError('electrostatics are available for 0-D and 3-D periodic systems only '
                                      '(this cell has %d periodic directions)' % nvec)
        alpha = self.alpha_scale/self.rcut if nvec == 3 else 0.0
        ei = PairPotEI(system.charges, alpha, self.rcut, self.tr if self.smooth_ei else None, dielectric,
                       system.radii)
        self.add_pair_part(system, ei, scalings)
        if self.reci_ei == 'ewald' and nvec == 3:
            groups = dict(nlow=self.nlow, nhigh=self.nhigh)
            self.parts.append(ForcePartEwaldReciprocal(system, alpha, self.gcut_scale*alpha, dielectric, **groups))
            self.parts.append(ForcePartEwaldCorrection(system, alpha, scalings, dielectric, **groups))
            self.parts.append(ForcePartEwaldNeutralizing(system, alpha, dielectric, **groups))


# ---------------------------------------------------------------------------
# Section specifications and the one reader

class SectionSpec(object):
    """What a section may contain.

    columns:  (name, type) of the parameters on a PARS line after the one
              atom type (float columns need a UNIT line each);
    pair_columns: the same for CPARS lines after two atom types;
    mixing:   {parameter: {rule name: (number of arguments, rule id)}};
    charges:  ATOM/BOND/DIELECTRIC definitions (FIXQ)."""

    def __init__(self, prefix, columns=(), pair_columns=(), mixing=None, charges=False):
        self.prefix = prefix
        self.columns = tuple(columns)
        self.pair_columns = tuple(pair_columns)
        self.mixing = mixing or {}
        self.charges = charges
        allowed = {'UNIT', 'SCALE'}
        if columns and not charges:
            allowed.add('PARS')
        if pair_columns:
            allowed.add('CPARS')
        if mixing:
            allowed.add('MIX')
        if charges:
            allowed |= {'ATOM', 'BOND', 'DIELECTRIC'}
        self.suffixes = frozenset(allowed)

    @property
    def unit_names(self):
        names = [n for n, t in self.columns + self.pair_columns if t is float]
        return list(dict.fromkeys(names))


SPECS = {
    'LJ': SectionSpec('LJ', columns=[('SIGMA', float), ('EPSILON', float)]),
    'MM3': SectionSpec('MM3', columns=[('SIGMA', float), ('EPSILON', float), ('ONLYPAULI', int)]),
    'EXPREP': SectionSpec('EXPREP', columns=[('A', float), ('B', float)], pair_columns=[('A', float), ('B', float)],
                          mixing={'A': {'GEOMETRIC': (0, 0), 'GEOMETRIC_COR': (1, 1)},
                                  'B': {'ARITHMETIC': (0, 0), 'ARITHMETIC_COR': (1, 1)}}),
    # Q0 (charge of an atom type), P (charge moved along a bond), R (Gaussian radius)
    'FIXQ': SectionSpec('FIXQ', columns=[('Q0', float), ('P', float), ('R', float)], charges=True),
}


class SectionData(object):
    """The tables read from one section (fields unused by a spec stay empty)."""

    def __init__(self):
        self.units = {}       # parameter name -> factor to atomic units
        self.pars = {}        # (type,) -> tuple of parameters
        self.cpars = {}       # (type0, type1) -> tuple of parameters, both orders
        self.scales = {}      # number of bonds (1..4) -> scale factor
        self.mixing = {}      # parameter -> (rule id, arguments)
        self.atoms = {}       # type -> (charge, radius)
        self.bonds = {}       # (type0, type1) -> charge moved from type1 to type0 (both orders, opposite signs)
        self.dielectric = None


def _fields(definition, what, nwords=None, at_least=None):
    """(line number, words) of each line of a definition, with the word
    count checked."""
    for number, data in definition:
        words = data.split()
        if nwords is not None and len(words) != nwords:
            definition.complain(number, '%s needs %d fields, found %d' % (what, nwords, len(words)))
        if at_least is not None and len(words) < at_least:
            definition.complain(number, '%s needs at least %d fields, found %d' % (what, at_least, len(words)))
        yield number, words


def _number(definition, number, word, kind, what):
    try:
        return kind(word)
    except ValueError:
        definition.complain(number, '%s: %r is not %s' % (what, word, 'an integer' if kind is int else 'a number'))


def _read_units(spec, definition):
    names = spec.unit_names
    units = {}
    for number, (name, unit) in _fields(definition, 'UNIT', nwords=2):
        name = name.upper()
        if name not in names:
            definition.complain(number, 'UNIT of unknown parameter %s (expected one of %s)'
                                % (name, ', '.join(names)))
        try:
            units[name] = parse_unit(unit)
        except (NameError, ValueError, ZeroDivisionError):
            definition.complain(number, 'UNIT %s: cannot interpret the unit %r' % (name, unit))
    missing = [n for n in names if n not in units]
    if missing:
        raise IOError('parameter file %s: section %s has no UNIT for %s'
                      % (definition.complain.filename, spec.prefix, ', '.join(missing)))
    return units


def _read_rows(spec, definition, ntype, columns, units, what):
    """Rows of ntype atom types plus parameter columns; a pair of types is
    stored under both orders, which must then agree.  One row per key: the
    nonbonded terms allow no superposition."""
    table = {}
    for number, words in _fields(definition, what, nwords=ntype + len(columns)):
        key = tuple(words[:ntype])
        row = tuple(_number(definition, number, w, kind, what)*(units[name] if kind is float else 1)
                    for w, (name, kind) in zip(words[ntype:], columns))
        keys = {key, key[::-1]}
        for k in keys:
            if k in table:
                definition.complain(number, '%s for %s is given twice (superposition is not allowed in a '
                                            'nonbonded section)' % (what, ' '.join(k)))
        for k in keys:
            table[k] = row
    return table


def _read_scales(spec, definition):
    scales = {}
    for number, words in _fields(definition, 'SCALE', nwords=2):
        nbond = _number(definition, number, words[0], int, 'SCALE bond count')
        factor = _number(definition, number, words[1], float, 'SCALE factor')
        if not 0.0 <= factor <= 1.0:
            definition.complain(number, 'SCALE factor %g lies outside [0, 1]' % factor)
        if scales.get(nbond, factor) != factor:
            definition.complain(number, 'SCALE for %d bonds is given twice with different factors' % nbond)
        scales[nbond] = factor
    if not {1, 2, 3} <= set(scales) or not set(scales) <= {1, 2, 3, 4}:
        definition.complain(None, 'section %s needs SCALE factors for 1, 2 and 3 bonds (and optionally 4), got '
                                  'bond counts %s' % (spec.prefix, sorted(scales)))
    scales.setdefault(4, 1.0)
    return scales


def _read_mixing(spec, definition):
    rules = {}
    for number, words in _fields(definition, 'MIX', at_least=2):
        par, rule = words[0].upper(), words[1].upper()
        if rule not in spec.mixing.get(par, {}):
            definition.complain(number, 'MIX %s %s is no known mixing rule' % (par, rule))
        nargs, rule_id = spec.mixing[par][rule]
        if len(words) != 2 + nargs:
            definition.complain(number, 'MIX %s %s takes %d argument(s)' % (par, rule, nargs))
        rules[par] = rule_id, tuple(_number(definition, number, w, float, 'MIX argument') for w in words[2:])
    if set(rules) != set(spec.mixing):
        definition.complain(None, 'section %s needs one MIX rule for each of %s'
                                  % (spec.prefix, ', '.join(sorted(spec.mixing))))
    return rules


def _read_charges(spec, section, units, data):
    for number, (ffatype, q, r) in _fields(section['ATOM'], 'ATOM', nwords=3):
        if ffatype in data.atoms:
            section['ATOM'].complain(number, 'ATOM for type %s is given twice' % ffatype)
        data.atoms[ffatype] = (_number(section['ATOM'], number, q, float, 'ATOM charge')*units['Q0'],
                               _number(section['ATOM'], number, r, float, 'ATOM radius')*units['R'])
    for number, (t0, t1, p) in _fields(section['BOND'], 'BOND', nwords=3):
        if (t0, t1) in data.bonds:
            section['BOND'].complain(number, 'BOND for types %s %s is given twice' % (t0, t1))
        moved = _number(section['BOND'], number, p, float, 'BOND charge increment')*units['P']
        data.bonds[(t0, t1)] = moved
        data.bonds[(t1, t0)] = -moved
    for number, words in _fields(section['DIELECTRIC'], 'DIELECTRIC', nwords=1):
        if data.dielectric is not None:
            section['DIELECTRIC'].complain(number, 'DIELECTRIC may be given only once')
        data.dielectric = _number(section['DIELECTRIC'], number, words[0], float, 'DIELECTRIC')


def read_section(spec, section):
    """Validates a ParameterSection against its spec and returns its
    SectionData."""
    extra = sorted(set(section.definitions) - spec.suffixes)
    if extra:
        section.complain(None, 'section %s does not accept the definition(s) %s (allowed: %s)'
                               % (spec.prefix, ', '.join(extra), ', '.join(sorted(spec.suffixes))))
    data = SectionData()
    data.units = _read_units(spec, section['UNIT'])
    if spec.charges:
        _read_charges(spec, section, data.units, data)
    else:
        data.pars = _read_rows(spec, section['PARS'], 1, spec.columns, data.units, 'PARS')
    if spec.pair_columns:
        data.cpars = _read_rows(spec, section['CPARS'], 2, spec.pair_columns, data.units, 'CPARS')
    if spec.mixing:
        data.mixing = _read_mixing(spec, section['MIX'])
    data.scales = _read_scales(spec, section['SCALE'])
    return data


# ---------------------------------------------------------------------------
# Builders: tables -> parts

def _scalings(system, scales):
    from .scaling import Scalings
    return Scalings(system, scales[1], scales[2], scales[3], scales[4])


def _per_atom(system, pars, ncol):
    """Parameter columns per atom from a one-type table (0 for atoms whose
    type has no PARS line)."""
    cols = np.zeros((ncol, system.natom))
    for i in range(system.natom):
        row = pars.get((system.get_ffatype(i),))
        if row is not None:
            cols[:, i] = row
    return cols


def _build_lj(system, data, ff_args):
    from .pair_pot import PairPotLJ
    sigmas, epsilons = _per_atom(system, data.pars, 2)
    ff_args.add_pair_part(system, PairPotLJ(sigmas, epsilons, ff_args.rcut, ff_args.tr),
                          _scalings(system, data.scales))


def _build_mm3(system, data, ff_args):
    from .pair_pot import PairPotMM3
    sigmas, epsilons, paulis = _per_atom(system, data.pars, 3)
    ff_args.add_pair_part(system, PairPotMM3(sigmas, epsilons, paulis.astype(np.int32), ff_args.rcut, ff_args.tr),
                          _scalings(system, data.scales))


def _build_exprep(system, data, ff_args):
    """Per-type A and B plus explicit pair values; pairs without CPARS are
    mixed by PairPotExpRep with the MIX rules (ext.pyx:1146-1172)."""
    from .pair_pot import PairPotExpRep
    types = list(system.ffatypes)
    nff = len(types)
    single = np.zeros((2, nff))
    for t, name in enumerate(types):
        if (name,) in data.pars:
            single[:, t] = data.pars[(name,)]
    cross = np.zeros((2, nff, nff))
    for t0 in range(nff):
        for t1 in range(nff):
            if (types[t0], types[t1]) in data.cpars:
                cross[:, t0, t1] = data.cpars[(types[t0], types[t1])]
    (amp_rule, amp_args), (b_rule, b_args) = data.mixing['A'], data.mixing['B']
    pot = PairPotExpRep(system.ffatype_ids, cross[0], cross[1], ff_args.rcut, ff_args.tr,
                        single[0], amp_rule, amp_args[0] if amp_args else 0.0,
                        single[1], b_rule, b_args[0] if b_args else 0.0)
    ff_args.add_pair_part(system, pot, _scalings(system, data.scales))


def _build_fixq(system, data, ff_args):
    """Charges = the type charge plus the bond charge increments (bonds
    between atoms of one type move nothing), radii per type; then the
    electrostatic parts."""
    if data.dielectric is None:
        raise TypeError('section FIXQ has no DIELECTRIC definition (a relative permittivity is required)')
    if system.charges is not None and np.any(system.charges != 0):
        warnings.warn('FIXQ replaces the charges of the system', RuntimeWarning)
    charges = np.zeros(system.natom)
    radii = np.zeros(system.natom)
    for i in range(system.natom):
        q, r = data.atoms.get(system.get_ffatype(i), (0.0, 0.0))
        charges[i] += q
        radii[i] = r
    for i0, i1 in system.iter_bonds():
        key = system.get_ffatype(i0), system.get_ffatype(i1)
        if key[0] != key[1] and key in data.bonds:
            charges[i0] += data.bonds[key]
            charges[i1] -= data.bonds[key]
    system.charges = charges
    system.radii = radii
    ff_args.add_electrostatic_parts(system, _scalings(system, data.scales), data.dielectric)


BUILDERS = {'LJ': _build_lj, 'MM3': _build_mm3, 'EXPREP': _build_exprep, 'FIXQ': _build_fixq}


# ---------------------------------------------------------------------------
# Generator objects (Yaff's public names) and the driver

class Generator(object):
    """A callable that adds the parts of one section to the force field:
    ``generator(system, section, ff_args)``."""
    prefix = None

    @property
    def spec(self):
        return SPECS[self.prefix]

    @property
    def suffixes(self):
        return sorted(self.spec.suffixes)

    @property
    def par_info(self):
        return list(self.spec.columns)

    def __call__(self, system, parsec, ff_args):
        BUILDERS[self.prefix](system, read_section(self.spec, parsec), ff_args)


class NonbondedGenerator(Generator):
    """Base of the pair-potential generators (SCALE factors, MIX rules)."""


class LJGenerator(NonbondedGenerator):
    prefix = 'LJ'


class MM3Generator(NonbondedGenerator):
    prefix = 'MM3'


class ExpRepGenerator(NonbondedGenerator):
    prefix = 'EXPREP'


class FixedChargeGenerator(NonbondedGenerator):
    prefix = 'FIXQ'


GENERATORS = {g.prefix: g for g in (LJGenerator, MM3Generator, ExpRepGenerator, FixedChargeGenerator)}


def apply_generators(system, parameters, ff_args, skip_other_sections=False):
    """Runs the generator of every section of ``parameters`` (sections of
    Yaff's other generators: NotImplementedError, or a RuntimeWarning and
    nothing with skip_other_sections; unknown prefixes: a RuntimeWarning),
    then adds tail corrections for the non-EI pair parts when
    ff_args.tailcorrections (3-D cells; a warning without a cell)."""
    from .ff import ForcePartPair, ForcePartTailCorrection
    from .pair_pot import PairPotEI
    for prefix, section in parameters.sections.items():
        if prefix in GENERATORS:
            GENERATORS[prefix]()(system, section, ff_args)
        elif prefix in YAFF_OTHER_PREFIXES:
            if not skip_other_sections:
                raise NotImplementedError('section %s of the parameter file is a term this engine does not build; '
                                          'pass skip_other_sections=True to leave it to Yaff' % prefix)
            warnings.warn('Section %s is left out (not a nonbonded term of this engine).' % prefix, RuntimeWarning)
        else:
            warnings.warn('Section %s has no generator and is ignored.' % prefix, RuntimeWarning)
    if not ff_args.tailcorrections:
        return
    nvec = system.cell.nvec
    if nvec == 0:
        warnings.warn('tail corrections need a periodic system; none are added', RuntimeWarning)
        return
    if nvec != 3:
        raise ValueError('tail corrections are defined for 3-D periodic systems only')
    for part in list(ff_args.parts):
        if isinstance(part, ForcePartPair) and not isinstance(part.pair_pot, PairPotEI):
            ff_args.parts.append(ForcePartTailCorrection(system, part, nlow=ff_args.nlow, nhigh=ff_args.nhigh))
