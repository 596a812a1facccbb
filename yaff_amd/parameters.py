"""Parameter files in Yaff's format (read as yaff/pes/parameters.py reads them).

The format is line based: ``PREFIX:SUFFIX DATA``; ``#`` starts a comment
that runs to the end of the line; blank lines and the order of lines do not
matter.  A prefix selects a section (one generator), a suffix a definition
inside it; prefixes and suffixes are case-insensitive (kept upper case), the
data are kept as written.  Every data line remembers its file and line number
so that the generators can point at the offending line.

Reading is one pass: each line is matched against the grammar below and the
first rule it breaks is reported as an ``IOError`` (the exception type of the
reference) naming the file and the line.
"""
import collections
import re

__all__ = ['Complain', 'Parameters', 'ParameterSection', 'ParameterDefinition']

# PREFIX ':' SUFFIX whitespace DATA, with optional blanks around the colon
_LINE = re.compile(r'^(?P<prefix>[^:]*):\s*(?P<suffix>\S*)(?P<sep>\s*)(?P<data>.*)$')


class Complain(object):
    """Raises IOError for a problem in a parameter file; ``where`` is a line
    number or None (a problem of a whole section or definition)."""

    def __init__(self, filename='__nofile__'):
        self.filename = filename

    def __call__(self, where, message):
        place = 'parameter file %s' % self.filename if where is None else \
            'parameter file %s, line %i' % (self.filename, where)
        raise IOError('%s: %s' % (place, message))


def _parse_line(text, number, complain):
    """(PREFIX, SUFFIX, data) of one non-empty, comment-stripped line."""
    m = _LINE.match(text)
    if m is None:
        complain(number, 'a line needs the form PREFIX:SUFFIX DATA (no colon found)')
    prefix = m.group('prefix').strip()
    if not prefix or len(prefix.split()) != 1:
        complain(number, 'the prefix before the colon must be a single word')
    if not m.group('suffix'):
        complain(number, 'nothing follows the colon')
    if not m.group('data'):
        complain(number, 'the definition %s has no data' % m.group('suffix'))
    return prefix.upper(), m.group('suffix').upper(), m.group('data')


class ParameterDefinition(object):
    """The data lines of one PREFIX:SUFFIX, as (line number, data) pairs in
    file order; iterating yields those pairs."""

    def __init__(self, suffix, lines=None, complain=None):
        self.suffix = suffix
        self.lines = [] if lines is None else lines
        self.complain = Complain() if complain is None else complain

    def __iter__(self):
        return iter(self.lines)

    def __getitem__(self, index):
        return self.lines[index]

    def __len__(self):
        return len(self.lines)

    def copy(self):
        return ParameterDefinition(self.suffix, list(self.lines), self.complain)


class ParameterSection(object):
    """The definitions of one prefix.  Looking up a suffix that the file
    does not have gives an empty definition."""

    def __init__(self, prefix, definitions=None, complain=None):
        self.prefix = prefix
        self.definitions = {} if definitions is None else definitions
        self.complain = Complain() if complain is None else complain

    def __getitem__(self, suffix):
        found = self.definitions.get(suffix.upper())
        return found if found is not None else ParameterDefinition(suffix.upper(), [], self.complain)

    def __contains__(self, suffix):
        return suffix.upper() in self.definitions

    def copy(self):
        return ParameterSection(self.prefix, {k: v.copy() for k, v in self.definitions.items()}, self.complain)


class Parameters(object):
    """All sections of one or more parameter files (later files add lines to
    the sections of earlier ones)."""

    def __init__(self, sections=None):
        self.sections = {} if sections is None else sections

    @classmethod
    def from_file(cls, filenames):
        if isinstance(filenames, str):
            filenames = [filenames]
        grouped = collections.OrderedDict()
        complains = {}
        for filename in filenames:
            complain = Complain(filename)
            with open(filename) as f:
                for number, raw in enumerate(f, 1):
                    text = raw.split('#', 1)[0].strip()
                    if not text:
                        continue
                    prefix, suffix, data = _parse_line(text, number, complain)
                    complains.setdefault(prefix, complain)
                    grouped.setdefault(prefix, collections.OrderedDict()).setdefault(suffix, []).append(
                        (number, data))
        sections = {}
        for prefix, defs in grouped.items():
            complain = complains[prefix]
            sections[prefix] = ParameterSection(
                prefix, {sfx: ParameterDefinition(sfx, lines, complain) for sfx, lines in defs.items()}, complain)
        return cls(sections)

    def __getitem__(self, prefix):
        found = self.sections.get(prefix.upper())
        return found if found is not None else ParameterSection(prefix.upper())

    def __contains__(self, prefix):
        return prefix.upper() in self.sections

    def copy(self):
        return Parameters({k: v.copy() for k, v in self.sections.items()})

    def write_to_file(self, filename):
        """Writes every data line back as PREFIX:SUFFIX DATA, section by
        section (line numbers are not kept)."""
        out = []
        for prefix, section in self.sections.items():
            for suffix, definition in section.definitions.items():
                out.extend('%s:%s %s' % (prefix, suffix, data) for _, data in definition)
            out.append('')
        with open(filename, 'w') as f:
            f.write('\n'.join(out) + '\n')
