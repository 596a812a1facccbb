"""Yaff parameter files (yaff/pes/parameters.py:35-206).

Line-based: ``PREFIX:SUFFIX DATA``, comments from ``#`` to the end of the
line, blank lines ignored, the order of lines free.  Prefixes name sections
(one generator each), suffixes name definitions; a definition keeps its data
lines with their line numbers so errors can name them.
"""
__all__ = ['Complain', 'Parameters', 'ParameterSection', 'ParameterDefinition']


class Complain(object):
    """parameters.py:35-43: raises IOError naming the file and line."""

    def __init__(self, filename='__nofile__'):
        self.filename = filename

    def __call__(self, counter, message):
        if counter is None:
            raise IOError('The parameter file %s %s.' % (self.filename, message))
        raise IOError('Line %i in the parameter file %s %s.' % (counter, self.filename, message))


class Parameters(object):
    """parameters.py:46-162"""

    def __init__(self, sections=None):
        self.sections = {} if sections is None else sections

    @classmethod
    def from_file(cls, filenames):
        if isinstance(filenames, str):
            filenames = [filenames]
        result = cls({})
        for filename in filenames:
            complain = Complain(filename)
            with open(filename) as f:
                for counter, line in enumerate(f, 1):
                    hash_at = line.find('#')
                    if hash_at >= 0:
                        line = line[:hash_at]
                    line = line.strip()
                    if not line:
                        continue
                    prefix, suffix, data = cls._split(line, counter, complain)
                    section = result.sections.get(prefix)
                    if section is None:
                        section = result.sections[prefix] = ParameterSection(prefix, {}, complain)
                    definition = section.definitions.get(suffix)
                    if definition is None:
                        definition = section.definitions[suffix] = ParameterDefinition(suffix, [], complain)
                    definition.lines.append((counter, data))
        return result

    @staticmethod
    def _split(line, counter, complain):
        colon = line.find(':')
        if colon == -1:
            complain(counter, 'does not contain a colon')
        prefix = line[:colon].upper()
        rest = line[colon + 1:].strip()
        if not rest:
            complain(counter, 'does not have text after the colon')
        if len(prefix.split()) > 1:
            complain(counter, 'has a prefix that contains whitespace')
        space = rest.find(' ')
        if space == 0:
            complain(counter, 'does not have a definition after the prefix')
        elif space == -1:
            complain(counter, 'does not have data after the definition')
        return prefix, rest[:space].upper(), rest[space + 1:]

    def copy(self):
        return Parameters({p: s.copy() for p, s in self.sections.items()})

    def write_to_file(self, filename):
        with open(filename, 'w') as f:
            for prefix, section in self.sections.items():
                for suffix, definition in section.definitions.items():
                    for counter, data in definition.lines:
                        f.write('%s:%s %s\n' % (prefix, suffix, data))
                    f.write('\n')
                f.write('\n\n')

    def __getitem__(self, prefix):
        s = self.sections.get(prefix.upper())
        return ParameterSection(prefix, {}) if s is None else s


class ParameterSection(object):
    """parameters.py:165-186"""

    def __init__(self, prefix, definitions=None, complain=None):
        self.prefix = prefix
        self.definitions = {} if definitions is None else definitions
        self.complain = Complain() if complain is None else complain

    def __getitem__(self, suffix):
        d = self.definitions.get(suffix.upper())
        return ParameterDefinition(suffix, [], self.complain) if d is None else d

    def copy(self):
        return ParameterSection(self.prefix, {k: d.copy() for k, d in self.definitions.items()}, self.complain)


class ParameterDefinition(object):
    """parameters.py:189-206"""

    def __init__(self, suffix, lines=None, complain=None):
        self.suffix = suffix
        self.lines = [] if lines is None else lines
        self.complain = Complain() if complain is None else complain

    def __getitem__(self, index):
        return self.lines[index]

    def __iter__(self):
        return iter(self.lines)

    def copy(self):
        return ParameterDefinition(self.suffix, list(self.lines), self.complain)
