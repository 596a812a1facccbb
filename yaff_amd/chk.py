"""Reader and writer of the ``.chk`` checkpoint text format that
``System.from_file`` loads (yaff/system.py:405-470 through molmod's chk
module): one field per header line ``NAME kind=KIND SHAPE`` followed, for
arrays, by the values separated by whitespace.  Kinds: ``int``, ``flt``,
``str`` (scalar on the header line), ``intar``, ``fltar``, ``strar``
(arrays of the comma-separated SHAPE) and ``none``.
"""
import numpy as np

__all__ = ['load_chk', 'dump_chk']

_DTYPES = {'intar': np.int64, 'fltar': np.float64}


def load_chk(filename):
    """Dict of the fields of a .chk file."""
    with open(filename) as f:
        tokens_by_line = [line.split() for line in f]
    result = {}
    i = 0
    n = len(tokens_by_line)
    while i < n:
        head = tokens_by_line[i]
        i += 1
        if not head:
            continue
        if len(head) < 2 or not head[1].startswith('kind='):
            raise IOError('%s: line %d is not a field header' % (filename, i))
        name, kind = head[0], head[1][5:]
        rest = head[2:]
        if kind == 'none':
            result[name] = None
        elif kind in ('int', 'flt', 'str'):
            if not rest:
                raise IOError('%s: field %s has no value' % (filename, name))
            result[name] = int(rest[0]) if kind == 'int' else (float(rest[0]) if kind == 'flt' else ' '.join(rest))
        elif kind in ('intar', 'fltar', 'strar'):
            shape = tuple(int(x) for x in rest[0].split(',')) if rest else (0,)
            size = int(np.prod(shape))
            words = []
            while len(words) < size:
                if i >= n:
                    raise IOError('%s: field %s is truncated' % (filename, name))
                words.extend(tokens_by_line[i])
                i += 1
            if len(words) != size:
                raise IOError('%s: field %s has %d values, expected %d' % (filename, name, len(words), size))
            if kind == 'strar':
                result[name] = np.array(words, dtype=object).reshape(shape).astype(str)
            else:
                result[name] = np.array(words, dtype=_DTYPES[kind]).reshape(shape)
        else:
            raise IOError('%s: field %s has an unknown kind %s' % (filename, name, kind))
    return result


def dump_chk(filename, fields):
    """Writes a dict of None, int, float, str and int/float/str arrays."""
    with open(filename, 'w') as f:
        for name in sorted(fields):
            v = fields[name]
            if v is None:
                f.write('%-40s  kind=none   None\n' % name)
            elif isinstance(v, (bool, np.bool_)):
                raise TypeError('%s: booleans are not a chk kind' % name)
            elif isinstance(v, (int, np.integer)):
                f.write('%-40s  kind=int   %d\n' % (name, v))
            elif isinstance(v, (float, np.floating)):
                f.write('%-40s  kind=flt   %s\n' % (name, repr(float(v))))
            elif isinstance(v, str):
                f.write('%-40s  kind=str   %s\n' % (name, v))
            else:
                a = np.asarray(v)
                if a.dtype.kind in 'iu':
                    kind, fmt = 'intar', lambda x: '%22d' % x
                elif a.dtype.kind == 'f':
                    kind, fmt = 'fltar', lambda x: '%22s' % repr(float(x))
                else:
                    kind, fmt = 'strar', lambda x: '%22s' % x
                f.write('%-40s  kind=%s %s\n' % (name, kind, ','.join(str(d) for d in a.shape)))
                flat = a.ravel()
                for k in range(0, len(flat), 4):
                    f.write(' '.join(fmt(x) for x in flat[k:k + 4]) + '\n')
