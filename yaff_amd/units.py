"""Atomic units used throughout (bohr, hartree, electron mass, atomic time).

The angstrom is the value Yaff's generator defaults are written in
(yaff/pes/generator.py:90: 10 A = 18.89726133921252 bohr).  The other
constants follow molmod's CODATA-2006 based definitions, which parameter
files are converted with (``parse_unit``, the units of ``UNIT`` lines).
"""
import ast
import math

__all__ = ['parse_unit']

au = 1.0
e = 1.0                                   # elementary charge
angstrom = 1.889726133921252              # 1 / 0.52917720859
meter = 1e10*angstrom
centimeter = 1e-2*meter
nanometer = nm = 1e-9*meter
picometer = pm = 1e-12*meter
bohr = 1.0
joule = 1.0/4.35974394e-18
mol = 6.02214179e23
calorie = 4.184*joule
kjmol = 1000.0*joule/mol
kcalmol = 1000.0*calorie/mol
coulomb = 1.0/1.602176487e-19
electronvolt = ev = joule/coulomb
hartree = 1.0
rydberg = 0.5
kilogram = 1.0/9.10938215e-31
gram = 1e-3*kilogram
unified = amu = 1.66053886e-27/9.10938215e-31
second = 1.0/2.418884326505e-17
nanosecond = 1e-9*second
picosecond = 1e-12*second
femtosecond = 1e-15/2.418884326505e-17
boltzmann = 1.3806504e-23/4.35974394e-18
kelvin = 1.0
rad = 1.0
deg = math.pi/180.0
newton = joule/meter
pascal = newton/meter**2
bar = 1e5*pascal
debye = 1e-21/299792458.0*coulomb*meter

_NAMES = {k: v for k, v in dict(globals()).items() if isinstance(v, float) and not k.startswith('_')}


def parse_unit(expression):
    """The value in atomic units of a unit expression such as 'kcalmol',
    '1/angstrom' or 'electronvolt*angstrom**6' (names above, numbers, * / **
    and parentheses); raises ValueError for anything else."""
    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name):
            if node.id not in _NAMES:
                raise ValueError('unknown unit %s' % node.id)
            return _NAMES[node.id]
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Mult, ast.Div, ast.Pow)):
            a, b = ev(node.left), ev(node.right)
            if isinstance(node.op, ast.Mult):
                return a*b
            if isinstance(node.op, ast.Div):
                return a/b
            return a**b
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.UAdd, ast.USub)):
            v = ev(node.operand)
            return v if isinstance(node.op, ast.UAdd) else -v
        raise ValueError('unsupported unit expression')
    try:
        return ev(ast.parse(expression.strip(), mode='eval'))
    except SyntaxError:
        raise ValueError('unparsable unit %r' % expression)
