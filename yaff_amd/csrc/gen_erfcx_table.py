#!/usr/bin/env python
"""Writes erfcx_table.h: power-series coefficients of a polynomial
approximation of the scaled complementary error function
erfcx(x) = exp(x^2) erfc(x) on x >= 0, in the variable y = 2 t - 1 with
t = 2 / (2 + x) (x = 0 -> y = 1, x -> inf -> y = -1).

The polynomial is the degree-N Chebyshev interpolant of scipy.special.erfcx
(Faddeeva-package implementation) at Chebyshev points, converted to powers of
y so that the device evaluates it by Horner's rule (one fused multiply-add per
term).  The monomial coefficients are small (their absolute sum is about 1),
so the conversion loses no accuracy; the script checks the Horner evaluation
in double precision against erfcx on x in [0, 27] (beyond, exp(-x^2)
underflows) and refuses to write a table worse than 5e-14 relative.

    python yaff_amd/csrc/gen_erfcx_table.py
"""
import os

import numpy as np
from numpy.polynomial import chebyshev as cheb
from scipy.special import erfcx

N = 24
HERE = os.path.dirname(os.path.abspath(__file__))


def f(y):
    t = (y + 1.0)/2.0
    safe = np.where(t <= 0.0, 1.0, t)
    return np.where(t <= 0.0, 0.0, erfcx(2.0/safe - 2.0))


def horner(m, y):
    r = np.full_like(y, m[-1])
    for c in m[-2::-1]:
        r = r*y + c
    return r


def main():
    c = cheb.chebinterpolate(f, N)
    m = cheb.cheb2poly(c)
    xs = np.linspace(0.0, 27.0, 400001)
    approx = horner(m, 2.0*(2.0/(2.0 + xs)) - 1.0)
    err = float(np.max(np.abs(approx - erfcx(xs))/erfcx(xs)))
    if err > 5e-14:
        raise SystemExit('polynomial too inaccurate: %.3e' % err)
    lines = ['// Generated by gen_erfcx_table.py: coefficients m_0..m_%d of erfcx(x) ~ sum_k m_k y^k,' % N,
             '// y = 2 t - 1, t = 2 / (2 + x) (Horner); max relative error %.1e on [0, 27].' % err,
             '#pragma once', '', 'namespace ynb {',
             'constexpr int ERFCX_NCOEF = %d;' % (N + 1),
             'constexpr double ERFCX_MONO[ERFCX_NCOEF] = {']
    lines += ['    %s,' % repr(float(v)) for v in m]
    lines += ['};', '}  // namespace ynb', '']
    with open(os.path.join(HERE, 'erfcx_table.h'), 'w') as fh:
        fh.write('\n'.join(lines))
    print('max relative error %.2e' % err)


if __name__ == '__main__':
    main()
