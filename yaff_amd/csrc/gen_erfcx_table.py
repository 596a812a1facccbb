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

A second, FP32 table serves the mixed-precision pair kernel: erfcx on
x in [0, 6] (beyond, erfc(x) < 2.2e-17 and the kernel clamps x to 6) as a
degree-9 polynomial in z = (t - 5/8) 8/3, t = 2 / (2 + x), z in [-1, 1],
checked with the Horner evaluation emulated in FP32 (numpy float32): it
must stay within 1.5e-6 relative (rounding of t and of the Horner steps;
about 8e-7 measured).

    python yaff_amd/csrc/gen_erfcx_table.py
"""
import os

import numpy as np
from numpy.polynomial import chebyshev as cheb
from scipy.special import erfcx

N = 24
N32 = 9
X32 = 6.0
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


def table32():
    """(coefficients, max relative error of the FP32 evaluation) of the
    FP32 polynomial."""
    t0 = 2.0/(2.0 + X32)

    def g(z):
        t = t0 + (z + 1.0)/2.0*(1.0 - t0)
        return erfcx(2.0/t - 2.0)
    m = cheb.cheb2poly(cheb.chebinterpolate(g, N32)).astype(np.float32)
    xs = np.linspace(0.0, X32, 200001).astype(np.float32)
    t = (np.float32(2.0)/(np.float32(2.0) + xs)).astype(np.float32)
    z = ((t - np.float32(0.625))*np.float32(8.0/3.0)).astype(np.float32)
    r = np.full_like(z, m[-1])
    for c in m[-2::-1]:
        r = (r*z + c).astype(np.float32)
    ref = erfcx(xs.astype(np.float64))
    return m, float(np.max(np.abs(r.astype(np.float64) - ref)/ref))


def main():
    m32, err32 = table32()
    if err32 > 1.5e-6:
        raise SystemExit('FP32 polynomial too inaccurate: %.3e' % err32)
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
    lines += ['};', '',
              '// FP32: erfcx(x) ~ sum_k c_k z^k for x in [0, %g], z = (t - 5/8) 8/3, t = 2 / (2 + x);' % X32,
              '// max relative error %.1e with FP32 Horner steps.' % err32,
              'constexpr int ERFCX32_NCOEF = %d;' % (N32 + 1),
              'constexpr float ERFCX32_XMAX = %.1ff;' % X32,
              'constexpr float ERFCX32_MONO[ERFCX32_NCOEF] = {']
    lines += ['    %sf,' % repr(float(v)) for v in m32]
    lines += ['};', '}  // namespace ynb', '']
    with open(os.path.join(HERE, 'erfcx_table.h'), 'w') as fh:
        fh.write('\n'.join(lines))
    print('max relative error %.2e (FP64 table), %.2e (FP32 table)' % (err, err32))


if __name__ == '__main__':
    main()
