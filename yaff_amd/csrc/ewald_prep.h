// Per-atom set-up of the Ewald reciprocal passes, shared by ynb_ewald_reci
// (ewald.hip) and the Verlet drift (md.hip), which prepares the atoms at
// their new positions in the same pass.
#pragma once

#include "common.h"

namespace ynb {

// Atom g (system S, NULL for padding; active: a real atom of S) at position
// (p0, p1, p2): fractional coordinates f = gvecs.r wrapped to [0, 1), its
// charge, cos, sin of 2 pi f0 (the phase step along a row) and of 2 pi f1
// (the step from a row to the next one in g1).  Padding atoms and atoms of
// systems without k vectors get f = 0, q = 0, steps 1.
__device__ inline void ewald_prep_atom(const ynb_sys *S, bool active, int64_t g, int64_t npad, double p0, double p1,
                                       double p2, const double *charges, double *frac, double *step) {
  double f[3] = {0.0, 0.0, 0.0}, q = 0.0, c = 1.0, s = 0.0, c1 = 1.0, s1 = 0.0;
  if (active && S->nk > 0) {
    for (int k = 0; k < 3; ++k) {
      const double x = fma(S->gvecs[3 * k + 2], p2, fma(S->gvecs[3 * k + 1], p1, S->gvecs[3 * k] * p0));
      f[k] = x - floor(x);
    }
    q = charges[g];
    sincospi(2.0 * f[0], &s, &c);
    sincospi(2.0 * f[1], &s1, &c1);
  }
  frac[g] = f[0];
  frac[npad + g] = f[1];
  frac[2 * npad + g] = f[2];
  frac[3 * npad + g] = q;
  step[g] = c;
  step[npad + g] = s;
  step[2 * npad + g] = c1;
  step[3 * npad + g] = s1;
}

}  // namespace ynb
