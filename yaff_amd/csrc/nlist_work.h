// Cell-binning workspace of the neighbour-list build and the per-atom binning
// step, shared by the build (nlist.hip) and the Verlet drift (md.hip), which
// bins the atoms at their new positions in the same pass.
#pragma once

#include "common.h"

namespace ynb {

struct Work {
  int32_t *slot_idx;   // [nbin][cb] atoms of each bin in arrival order
  int32_t *bin_count;  // [nbin]
  double4 *xw, *sxw;   // wrapped coordinates: per atom, and per bin slot
  int4 *wr, *sw;       // wrap counts (+ atom index in .w)
  int32_t *bin_end;    // [nbin] end of each bin's slots (they start at bin * cb)
  size_t total;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline Work carve(void *base, int64_t npad, int32_t nbin, int32_t cb) {
  Work w;
  char *p = (char *)base;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = p ? p + off : nullptr;
    off += align256(bytes);
    return q;
  };
  const size_t nb = (size_t)(nbin > 0 ? nbin : 1);
  const size_t slots = nb * (size_t)(cb > 0 ? cb : 1);
  // the bin counters first: their place does not depend on npad, nbin or cb,
  // so a batch whose extents change keeps them (zero between builds)
  w.bin_count = (int32_t *)take(4 * nb);
  w.slot_idx = (int32_t *)take(4 * slots);
  w.xw = (double4 *)take(32 * npad);
  w.wr = (int4 *)take(16 * npad);
  w.sxw = (double4 *)take(32 * slots);
  w.sw = (int4 *)take(16 * slots);
  w.bin_end = (int32_t *)take(4 * nb);
  w.total = off;
  return w;
}

// Atom g (its system S; active: a real atom of S) at position (p0, p1, p2):
// wrapped into the cell (Yaff's fractional floor per axis), its bin from the
// wrapped fractional coordinates, an arrival slot in the bin by an atomic
// counter (k_bin_order fixes the order afterwards).  Called by every lane of
// a wavefront (the overflow status is a wavefront maximum).
__device__ inline void bin_atom(const ynb_sys &S, bool active, int64_t g, double p0, double p1, double p2, int32_t cb,
                                int32_t *bin_count, int32_t *slot_idx, double4 *xw, int4 *wr, int32_t *status) {
  int fill = 0;
  if (active && S.nbin != 0) {
    int c[3], w[3];
    double x = p0, y = p1, z = p2;
    for (int k = 0; k < 3; ++k) {
      const double f = fma(S.gvecs[3 * k + 2], p2, fma(S.gvecs[3 * k + 1], p1, S.gvecs[3 * k] * p0));
      const double fl = floor(f);
      int ci = (int)((f - fl) * S.grid[k]);
      ci = ci < 0 ? 0 : (ci >= S.grid[k] ? S.grid[k] - 1 : ci);
      c[k] = ci;
      w[k] = (int)fl;
      x = fma(-fl, S.rvecs[3 * k], x);
      y = fma(-fl, S.rvecs[3 * k + 1], y);
      z = fma(-fl, S.rvecs[3 * k + 2], z);
    }
    const int key = S.bin0 + (c[0] * S.grid[1] + c[1]) * S.grid[2] + c[2];
    xw[g] = make_double4(x, y, z, 0.0);
    wr[g] = make_int4(w[0], w[1], w[2], (int)g);
    const int slot = atomicAdd(bin_count + key, 1);
    if (slot < cb)
      slot_idx[(int64_t)key * cb + slot] = (int32_t)g;
    else
      atomicOr(status, 4);
    fill = slot + 1;
  }
  // the fullest bin, recorded only on overflow (the caller sizes the retry
  // from it; bin counts are cleared by k_bin_order)
  wave_status_max(status + 1, fill > cb ? fill : 0);
}

}  // namespace ynb
