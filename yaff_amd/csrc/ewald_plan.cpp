// Host-side planning of the Ewald reciprocal tables of one system: the k
// vectors compute_ewald_reci visits and their rows, and the greedy row
// blocks of the two passes.  Host C++ (no device code); the engine calls it
// once per system and cell.
#include <cmath>
#include <cstdint>

#include "yaff_nb.h"

namespace {
constexpr double TWO_PI = 6.2831853071795864769;
constexpr double FOUR_PI = 12.5663706143591729538;
}  // namespace

// k = g0 b0 + g1 b1 + g2 b2 with b = 2 pi gvecs, |k|^2 and the cutoff test
// exactly as ewald.c:38-58 forms them (this file is built without
// contraction: no fused multiply-adds); the half space g2 > 0, or g2 = 0 and
// g1 > 0, or g2 = g1 = 0 and g0 > 0.  Visited in row order: g2, then g1,
// then g0 ascending (ewald.c loops g0 outermost; the set is the same).
extern "C" int ynb_ewald_kvectors(const double *gvecs, double volume, double alpha, double gcut,
                                  const int64_t *gmax, int64_t cap, double *kvec, int32_t *g, int32_t *krow,
                                  int32_t *krow_k, int64_t *nk_out, int64_t *nrow_out) {
  if (!gvecs || !gmax || !kvec || !g || !krow || !krow_k || !nk_out || !nrow_out || !(volume > 0.0) ||
      !(alpha > 0.0) || gmax[0] < 0 || gmax[1] < 0 || gmax[2] < 0 || gmax[0] > (1 << 20) || gmax[1] > (1 << 20) ||
      gmax[2] > (1 << 20))
    return YNB_EINVAL;
  double kv[9];
  for (int i = 0; i < 9; ++i) kv[i] = TWO_PI * gvecs[i];
  double gc = gcut * TWO_PI;
  gc *= gc;
  const double fac1 = FOUR_PI / volume;
  const double fac2 = 0.25 / alpha / alpha;
  int64_t nk = 0, nrow = 0;
  for (int64_t g2 = 0; g2 <= gmax[2]; ++g2) {
    for (int64_t g1 = -gmax[1]; g1 <= gmax[1]; ++g1) {
      if (g2 == 0 && g1 < 0) continue;
      int64_t prev = INT64_MIN;  // g0 of the last k taken in this (g1, g2) line
      for (int64_t g0 = -gmax[0]; g0 <= gmax[0]; ++g0) {
        if (g2 == 0 && g1 == 0 && g0 <= 0) continue;
        const double f0 = (double)g0, f1 = (double)g1, f2 = (double)g2;
        const double kx = f0 * kv[0] + f1 * kv[3] + f2 * kv[6];
        const double ky = f0 * kv[1] + f1 * kv[4] + f2 * kv[7];
        const double kz = f0 * kv[2] + f1 * kv[5] + f2 * kv[8];
        const double ksq = kx * kx + ky * ky + kz * kz;
        if (ksq > gc) continue;
        if (nk >= cap) return YNB_ECAPACITY;
        if (prev != g0 - 1) {  // a new row: a run of consecutive g0
          krow[4 * nrow] = (int32_t)g0;
          krow[4 * nrow + 1] = (int32_t)g1;
          krow[4 * nrow + 2] = (int32_t)g2;
          krow[4 * nrow + 3] = 0;
          krow_k[nrow] = (int32_t)nk;
          ++nrow;
        }
        ++krow[4 * (nrow - 1) + 3];
        prev = g0;
        kvec[4 * nk] = kx;
        kvec[4 * nk + 1] = ky;
        kvec[4 * nk + 2] = kz;
        kvec[4 * nk + 3] = fac1 * std::exp(-ksq * fac2) / ksq;
        g[3 * nk] = (int32_t)g0;
        g[3 * nk + 1] = (int32_t)g1;
        g[3 * nk + 2] = (int32_t)g2;
        ++nk;
      }
    }
  }
  *nk_out = nk;
  *nrow_out = nrow;
  return YNB_OK;
}

// Consecutive rows in blocks of at most kmax k vectors (a longer row is a
// block of its own): blk[b] = {first row, end row}.
extern "C" int ynb_row_blocks(const int32_t *krow, int64_t nrow, int32_t kmax, int32_t *blk, int64_t *nblk_out) {
  if ((nrow > 0 && (!krow || !blk)) || !nblk_out || kmax < 1 || nrow < 0) return YNB_EINVAL;
  int64_t nb = 0, r0 = 0, acc = 0;
  for (int64_t r = 0; r < nrow; ++r) {
    const int64_t c = krow[4 * r + 3];
    if (acc > 0 && acc + c > kmax) {
      blk[2 * nb] = (int32_t)r0;
      blk[2 * nb + 1] = (int32_t)r;
      ++nb;
      r0 = r;
      acc = 0;
    }
    acc += c;
  }
  if (nrow > 0) {
    blk[2 * nb] = (int32_t)r0;
    blk[2 * nb + 1] = (int32_t)nrow;
    ++nb;
  }
  *nblk_out = nb;
  return YNB_OK;
}

// Force records: each row's k vectors padded to a multiple of 4 records
// (the force pass reads 4 at a time); krow_f[r] = the row's first record,
// kpos[k] = the record of k (-1 for the padding k from nk to nkp).
extern "C" int ynb_ewald_records(const int32_t *krow, int64_t nrow, int64_t nkp, int32_t *krow_f, int32_t *kpos,
                                 int64_t *nrec_out) {
  if ((nrow > 0 && (!krow || !krow_f)) || !kpos || !nrec_out || nrow < 0 || nkp < 0) return YNB_EINVAL;
  int64_t rec = 0, k = 0;
  for (int64_t r = 0; r < nrow; ++r) {
    const int64_t c = krow[4 * r + 3];
    if (c < 0 || k + c > nkp) return YNB_EINVAL;
    krow_f[r] = (int32_t)rec;
    for (int64_t j = 0; j < c; ++j) kpos[k++] = (int32_t)(rec + j);
    rec += (c + 3) / 4 * 4;
  }
  for (; k < nkp; ++k) kpos[k] = -1;
  *nrec_out = rec;
  return YNB_OK;
}

// Structure-factor work items of one system: every (row block, atom chunk),
// block major: {sys, blk0 + block, chunk0 + chunk, 0}.
extern "C" int ynb_ewald_work(int64_t nblk, int64_t nchunk, int32_t sys, int32_t blk0, int32_t chunk0,
                              int32_t *work) {
  if (nblk < 0 || nchunk < 0 || (nblk * nchunk > 0 && !work)) return YNB_EINVAL;
  int64_t w = 0;
  for (int64_t b = 0; b < nblk; ++b)
    for (int64_t c = 0; c < nchunk; ++c, ++w) {
      work[4 * w] = sys;
      work[4 * w + 1] = (int32_t)(blk0 + b);
      work[4 * w + 2] = (int32_t)(chunk0 + c);
      work[4 * w + 3] = 0;
    }
  return YNB_OK;
}
