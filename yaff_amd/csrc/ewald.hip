// Ewald reciprocal space (yaff/pes/ewald.c:32-150) and the scaled-pair /
// self correction (yaff/pes/ewald.c:278-349).
//
// Phases.  With f = gvecs.r the fractional coordinates, k.r = 2 pi g.f, so
// exp(i k.r) for consecutive g0 at fixed (g1, g2) follows by multiplying with
// exp(2 pi i f0).  The k vectors are grouped in rows (runs of consecutive g0);
// at a row start each lane seeds exp(i k.r) of its atoms from the reduced
// phase g.f - rint(g.f) (FP64 in both precisions) and steps it along the row.
// The FP64 passes step by the three-term recurrence
// z_{j+1} = 2 cos(2 pi f0) z_j - z_{j-1} (one FMA per component instead of a
// complex product; z_{-1} = z_0 exp(-2 pi i f0) at the row start).  Its
// rounding error after j steps is below ~j^2 eps whatever f0 (1e-13 for
// the 45-k rows of the 6912-atom rocksalt, against ~j eps for the complex
// product).  Fractional coordinates are wrapped to [0, 1), which leaves
// every phase unchanged modulo 2 pi.
//
// Pass 1 (k_ewald_sk): lanes hold atoms (apl per lane) of one chunk and walk
//   a block of whole rows; per k a lane sums its atoms, and every 8 k the
//   wavefront transposes its per-lane sums through LDS so that each k's sum
//   over the chunk is one fixed-order tree.  Partial S(k) per (chunk, k).
//   Mixed precision (k_ewald_sk_mixed): the terms in FP32 between FP64
//   phase anchors every 8 k of a row, the sums from the LDS tile on in FP64.
// Pass 2 (k_ewald_kreduce): one thread per k sums its chunks in a fixed
//   order per group (atoms below nlow, between, above nhigh), forms the
//   energy and virial terms of ewald.c:82-127 and the coefficients 2 c S(k).
// Pass 3 (k_ewald_force): lanes hold 1, 2 or 4 atoms each; the per-k
//   coefficients are uniform and come through scalar loads; the k vector
//   weighting is carried by two running sums per row; the 4 wavefronts of a
//   block walk 4 row blocks of the system and are summed in LDS.
// Pass 4: one extra block of the force launch per system sums its energy
//   and virial from the k-block sums (ewald_system_reduce); the last force
//   block of an atom block sums its partial gradients (ewald_gsum_block), or
//   the caller leaves them to the Verlet kick.  Fixed-order sums.
// No floating-point atomics: results do not depend on scheduling or on the
// other systems of a batch.
#include "common.h"
#include "ewald_prep.h"

namespace ynb {
namespace {

__global__ void k_ewald_prep(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad, const double *pos,
                             const double *charges, double *frac, double *step) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= npad) return;
  const int sid = wave_sys[g >> 6];
  const ynb_sys *S = sid >= 0 ? &sys[sid] : nullptr;
  const bool active = S && g - S->atom0 < S->natom;
  ewald_prep_atom(S, active, g, npad, active ? pos[3 * g] : 0.0, active ? pos[3 * g + 1] : 0.0,
                  active ? pos[3 * g + 2] : 0.0, charges, frac, step);
}

// sin and cos of 2 pi t for |t| <= 1/2 (a phase in turns), FP64: quadrant
// q = rint(4 t), remainder x = 2 pi (t - q/4) with |x| <= pi/4, Taylor series
// of sin (to x^17) and cos (to x^16), whose truncation errors on that range
// are below 1e-17; then a rotation by q quarter turns.
__device__ inline void sincos_turns(double t, double *s, double *c) {
  constexpr double S3 = -1.0 / 6.0, S5 = 1.0 / 120.0, S7 = -1.0 / 5040.0, S9 = 1.0 / 362880.0,
                   S11 = -1.0 / 39916800.0, S13 = 1.0 / 6227020800.0, S15 = -1.0 / 1307674368000.0,
                   S17 = 1.0 / 355687428096000.0;
  constexpr double C4 = 1.0 / 24.0, C6 = -1.0 / 720.0, C8 = 1.0 / 40320.0, C10 = -1.0 / 3628800.0,
                   C12 = 1.0 / 479001600.0, C14 = -1.0 / 87178291200.0, C16 = 1.0 / 20922789888000.0;
  const double qd = rint(4.0 * t);
  const double x = fma(-0.25, qd, t) * 6.283185307179586477;
  const double x2 = x * x;
  double ps = fma3(x2, S17, S15);
  ps = fma3(x2, ps, S13);
  ps = fma3(x2, ps, S11);
  ps = fma3(x2, ps, S9);
  ps = fma3(x2, ps, S7);
  ps = fma3(x2, ps, S5);
  ps = fma3(x2, ps, S3);
  const double sn = fma(x * x2, ps, x);
  double pc = fma3(x2, C16, C14);
  pc = fma3(x2, pc, C12);
  pc = fma3(x2, pc, C10);
  pc = fma3(x2, pc, C8);
  pc = fma3(x2, pc, C6);
  pc = fma3(x2, pc, C4);
  pc = fma3(x2, pc, -0.5);
  const double cn = fma3(x2, pc, 1.0);
  const int q = (int)qd & 3;
  const double a = (q & 1) ? cn : sn;
  const double b = (q & 1) ? sn : cn;
  *s = (q & 2) ? -a : a;
  *c = ((q + 1) & 2) ? -b : b;
}

// Seed of a recurrence: sin and cos of 2 pi t for a reduced phase t in
// [-1/2, 1/2] turns.  Mixed mode evaluates it in FP32 on the FP64-reduced
// argument; the force pass's recurrences and sums stay FP64 in both modes
// (the mixed structure-factor pass is k_ewald_sk_mixed).
template <int PREC>
__device__ inline void seed_sincos(double t, double *s, double *c) {
  if (PREC == YNB_FP64) {
    sincos_turns(t, s, c);
  } else {
    float sf, cf;
    sincospif(2.0f * (float)t, &sf, &cf);
    *s = sf;
    *c = cf;
  }
}

// Phase g0 f0 + g1 f1 + g2 f2 in turns reduced to [-1/2, 1/2] in FP64.
__device__ inline double reduced_phase(int g0, int g1, int g2, double f0, double f1, double f2) {
  const double t = fma((double)g2, f2, fma((double)g1, f1, (double)g0 * f0));
  return t - rint(t);
}

// Row seeds by continuation.  The next row of a block usually lies in the
// same g2 plane one step up in g1 and starts within two steps of the previous
// row's first g0: then its seed is the previous seed times exp(2 pi i f1)
// and exp(+-2 pi i f0) once or twice (complex products, a few ulp each),
// instead of a sine and cosine from scratch.  Other rows are seeded from the
// reduced phase.  The test depends only on the row records (uniform).
__device__ inline int row_step(int4 prev, int4 row, bool have_prev) {
  const int dx = row.x - prev.x;
  return (have_prev && row.z == prev.z && row.y == prev.y + 1 && dx >= -2 && dx <= 2) ? dx : 3;
}

// z *= (br + i bi)
__device__ inline void cmul(double &zr, double &zi, double br, double bi) {
  const double nr = zr * br - zi * bi;
  zi = zr * bi + zi * br;
  zr = nr;
}

// z *= exp(2 pi i (f1 + dx f0)) for |dx| <= 2 (dx uniform).
__device__ inline void continue_seed(double &zr, double &zi, double c1, double s1, double c0, double s0, int dx) {
  cmul(zr, zi, c1, s1);
  if (dx != 0) {
    const double sd = dx > 0 ? s0 : -s0;
    cmul(zr, zi, c0, sd);
    if (dx == 2 || dx == -2) cmul(zr, zi, c0, sd);
  }
}

// Pointers into the constant address space (data that no kernel of the
// launch writes: loads at uniform addresses become scalar loads).
typedef const __attribute__((address_space(4))) double cdouble;
__device__ inline double4 load_c4(const cdouble *p) { return make_double4(p[0], p[1], p[2], p[3]); }

// LDS tile of the structure-factor passes: T = 8 k rows of RS = 66 entries
// (one per lane), padded so that both the row writes and the strided reads
// are free of bank conflicts.
constexpr int SK_T = 8, SK_RS = 66;

// Sums the m filled rows of the tile: lane (part p, row r) adds entries p,
// p + 8, ..., p + 56 of row r, three xor-shuffle levels combine the 8 parts,
// and part 0 writes S(k) of row r to out[2 r], out[2 r + 1].  Fixed order.
__device__ inline void sk_flush_tile(double2 *l_t, int m, double *out) {
  const int lane = threadIdx.x;
  const int m_r = lane & (SK_T - 1), part = lane >> 3;
  __syncthreads();
  double ar = 0.0, ai = 0.0;
  if (m_r < m) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double2 v = l_t[m_r * SK_RS + part + 8 * u];
      ar += v.x;
      ai += v.y;
    }
  }
#pragma unroll
  for (int off = 8; off < 64; off <<= 1) {
    ar += __shfl_xor(ar, off, 64);
    ai += __shfl_xor(ai, off, 64);
  }
  if (part == 0 && m_r < m) {
    out[2 * m_r] = ar;
    out[2 * m_r + 1] = ai;
  }
  __syncthreads();
}

// Structure factors.  One wavefront per work item (row block, atom chunk):
// lane l holds atoms l, l + 64, ... of the chunk (P per lane).  Per k the
// lane adds its atoms' q exp(i k.r) and stores the sum in row m of an LDS
// tile of T = 8 k; when the tile is full (or the block ends), lane (part p,
// row m) adds entries p, p + 8, ..., p + 56 of row m and three xor-shuffle
// levels combine the 8 parts.  Every sum has a fixed order.  Tile rows are
// RS = 66 entries apart so that both the row writes and the strided reads
// are free of bank conflicts.  Rows are walked with a countdown; the next
// row record is prefetched.
template <int PREC, int P>
__global__ __launch_bounds__(64) void k_ewald_sk(const ynb_sys *__restrict__ sys, const int32_t *__restrict__ work,
                                                 const int32_t *__restrict__ chunk,
                                                 const int32_t *__restrict__ skblk, const int32_t *__restrict__ krow,
                                                 const int32_t *__restrict__ krow_k,
                                                 const double *__restrict__ frac, const double *__restrict__ step,
                                                 double *__restrict__ spart, int nwork, int64_t npad) {
#pragma clang fp contract(fast)
  constexpr int T = SK_T, RS = SK_RS;
  __shared__ double2 l_t[T * RS];
  const int lane = threadIdx.x;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int4 wk = *(const int4 *)(work + 4 * w);
    if (wk.x < 0) continue;  // an item of no system (the same for the wavefront)
    const ynb_sys &S = sys[wk.x];
    const int4 ch = *(const int4 *)(chunk + 4 * wk.z);
    const int rb = skblk[2 * wk.y], re = skblk[2 * wk.y + 1];
    double f0[P], f1[P], f2[P], q[P], cr[P], ci[P], tc[P], zr[P], zi[P], pr[P], pi[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int la = lane + 64 * p;
      const bool valid = la < ch.z;
      const int64_t g = S.atom0 + ch.y + (valid ? la : 0);
      f0[p] = frac[g];
      f1[p] = frac[npad + g];
      f2[p] = frac[2 * npad + g];
      q[p] = valid ? frac[3 * npad + g] : 0.0;
      cr[p] = step[g];
      ci[p] = step[npad + g];
      tc[p] = 2.0 * cr[p];
      zr[p] = zi[p] = pr[p] = pi[p] = 0.0;
    }
    double *out = spart + 2 * (S.spart0 + (int64_t)(wk.z - S.chunk0) * S.nk - S.k0);
    int kt = krow_k[rb];  // first k of the current tile
    int m = 0;            // k stored in the tile
    int4 rown = *(const int4 *)(krow + 4 * rb);
    for (int r = rb; r < re; ++r) {
      const int4 row = rown;
      if (r + 1 < re) rown = *(const int4 *)(krow + 4 * (r + 1));
      // Row blocks of this pass hold few rows (many work items keep the
      // chip busy), so every row is seeded from its reduced phase.
#pragma unroll
      for (int p = 0; p < P; ++p) {  // row start: seed q exp(2 pi i g.f), the term before it z exp(-2 pi i f0)
        double sn, cs;
        seed_sincos<PREC>(reduced_phase(row.x, row.y, row.z, f0[p], f1[p], f2[p]), &sn, &cs);
        zr[p] = q[p] * cs;
        zi[p] = q[p] * sn;
        pr[p] = zr[p] * cr[p] + zi[p] * ci[p];
        pi[p] = zi[p] * cr[p] - zr[p] * ci[p];
      }
      for (int j = 0; j < row.w;) {
        // k up to the end of the row or of the tile: a loop without branches
        const int n = min(row.w - j, T - m);
#pragma unroll 2
        for (int i = 0; i < n; ++i) {
          double sr = zr[0], si = zi[0];
#pragma unroll
          for (int p = 1; p < P; ++p) {
            sr += zr[p];
            si += zi[p];
          }
          l_t[(m + i) * RS + lane] = make_double2(sr, si);
          // the next term by the three-term recurrence
          // z_{j+1} = 2 cos(2 pi f0) z_j - z_{j-1} (2 FP64 instructions)
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const double nr = fma(tc[p], zr[p], -pr[p]);
            const double ni = fma(tc[p], zi[p], -pi[p]);
            pr[p] = zr[p];
            pi[p] = zi[p];
            zr[p] = nr;
            zi[p] = ni;
          }
        }
        j += n;
        m += n;
        if (m == T || (r + 1 == re && j == row.w)) {  // flush the tile
          sk_flush_tile(l_t, m, out + 2 * kt);
          kt += m;
          m = 0;
        }
      }
    }
  }
}

// Structure factors in mixed precision: the terms q exp(i k.r) in FP32,
// their phases anchored in FP64.  Along a row, every 8th term (j = 0, 8, 16,
// ... from the row start) is taken from an FP64 anchor: at the row start the
// anchor is q exp(2 pi i t) of the FP64-reduced phase t (FP32 sine and cosine
// of t), then each anchor is the previous one times exp(2 pi i 8 f0) in FP64
// (exp(2 pi i f0) squared three times per atom).  Between anchors a term
// steps in FP32 by exp(2 pi i f0): at most 7 FP32 complex products, so the
// phase error of a term stays below ~5e-7 rad instead of growing along the
// whole row.  A lane sums its atoms' terms in FP32; from the LDS tile on
// every sum is FP64, in the same fixed order as the FP64 pass.
template <int P>
__global__ __launch_bounds__(64) void k_ewald_sk_mixed(const ynb_sys *__restrict__ sys,
                                                       const int32_t *__restrict__ work,
                                                       const int32_t *__restrict__ chunk,
                                                       const int32_t *__restrict__ skblk,
                                                       const int32_t *__restrict__ krow,
                                                       const int32_t *__restrict__ krow_k,
                                                       const double *__restrict__ frac,
                                                       const double *__restrict__ step,
                                                       double *__restrict__ spart, int nwork, int64_t npad) {
#pragma clang fp contract(fast)
  constexpr int T = SK_T, RS = SK_RS, SEG = 8;
  __shared__ double2 l_t[T * RS];
  const int lane = threadIdx.x;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int4 wk = *(const int4 *)(work + 4 * w);
    if (wk.x < 0) continue;  // an item of no system (the same for the wavefront)
    const ynb_sys &S = sys[wk.x];
    const int4 ch = *(const int4 *)(chunk + 4 * wk.z);
    const int rb = skblk[2 * wk.y], re = skblk[2 * wk.y + 1];
    double f0[P], f1[P], f2[P], q[P], ar[P], ai[P], w8r[P], w8i[P];
    float cr[P], ci[P], zr[P], zi[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int la = lane + 64 * p;
      const bool valid = la < ch.z;
      const int64_t g = S.atom0 + ch.y + (valid ? la : 0);
      f0[p] = frac[g];
      f1[p] = frac[npad + g];
      f2[p] = frac[2 * npad + g];
      q[p] = valid ? frac[3 * npad + g] : 0.0;
      double c = step[g], sn = step[npad + g];  // exp(2 pi i f0)
      cr[p] = (float)c;
      ci[p] = (float)sn;
#pragma unroll
      for (int e = 0; e < 3; ++e) {  // exp(2 pi i 8 f0)
        const double nc = c * c - sn * sn;
        sn = 2.0 * c * sn;
        c = nc;
      }
      w8r[p] = c;
      w8i[p] = sn;
      ar[p] = ai[p] = 0.0;
      zr[p] = zi[p] = 0.0f;
    }
    double *out = spart + 2 * (S.spart0 + (int64_t)(wk.z - S.chunk0) * S.nk - S.k0);
    int kt = krow_k[rb];  // first k of the current tile
    int m = 0;            // k stored in the tile
    int4 rown = *(const int4 *)(krow + 4 * rb);
    for (int r = rb; r < re; ++r) {
      const int4 row = rown;
      if (r + 1 < re) rown = *(const int4 *)(krow + 4 * (r + 1));
#pragma unroll
      for (int p = 0; p < P; ++p) {  // row start: the anchor q exp(2 pi i g.f)
        float sf, cf;
        sincospif(2.0f * (float)reduced_phase(row.x, row.y, row.z, f0[p], f1[p], f2[p]), &sf, &cf);
        ar[p] = q[p] * (double)cf;
        ai[p] = q[p] * (double)sf;
        zr[p] = (float)ar[p];
        zi[p] = (float)ai[p];
      }
      for (int j = 0; j < row.w;) {
        // k up to the end of the row, of the tile or of the anchor's segment
        const int n = min(min(row.w - j, T - m), SEG - (j & (SEG - 1)));
#pragma unroll 2
        for (int i = 0; i < n; ++i) {
          float sr = zr[0], si = zi[0];
#pragma unroll
          for (int p = 1; p < P; ++p) {
            sr += zr[p];
            si += zi[p];
          }
          l_t[(m + i) * RS + lane] = make_double2((double)sr, (double)si);
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const float nr = zr[p] * cr[p] - zi[p] * ci[p];
            zi[p] = zr[p] * ci[p] + zi[p] * cr[p];
            zr[p] = nr;
          }
        }
        j += n;
        m += n;
        if ((j & (SEG - 1)) == 0 && j < row.w) {  // next anchor: 8 steps in FP64
#pragma unroll
          for (int p = 0; p < P; ++p) {
            const double nr = ar[p] * w8r[p] - ai[p] * w8i[p];
            ai[p] = ar[p] * w8i[p] + ai[p] * w8r[p];
            ar[p] = nr;
            zr[p] = (float)ar[p];
            zi[p] = (float)ai[p];
          }
        }
        if (m == T || (r + 1 == re && j == row.w)) {  // flush the tile
          sk_flush_tile(l_t, m, out + 2 * kt);
          kt += m;
          m = 0;
        }
      }
    }
  }
}

// out[s] = {E, V (3x3)}; ewald.c:128-149 (virial diagonal -= E, then 1/eps_r),
// from the system's k-block sums, by one extra block of the force launch
// (256 threads): strided sums over the blocks, wavefront trees, then the 4
// wavefronts in order.
__device__ void ewald_system_reduce(const ynb_sys &S, int s, const double *kpart, double *out) {
  const int t = threadIdx.x;
  __shared__ double l_red[4][8];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int b0 = S.k0 / 32, b1 = (S.k0 + S.nk) / 32;
  for (int b = b0 + t; b < b1; b += 256)
#pragma unroll
    for (int c = 0; c < 7; ++c) acc[c] += kpart[8 * (int64_t)b + c];
#pragma unroll
  for (int c = 0; c < 7; ++c) acc[c] = wave_sum(acc[c]);
  if ((t & 63) == 0)
#pragma unroll
    for (int c = 0; c < 7; ++c) l_red[t >> 6][c] = acc[c];
  __syncthreads();
  if (t == 0) {
    double r[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) r[c] = ((l_red[0][c] + l_red[1][c]) + l_red[2][c]) + l_red[3][c];
    const double e = r[0];
    const double f = 1.0 / S.dielectric;
    double *o = out + 10 * (int64_t)s;
    o[0] = e * f + S.ew_const;
    o[1] = (r[1] - e) * f; o[2] = r[4] * f; o[3] = r[5] * f;
    o[4] = r[4] * f; o[5] = (r[2] - e) * f; o[6] = r[6] * f;
    o[7] = r[5] * f; o[8] = r[6] * f; o[9] = (r[3] - e) * f;
  }
}

// Block of 32 k vectors x 8 chunk parts: part p sums the chunks p, p+8, ...
// of each group (4 independent accumulators), then part sums are added in a
// fixed order; thread (k, part 0) forms the energy/virial terms of k and the
// coefficients; a fixed LDS tree sums energy and virial over the block.
__global__ __launch_bounds__(256) void k_ewald_kreduce(const ynb_sys *sys, ynb_ewald ew) {
  const int t = threadIdx.x;
  const int kl = t & 31, part = t >> 5;
  const int k = blockIdx.x * 32 + kl;
  const int sid = ew.kblk_sys[k >> 8];  // the same for the block
  if (sid < 0) return;                   // k vectors of no system
  const ynb_sys &S = sys[sid];
  const int lk = k - S.k0;
  const double4 kv = *(const double4 *)(ew.kvec + 4 * (int64_t)k);
  const int kp = ew.kpos[k];  // requested with kv, used at the end
  __shared__ double l_s[6][8][32];
  // chunk groups: [0, lo) low (1), [lo, nchunk - hi) mid (0), [nchunk - hi, nchunk) high (2)
  const int lo = S.nchunk_low, hi = S.nchunk_high, nmid_end = S.nchunk - hi;
  const int gb[3] = {lo, 0, nmid_end}, ge[3] = {nmid_end, lo, S.nchunk};
#pragma unroll
  for (int grp = 0; grp < 3; ++grp) {
    double ac[4] = {0, 0, 0, 0}, as[4] = {0, 0, 0, 0};
    // the partial sums are read whether or not k is a padding vector (the
    // addresses are valid; a padding vector's entries are never written), so
    // the loads do not wait for kv; a padding vector's sums are dropped below
    for (int c = gb[grp] + part; c < ge[grp]; c += 32) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int cu = c + 8 * u;
        if (cu < ge[grp]) {
          const double2 v = *(const double2 *)(ew.spart + 2 * (S.spart0 + (int64_t)cu * S.nk + lk));
          ac[u] += v.x;
          as[u] += v.y;
        }
      }
    }
    const bool real_k = kv.w != 0.0;
    l_s[2 * grp][part][kl] = real_k ? (ac[0] + ac[1]) + (ac[2] + ac[3]) : 0.0;
    l_s[2 * grp + 1][part][kl] = real_k ? (as[0] + as[1]) + (as[2] + as[3]) : 0.0;
  }
  __syncthreads();
  double e = 0.0, v[6] = {0, 0, 0, 0, 0, 0};
  if (part == 0) {
    double sc[3], ss[3];
#pragma unroll
    for (int grp = 0; grp < 3; ++grp) {
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        a += l_s[2 * grp][p][kl];
        b += l_s[2 * grp + 1][p][kl];
      }
      sc[grp] = a;
      ss[grp] = b;
    }
    double cf[6] = {0, 0, 0, 0, 0, 0};
    if (kv.w != 0.0) {
      // ewald.c:82-95: S0 (mid), S1 (low), S2 (high).
      double s = sc[0] * (sc[0] + 2.0 * (sc[1] + sc[2])) + ss[0] * (ss[0] + 2.0 * (ss[1] + ss[2]));
      s += 2.0 * (sc[1] * sc[2] + ss[1] * ss[2]);
      e = kv.w * s;
      const double x = 2.0 * kv.w;
      cf[0] = x * sc[0]; cf[1] = x * ss[0];
      cf[2] = x * sc[1]; cf[3] = x * ss[1];
      cf[4] = x * sc[2]; cf[5] = x * ss[2];
      // ewald.c:113-127
      const double ksq = kv.x * kv.x + kv.y * kv.y + kv.z * kv.z;
      const double fac2 = 0.25 / S.alpha / S.alpha;
      const double c = kv.w * (2.0 * (1.0 / ksq + fac2) * s);
      v[0] = c * kv.x * kv.x;
      v[1] = c * kv.y * kv.y;
      v[2] = c * kv.z * kv.z;
      v[3] = c * kv.y * kv.x;
      v[4] = c * kv.z * kv.x;
      v[5] = c * kv.z * kv.y;
    }
    double *o = ew.coef + 6 * (int64_t)k;
#pragma unroll
    for (int c = 0; c < 6; ++c) o[c] = cf[c];
    if (kp >= 0) *(double2 *)(ew.frec + 2 * (int64_t)kp) = make_double2(cf[1], cf[0]);
    // energy and virial sums of the block's 32 k: a tree over lanes 0-31 of
    // the first wavefront (lane l adds lane l + off for off = 16, 8, ..., 1)
    double r[7] = {e, v[0], v[1], v[2], v[3], v[4], v[5]};
#pragma unroll
    for (int off = 16; off > 0; off >>= 1)
#pragma unroll
      for (int c = 0; c < 7; ++c) r[c] += __shfl_down(r[c], off, 32);
    if (kl == 0)
#pragma unroll
      for (int c = 0; c < 7; ++c) ew.kpart[8 * (int64_t)blockIdx.x + c] = r[c];
  }
}

// Gradient of one atom block (na atoms from g0): the sum of the system's
// nslot partial gradients, one thread per (atom, component), four
// interleaved accumulators combined in a fixed order; written or, with
// accumulate, added to gpos.  Run by the atom block's last force block.
__device__ void ewald_gsum_block(const ynb_sys &S, int64_t g0, int na, const double *gpart, int64_t npad,
                                 double *gpos, int accumulate) {
  const int ns = S.nslot;
  const int64_t stride = 3 * npad;
  for (int t = threadIdx.x; t < 3 * na; t += blockDim.x) {
    const int64_t e = 3 * g0 + t;  // gpart[(slot * npad + g) * 3 + c] = src[slot * 3 npad]
    if (g0 + t / 3 - S.atom0 >= S.natom) continue;
    const double *src = gpart + e;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int sl = 0;
    for (; sl + 4 <= ns; sl += 4) {
      a0 += gather(src + (int64_t)sl * stride);
      a1 += gather(src + (int64_t)(sl + 1) * stride);
      a2 += gather(src + (int64_t)(sl + 2) * stride);
      a3 += gather(src + (int64_t)(sl + 3) * stride);
    }
    if (sl < ns) {  // the last 1-3 slots into a0 in slot order, their loads issued together
      double v[3];
#pragma unroll
      for (int u = 0; u < 3; ++u) v[u] = gather(src + (int64_t)min(sl + u, ns - 1) * stride);
#pragma unroll
      for (int u = 0; u < 3; ++u)
        if (sl + u < ns) a0 += v[u];
    }
    const double G = (a0 + a1) + (a2 + a3);
    if (accumulate)
      gpos[e] += G;
    else
      gpos[e] = G;
  }
}

// Forces (ewald.c:104-111), no nlow/nhigh groups.  Block (atom block of 64 P
// atoms, slot y) of 4 wavefronts: wavefront w walks row block 4 y + w of the
// system with P atoms per lane (atoms g + 64 p).
//
// Along a row k_j = k_0 + j b (b = 2 pi times the first reciprocal vector),
// so with x_j = 2c (S_s cos - S_c sin) at k_j the row adds
//   sum_j k_j x_j = k_0 T + b (L T - U),  T = sum_j x_j,  U = sum_j T_j,
// T_j the running sum and L the number of steps; the inner loop carries T
// and U (two adds) instead of three k-weighted sums.  The only per-k data
// are the two coefficients: 16-byte records (k_ewald_kreduce), rows padded
// to a multiple of 4 records with zeros, read 4 at a time with one scalar
// load that is issued a group ahead.  A zero record adds x = 0 and leaves
// the row sum unchanged when L counts the padded steps.  The 4 partial
// gradients of a block are added in LDS in wavefront order and scaled by
// q / dielectric.
template <int PREC, int P>
__global__ __launch_bounds__(256) void k_ewald_force(const ynb_sys *__restrict__ sys,
                                                     const int32_t *__restrict__ wave_sys,
                                                     const double *__restrict__ frec, const int32_t *__restrict__ krow,
                                                     const int32_t *__restrict__ krow_f,
                                                     const int32_t *__restrict__ fblk,
                                                     const double *__restrict__ frac, const double *__restrict__ step,
                                                     double *__restrict__ gpart, int64_t npad, int32_t *fctr,
                                                     double *gpos, int accumulate, const double *__restrict__ kpart,
                                                     double *__restrict__ out) {
#pragma clang fp contract(fast)
  constexpr int NA = 64 * P;
  constexpr double TWO_PI = 6.283185307179586477;
  __shared__ double l_g[3][4][NA];
  const int wv = uniform_int(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t blk = blockIdx.x;
  if (blk >= npad / NA) {  // one block per system: energy and virial from the k-block sums
    const int s = (int)(blk - npad / NA);
    if (blockIdx.y == 0) {
      if (sys[s].nk > 0)
        ewald_system_reduce(sys[s], s, kpart, out);
      else if (threadIdx.x < 10)  // no k vectors: a zero term
        out[10 * (int64_t)s + threadIdx.x] = 0.0;
    }
    return;
  }
  const int sid = wave_sys[P * blk];
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (S.nk == 0 || (int)blockIdx.y >= S.nslot) return;  // the same for the whole block
  double f0[P], f1[P], f2[P], cr[P], ci[P], tc[P], c1[P], s1[P], zr[P], zi[P], pr[P], pi[P], sr0[P], si0[P], T[P],
      U[P], G0[P], G1[P], G2[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int64_t g = blk * NA + lane + 64 * p;
    f0[p] = frac[g];
    f1[p] = frac[npad + g];
    f2[p] = frac[2 * npad + g];
    cr[p] = step[g];
    ci[p] = step[npad + g];
    tc[p] = 2.0 * cr[p];
    c1[p] = step[2 * npad + g];
    s1[p] = step[3 * npad + g];
    G0[p] = G1[p] = G2[p] = 0.0;
    sr0[p] = si0[p] = 0.0;
  }
  const int fb = 4 * (int)blockIdx.y + wv;
  if (fb < S.nfrb) {
    const double b0 = TWO_PI * S.gvecs[0], b1 = TWO_PI * S.gvecs[1], b2 = TWO_PI * S.gvecs[2];
    const int rb = fblk[2 * (S.frb0 + fb)], re = fblk[2 * (S.frb0 + fb) + 1];
    // the records are read-only for the whole launch: read through the
    // constant address space, so the loads stay scalar loads
    const cdouble *rp = (const cdouble *)(frec + 2 * (int64_t)krow_f[rb]);
    double4 na = load_c4(rp), nb = load_c4(rp + 4);  // next 4 records
    int4 rown = *(const int4 *)(krow + 4 * rb);
    int4 prev = make_int4(0, 0, 0, 0);
    for (int r = rb; r < re; ++r) {
      const int4 row = rown;
      if (r + 1 < re) rown = *(const int4 *)(krow + 4 * (r + 1));
      const int dx = row_step(prev, row, r > rb);
      prev = row;
      if (dx != 3) {
#pragma unroll
        for (int p = 0; p < P; ++p) continue_seed(sr0[p], si0[p], c1[p], s1[p], cr[p], ci[p], dx);
      } else {
#pragma unroll
        for (int p = 0; p < P; ++p)  // row start: seed exp(2 pi i g.f)
          seed_sincos<PREC>(reduced_phase(row.x, row.y, row.z, f0[p], f1[p], f2[p]), &si0[p], &sr0[p]);
      }
#pragma unroll
      for (int p = 0; p < P; ++p) {
        zr[p] = sr0[p];
        zi[p] = si0[p];
        // the term before the row start, z exp(-2 pi i f0)
        pr[p] = zr[p] * cr[p] + zi[p] * ci[p];
        pi[p] = zi[p] * cr[p] - zr[p] * ci[p];
        T[p] = U[p] = 0.0;
      }
      const int ng = (row.w + 3) >> 2;
      // Scalar loads complete out of order, so a wait for one of them waits
      // for all.  Waiting here (the requests are long done after the seeds)
      // leaves nothing in flight at the loop head, so the only wait in the
      // loop falls after its arithmetic.
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
      // one group of 4 records: T += cs zr - cc zi as two fused steps,
      // U += T, and the next term by the three-term recurrence
      // z_{j+1} = 2 cos(2 pi f0) z_j - z_{j-1} (5 FP64 instructions per
      // (k, atom)); z and p swap roles every step
      auto group = [&](const double4 ca, const double4 cb) {
        const double cs[4] = {ca.x, ca.z, cb.x, cb.z}, cc[4] = {ca.y, ca.w, cb.y, cb.w};
#pragma unroll
        for (int m = 0; m < 4; m += 2) {
#pragma unroll
          for (int p = 0; p < P; ++p) {
            T[p] = fma(cs[m], zr[p], fma(-cc[m], zi[p], T[p]));
            U[p] += T[p];
            pr[p] = fma(tc[p], zr[p], -pr[p]);
            pi[p] = fma(tc[p], zi[p], -pi[p]);
            T[p] = fma(cs[m + 1], pr[p], fma(-cc[m + 1], pi[p], T[p]));
            U[p] += T[p];
            zr[p] = fma(tc[p], pr[p], -zr[p]);
            zi[p] = fma(tc[p], pi[p], -zi[p]);
          }
        }
      };
      // Two groups per trip, the records alternating between two register
      // sets (no copies): the request for the next group is issued before a
      // group's arithmetic (a scheduling barrier keeps it there) and waited
      // for after it.
      int q = 0;
      for (; q + 2 <= ng; q += 2) {
        const double4 nc = load_c4(rp + 8), nd = load_c4(rp + 12);
        __builtin_amdgcn_sched_barrier(0);
        group(na, nb);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        rp += 16;
        na = load_c4(rp);
        nb = load_c4(rp + 4);
        __builtin_amdgcn_sched_barrier(0);
        group(nc, nd);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
      }
      if (q < ng) {
        const double4 ca = na, cb = nb;
        rp += 8;
        na = load_c4(rp);
        nb = load_c4(rp + 4);
        __builtin_amdgcn_sched_barrier(0);
        group(ca, cb);
      }
      // k_0 of the row and the row's contribution
      const double g0 = (double)row.x, g1 = (double)row.y, g2 = (double)row.z;
      const double k0x = TWO_PI * (g0 * S.gvecs[0] + g1 * S.gvecs[3] + g2 * S.gvecs[6]);
      const double k0y = TWO_PI * (g0 * S.gvecs[1] + g1 * S.gvecs[4] + g2 * S.gvecs[7]);
      const double k0z = TWO_PI * (g0 * S.gvecs[2] + g1 * S.gvecs[5] + g2 * S.gvecs[8]);
      const double Lp = (double)(4 * ng);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const double sj = Lp * T[p] - U[p];
        G0[p] += k0x * T[p] + b0 * sj;
        G1[p] += k0y * T[p] + b1 * sj;
        G2[p] += k0z * T[p] + b2 * sj;
      }
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
    l_g[0][wv][lane + 64 * p] = G0[p];
    l_g[1][wv][lane + 64 * p] = G1[p];
    l_g[2][wv][lane + 64 * p] = G2[p];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * NA; t += 256) {
    const int a = t / 3, c = t - 3 * a;
    const int64_t g = blk * NA + a;
    const double v = ((l_g[c][0][a] + l_g[c][1][a]) + l_g[c][2][a]) + l_g[c][3][a];
    const double sc = (g - S.atom0 < S.natom) ? frac[3 * npad + g] / S.dielectric : 0.0;
    publish(gpart + ((int64_t)blockIdx.y * npad + blk * NA) * 3 + t, v * sc);
  }
  if (gpos && last_block(fctr + blk, S.nslot)) ewald_gsum_block(S, blk * NA, NA, gpart, npad, gpos, accumulate);
}

// Forces with the nlow/nhigh groups (ewald.c:104-111 with the per-atom
// coefficients of ewald.c:105-107): the same blocks as k_ewald_force, k
// vectors and the 6 group coefficients read per k from kvec and coef.
template <int PREC, int P>
__global__ __launch_bounds__(256) void k_ewald_force_groups(const ynb_sys *__restrict__ sys,
                                                            const int32_t *__restrict__ wave_sys,
                                                            const double *__restrict__ kvec,
                                                            const double *__restrict__ coef,
                                                            const int32_t *__restrict__ krow,
                                                            const int32_t *__restrict__ krow_k,
                                                            const int32_t *__restrict__ fblk,
                                                            const double *__restrict__ frac,
                                                            const double *__restrict__ step,
                                                            double *__restrict__ gpart, int64_t npad,
                                                            int32_t *fctr, double *gpos, int accumulate, const double *__restrict__ kpart,
    double *__restrict__ out) {
#pragma clang fp contract(fast)
  constexpr int NA = 64 * P;
  __shared__ double l_g[3][4][NA];
  const int wv = uniform_int(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t blk = blockIdx.x;
  if (blk >= npad / NA) {  // one block per system: energy and virial from the k-block sums
    const int s = (int)(blk - npad / NA);
    if (blockIdx.y == 0) {
      if (sys[s].nk > 0)
        ewald_system_reduce(sys[s], s, kpart, out);
      else if (threadIdx.x < 10)  // no k vectors: a zero term
        out[10 * (int64_t)s + threadIdx.x] = 0.0;
    }
    return;
  }
  const int sid = wave_sys[P * blk];
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (S.nk == 0 || (int)blockIdx.y >= S.nslot) return;  // the same for the whole block
  double f0[P], f1[P], f2[P], cr[P], ci[P], zr[P], zi[P], wh[P], wl[P], G0[P], G1[P], G2[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int64_t g = blk * NA + lane + 64 * p;
    f0[p] = frac[g];
    f1[p] = frac[npad + g];
    f2[p] = frac[2 * npad + g];
    cr[p] = step[g];
    ci[p] = step[npad + g];
    const int la = (int)(g - S.atom0);
    wh[p] = la < S.ew_nhigh ? 1.0 : 0.0;
    wl[p] = la >= S.ew_nlow ? 1.0 : 0.0;
    G0[p] = G1[p] = G2[p] = 0.0;
  }
  const int fb = 4 * (int)blockIdx.y + wv;
  if (fb < S.nfrb) {
    const int rb = fblk[2 * (S.frb0 + fb)], re = fblk[2 * (S.frb0 + fb) + 1];
    int k = krow_k[rb];
    for (int r = rb; r < re; ++r) {
      const int4 row = *(const int4 *)(krow + 4 * r);
#pragma unroll
      for (int p = 0; p < P; ++p)  // row start: seed exp(2 pi i g.f)
        seed_sincos<PREC>(reduced_phase(row.x, row.y, row.z, f0[p], f1[p], f2[p]), &zi[p], &zr[p]);
      for (int j = 0; j < row.w; ++j, ++k) {
        const double4 kv = *(const double4 *)(kvec + 4 * (int64_t)k);
        const double *cf = coef + 6 * (int64_t)k;
        const double c0 = cf[0], c1 = cf[1], c2 = cf[2], c3 = cf[3], c4 = cf[4], c5 = cf[5];
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const double ac = c0 + wh[p] * c4 + wl[p] * c2;
          const double as = c1 + wh[p] * c5 + wl[p] * c3;
          // x = 2c (S_s cos - S_c sin); q is applied at the end
          const double x = as * zr[p] - ac * zi[p];
          G0[p] = fma(kv.x, x, G0[p]);
          G1[p] = fma(kv.y, x, G1[p]);
          G2[p] = fma(kv.z, x, G2[p]);
          const double nr = zr[p] * cr[p] - zi[p] * ci[p];
          zi[p] = zr[p] * ci[p] + zi[p] * cr[p];
          zr[p] = nr;
        }
      }
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
    l_g[0][wv][lane + 64 * p] = G0[p];
    l_g[1][wv][lane + 64 * p] = G1[p];
    l_g[2][wv][lane + 64 * p] = G2[p];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * NA; t += 256) {
    const int a = t / 3, c = t - 3 * a;
    const int64_t g = blk * NA + a;
    const double v = ((l_g[c][0][a] + l_g[c][1][a]) + l_g[c][2][a]) + l_g[c][3][a];
    const double sc = (g - S.atom0 < S.natom) ? frac[3 * npad + g] / S.dielectric : 0.0;
    publish(gpart + ((int64_t)blockIdx.y * npad + blk * NA) * 3 + t, v * sc);
  }
  if (gpos && last_block(fctr + blk, S.nslot)) ewald_gsum_block(S, blk * NA, NA, gpart, npad, gpos, accumulate);
}

// Sums of the block partials [b0, b0 + nb) of one group (8 doubles per
// block, values 0..6 used), by one block of 256 threads: thread (value q,
// part p) adds blocks b0 + p, b0 + p + 32, ... in order, then a fixed LDS
// tree over the 32 parts.  The result of value q is in l_red[q] afterwards.
__device__ void group_block_sum(const double *partial, int64_t b0, int64_t nb, double *l_red) {
  const int t = threadIdx.x;
  const int q = t & 7, part = t >> 3;
  double a = 0.0;
  if (q < 7)
    for (int64_t b = part; b < nb; b += 32) a += gather(partial + (b0 + b) * 8 + q);
  l_red[t] = a;
  __syncthreads();
  for (int off = 16; off > 0; off >>= 1) {
    if (part < off) l_red[t] += l_red[t + 8 * off];
    __syncthreads();
  }
}

// The scaled-pair rows of ewald.c:291-334, one thread per row: delta = mic(pos_b
// - pos_a), g = -fac (2/sqrt(pi) alpha exp(-x^2) - erf(x)/d) / d^2 with
// fac = (1 - scale) q_a q_b, the row's delta * g to rowg, its energy and
// virial to the block sums; the last row block of each system adds the
// system's block sums in a fixed order and writes (or adds to) out with the
// self term and the 1/dielectric factor of ewald.c:336-347.
__global__ __launch_bounds__(256) void k_corr_rows(const ynb_sys *sys, ynb_corr c, const double *pos,
                                                   const double *q, double *out, int out_accumulate) {
  const int blk = blockIdx.x;
  const int sid = c.blk_sys[blk];
  const ynb_sys &S = sys[sid];
  const ynb_corr_sys &P = c.csys[sid];
  const int t = threadIdx.x;
  const int lr = (blk - P.blk0) * 256 + t;
  double e = 0.0, v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (lr < P.nrow) {
    const int64_t r = (int64_t)P.row0 + lr;
    const ynb_scaling_row w = c.stab[r];
    double h[3] = {0.0, 0.0, 0.0};
    if (!((w.a < P.nlow && w.b < P.nlow) || (w.a >= P.nhigh && w.b >= P.nhigh))) {
      const int64_t ga = S.atom0 + w.a, gb = S.atom0 + w.b;
      double dl[3] = {pos[3 * gb] - pos[3 * ga], pos[3 * gb + 1] - pos[3 * ga + 1], pos[3 * gb + 2] - pos[3 * ga + 2]};
      int m[3];
      mic_exact(dl, S.rvecs, S.gvecs, S.nvec, m);
      const double d = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
      const double x = P.alpha * d;
      const double pot = erf(x) / d;
      const double fac = (1 - w.scale) * q[gb] * q[ga];
      const double g = -fac * (1.1283791670955125739 * P.alpha * exp(-x * x) - pot) / d / d;
      h[0] = dl[0] * g;
      h[1] = dl[1] * g;
      h[2] = dl[2] * g;
      v[0] = dl[0] * dl[0] * g;
      v[1] = dl[1] * dl[1] * g;
      v[2] = dl[2] * dl[2] * g;
      v[3] = dl[1] * dl[0] * g;
      v[4] = dl[2] * dl[0] * g;
      v[5] = dl[2] * dl[1] * g;
      e = -fac * pot;
    }
    c.rowg[3 * r] = h[0];
    c.rowg[3 * r + 1] = h[1];
    c.rowg[3 * r + 2] = h[2];
  }
  // block sums: a fixed tree over the 256 rows
  __shared__ double l_r[7][256];
  l_r[0][t] = e;
#pragma unroll
  for (int k = 0; k < 6; ++k) l_r[1 + k][t] = v[k];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off)
#pragma unroll
      for (int k = 0; k < 7; ++k) l_r[k][t] += l_r[k][t + off];
    __syncthreads();
  }
  if (t < 8) publish(c.partial + (int64_t)blk * 8 + t, t < 7 ? l_r[t][0] : 0.0);
  if (last_block(c.ctr + sid, P.nblk)) {
    __shared__ double l_red[256];
    group_block_sum(c.partial, P.blk0, P.nblk, l_red);
    if (t == 0) {
      const double f = 1.0 / P.dielectric;
      const double r[10] = {(P.self_energy + l_red[0]) * f,
                            l_red[1] * f, l_red[4] * f, l_red[5] * f,
                            l_red[4] * f, l_red[2] * f, l_red[6] * f,
                            l_red[5] * f, l_red[6] * f, l_red[3] * f};
      double *o = out + 10 * (int64_t)sid;
      for (int k = 0; k < 10; ++k) o[k] = out_accumulate ? o[k] + r[k] : r[k];
    }
  }
}

// Gradient of the correction, one thread per (atom, component): the atom's
// rows in row order (+ delta g where it is b, - where it is a), as
// compute_ewald_corr accumulates them into its zeroed buffer, then
// 1/dielectric (ewald.c:336-341).
__global__ __launch_bounds__(256) void k_corr_atoms(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad,
                                                    ynb_corr c, double *gpos, int accumulate) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 3 * npad) return;
  const int64_t g = e / 3;
  const int comp = (int)(e - 3 * g);
  const int sid = wave_sys[g >> 6];
  if (sid < 0) return;
  if (g - sys[sid].atom0 >= sys[sid].natom) return;
  const int k0 = c.atom_ref0[g], k1 = c.atom_ref0[g + 1];
  double acc = 0.0;
  for (int k = k0; k < k1; ++k) {
    const int ref = c.atom_ref[k];
    const double x = c.rowg[3 * (int64_t)(ref >> 1) + comp];
    acc = (ref & 1) ? acc - x : acc + x;
  }
  const double G = acc * (1.0 / c.csys[sid].dielectric);
  gpos[e] = accumulate ? gpos[e] + G : G;
}

}  // namespace
}  // namespace ynb

using namespace ynb;

extern "C" int ynb_ewald_reci(const ynb_batch *b, const ynb_ewald *ew, const double *pos, const double *charges,
                              int32_t precision, double *gpos, int32_t accumulate, double *out, ynb_stream stream) {
  if (!b || !ew || !pos || !charges || !out || !ew->fctr || b->npad % 256 != 0 || ew->nk_tot % 256 != 0 ||
      ew->fslots < 0 ||
      (ew->apl != 2 && ew->apl != 4 && ew->apl != 8) || (ew->fapl != 1 && ew->fapl != 2 && ew->fapl != 4) ||
      (precision != YNB_FP64 && precision != YNB_MIXED)) {
    set_error("ynb_ewald_reci: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t npad = b->npad;
  if (!(ew->flags & YNB_EWALD_PREPARED)) {
    YNB_LAUNCH("ewald_prep", k_ewald_prep, dim3((unsigned)ceil_div(npad, 256)), dim3(256), 0, st, b->sys, b->wave_sys,
               npad, pos, charges, ew->frac, ew->step);
    YNB_LAUNCHED();
  }
  if (ew->nwork > 0) {
    const dim3 grid((unsigned)(ew->nwork < (1 << 30) ? ew->nwork : (1 << 30)));
#define YNB_SK(PR, PP)                                                                                        \
  YNB_LAUNCH("ewald_sk", (k_ewald_sk<PR, PP>), grid, dim3(64), 0, st, b->sys, ew->work, ew->chunk, ew->skblk,  \
             ew->krow, ew->krow_k, ew->frac, ew->step, ew->spart, ew->nwork, npad)
#define YNB_SKM(PP)                                                                                           \
  YNB_LAUNCH("ewald_sk", (k_ewald_sk_mixed<PP>), grid, dim3(64), 0, st, b->sys, ew->work, ew->chunk, ew->skblk, \
             ew->krow, ew->krow_k, ew->frac, ew->step, ew->spart, ew->nwork, npad)
    if (precision == YNB_FP64) {
      if (ew->apl == 2) YNB_SK(YNB_FP64, 2);
      else if (ew->apl == 4) YNB_SK(YNB_FP64, 4);
      else YNB_SK(YNB_FP64, 8);
    } else {
      if (ew->apl == 2) YNB_SKM(2);
      else if (ew->apl == 4) YNB_SKM(4);
      else YNB_SKM(8);
    }
#undef YNB_SKM
#undef YNB_SK
    YNB_LAUNCHED();
  }
  if (ew->nk_tot > 0) {
    YNB_LAUNCH("ewald_kreduce", k_ewald_kreduce, dim3(ew->nk_tot / 32), dim3(256), 0, st, b->sys, *ew);
    YNB_LAUNCHED();
  }
  if (ew->fslots > 0) {
    const int P = ew->fapl;
    // npad / (64 P) atom blocks, then one block per system for its energy and virial
    const dim3 fgrid((unsigned)(npad / (64 * P) + b->nsys), (unsigned)ew->fslots);
#define YNB_FORCE(PR, PP)                                                                                     \
  YNB_LAUNCH("ewald_force", (k_ewald_force<PR, PP>), fgrid, dim3(256), 0, st, b->sys, b->wave_sys, ew->frec,    \
             ew->krow, ew->krow_f, ew->fblk, ew->frac, ew->step, ew->gpart, npad, ew->fctr, gpos, accumulate,      \
             ew->kpart, out)
#define YNB_FORCE_G(PR, PP)                                                                                   \
  YNB_LAUNCH("ewald_force", (k_ewald_force_groups<PR, PP>), fgrid, dim3(256), 0, st, b->sys, b->wave_sys,       \
             ew->kvec, ew->coef, ew->krow, ew->krow_k, ew->fblk, ew->frac, ew->step, ew->gpart, npad, ew->fctr, gpos,    \
             accumulate, ew->kpart, out)
#define YNB_FORCE_P(PR)                                            \
  if (ew->groups) {                                                \
    if (P == 1) YNB_FORCE_G(PR, 1); else if (P == 2) YNB_FORCE_G(PR, 2); else YNB_FORCE_G(PR, 4); \
  } else {                                                         \
    if (P == 1) YNB_FORCE(PR, 1); else if (P == 2) YNB_FORCE(PR, 2); else YNB_FORCE(PR, 4); \
  }
    if (precision == YNB_FP64) {
      YNB_FORCE_P(YNB_FP64);
    } else {
      YNB_FORCE_P(YNB_MIXED);
    }
#undef YNB_FORCE_P
#undef YNB_FORCE_G
#undef YNB_FORCE
    YNB_LAUNCHED();
  }
  return YNB_OK;
}

extern "C" int ynb_ewald_corr(const ynb_batch *b, const ynb_corr *c, const double *pos, const double *charges,
                              double *gpos, int32_t accumulate, double *out, int32_t out_accumulate,
                              ynb_stream stream) {
  if (!b || !c || !pos || !charges || !out || !c->csys || !c->blk_sys || !c->ctr || !c->partial ||
      c->nrow_tot < 0 || c->nblk_tot < b->nsys || (c->nrow_tot > 0 && (!c->stab || !c->rowg)) ||
      (gpos && (!c->atom_ref0 || (c->nrow_tot > 0 && !c->atom_ref))) || b->npad % 64 != 0) {
    set_error("ynb_ewald_corr: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  YNB_LAUNCH("ewald_corr_rows", k_corr_rows, dim3((unsigned)c->nblk_tot), dim3(256), 0, st, b->sys, *c, pos,
             charges, out, out_accumulate);
  YNB_LAUNCHED();
  if (gpos) {
    YNB_LAUNCH("ewald_corr_atoms", k_corr_atoms, dim3((unsigned)ceil_div(3 * b->npad, 256)), dim3(256), 0, st,
               b->sys, b->wave_sys, b->npad, *c, gpos, accumulate);
    YNB_LAUNCHED();
  }
  return YNB_OK;
}
