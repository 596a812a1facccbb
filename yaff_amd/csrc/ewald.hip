// Ewald reciprocal space (yaff/pes/ewald.c:32-150) and the scaled-pair /
// self correction (yaff/pes/ewald.c:278-349).
//
// Phases.  With f = gvecs.r the fractional coordinates, k.r = 2 pi g.f, so
// exp(i k.r) for consecutive g0 at fixed (g1, g2) follows by multiplying with
// exp(2 pi i f0).  The k vectors are grouped in rows (runs of consecutive g0);
// a row or segment starts from one sincospi of the reduced phase
// g.f - rint(g.f) (computed in FP64 in both precisions) and continues by the
// recurrence.  Fractional coordinates are wrapped to [0, 1), which leaves
// every phase unchanged modulo 2 pi.
//
// Pass 1 (k_ewald_sk): one lane per segment of <= L k vectors, atoms of a
//   chunk read with scalar loads; partial S(k) per (chunk, k).
// Pass 2 (k_ewald_kreduce): one thread per k sums its chunks in a fixed
//   order per group (atoms below nlow, between, above nhigh), forms the
//   energy and virial terms of ewald.c:82-127 and the coefficients 2 c S(k).
// Pass 3 (k_ewald_force): one lane per atom, the system's rows split in
//   kchunks parts over blockIdx.y; partial gradients per part.
// Pass 4 (k_ewald_gsum, k_ewald_sysreduce): fixed-order sums.
// No floating-point atomics: results do not depend on scheduling or on the
// other systems of a batch.
#include "common.h"

namespace ynb {
namespace {

constexpr int CHUNK = 256;  // atoms per structure-factor chunk (LDS)

__global__ void k_ewald_prep(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad, const double *pos,
                             const double *charges, double *frac, double *step) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= npad) return;
  const int sid = wave_sys[g >> 6];
  double f[3] = {0.0, 0.0, 0.0}, q = 0.0, c = 1.0, s = 0.0;
  if (sid >= 0) {
    const ynb_sys &S = sys[sid];
    if (g - S.atom0 < S.natom && S.nk > 0) {
      const double p0 = pos[3 * g], p1 = pos[3 * g + 1], p2 = pos[3 * g + 2];
      for (int k = 0; k < 3; ++k) {
        const double x = fma(S.gvecs[3 * k + 2], p2, fma(S.gvecs[3 * k + 1], p1, S.gvecs[3 * k] * p0));
        f[k] = x - floor(x);
      }
      q = charges[g];
      sincospi(2.0 * f[0], &s, &c);
    }
  }
  frac[g] = f[0];
  frac[npad + g] = f[1];
  frac[2 * npad + g] = f[2];
  frac[3 * npad + g] = q;
  step[g] = c;
  step[npad + g] = s;
}

// Seed of a recurrence: sin and cos of pi x for x = 2 (phase in turns) in
// [-1, 1].  Mixed mode evaluates it in FP32 on the FP64-reduced argument;
// the recurrences and all sums stay in FP64 in both modes.
template <int PREC>
__device__ inline void seed_sincospi(double x, double *s, double *c) {
  if (PREC == YNB_FP64) {
    sincospi(x, s, c);
  } else {
    float sf, cf;
    sincospif((float)x, &sf, &cf);
    *s = sf;
    *c = cf;
  }
}

// Phase g0 f0 + g1 f1 + g2 f2 in turns reduced to [-1/2, 1/2] in FP64.
__device__ inline double reduced_phase(int g0, int g1, int g2, double f0, double f1, double f2) {
  const double t = fma((double)g2, f2, fma((double)g1, f1, (double)g0 * f0));
  return t - rint(t);
}

// One lane per segment of <= L consecutive k of a row; the atoms of the work
// item's chunk are walked in order and their data are the same for every
// lane, so they come through scalar loads (no LDS staging).
template <int PREC, int L>
__global__ __launch_bounds__(64) void k_ewald_sk(const ynb_sys *__restrict__ sys, const int32_t *__restrict__ work,
                                                 const int32_t *__restrict__ chunk, const int32_t *__restrict__ seg,
                                                 const int32_t *__restrict__ seg_len,
                                                 const double *__restrict__ frac, const double *__restrict__ step,
                                                 double *__restrict__ spart, int nwork, int64_t npad) {
#pragma clang fp contract(fast)
  const int lane = threadIdx.x;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int4 wk = *(const int4 *)(work + 4 * w);
    const ynb_sys &S = sys[wk.x];
    const int4 ch = *(const int4 *)(chunk + 4 * wk.z);
    const int64_t a0 = S.atom0 + ch.y;
    const int na = ch.z;
    const bool valid = lane < wk.w;
    int4 sg = make_int4(0, 0, 0, 0);
    int len = 0;
    if (valid) {
      sg = *(const int4 *)(seg + 4 * (wk.y + lane));
      len = seg_len[wk.y + lane];
    }
    double ar[L], ai[L];
#pragma unroll
    for (int m = 0; m < L; ++m) ar[m] = ai[m] = 0.0;
#pragma unroll 2
    for (int t = 0; t < na; ++t) {
      const int64_t g = a0 + t;
      const double ph = reduced_phase(sg.x, sg.y, sg.z, frac[g], frac[npad + g], frac[2 * npad + g]);
      double sn, cs;
      seed_sincospi<PREC>(2.0 * ph, &sn, &cs);
      const double q = frac[3 * npad + g];
      double zr = q * cs, zi = q * sn;
      const double cr = step[g], ci = step[npad + g];
#pragma unroll
      for (int m = 0; m < L; ++m) {
        ar[m] += zr;
        ai[m] += zi;
        const double nr = zr * cr - zi * ci;
        zi = zr * ci + zi * cr;
        zr = nr;
      }
    }
    if (valid) {
      double *out = spart + 2 * (S.spart0 + (int64_t)(wk.z - S.chunk0) * S.nk + (sg.w - S.k0));
#pragma unroll
      for (int m = 0; m < L; ++m) {
        if (m < len) {
          out[2 * m] = ar[m];
          out[2 * m + 1] = ai[m];
        }
      }
    }
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
  const int sid = ew.kblk_sys[k >> 8];
  const ynb_sys &S = sys[sid];
  const int lk = k - S.k0;
  const double4 kv = *(const double4 *)(ew.kvec + 4 * (int64_t)k);
  __shared__ double l_s[6][8][32];
  // chunk groups: [0, lo) low (1), [lo, nchunk - hi) mid (0), [nchunk - hi, nchunk) high (2)
  const int lo = S.nchunk_low, hi = S.nchunk_high, nmid_end = S.nchunk - hi;
  const int gb[3] = {lo, 0, nmid_end}, ge[3] = {nmid_end, lo, S.nchunk};
#pragma unroll
  for (int grp = 0; grp < 3; ++grp) {
    double ac[4] = {0, 0, 0, 0}, as[4] = {0, 0, 0, 0};
    if (kv.w != 0.0) {
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
    }
    l_s[2 * grp][part][kl] = (ac[0] + ac[1]) + (ac[2] + ac[3]);
    l_s[2 * grp + 1][part][kl] = (as[0] + as[1]) + (as[2] + as[3]);
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
  }
  __syncthreads();
  __shared__ double red[7][32];
  if (part == 0) {
    red[0][kl] = e;
#pragma unroll
    for (int c = 0; c < 6; ++c) red[1 + c][kl] = v[c];
  }
  __syncthreads();
  for (int off = 16; off > 0; off >>= 1) {
    if (part == 0 && kl < off)
      for (int c = 0; c < 7; ++c) red[c][kl] += red[c][kl + off];
    __syncthreads();
  }
  if (t < 7) ew.kpart[8 * (int64_t)blockIdx.x + t] = red[t][0];
}

// One lane per atom; blockIdx.y selects a contiguous part of the system's
// k rows.  The part's k data (k vector, coefficients, g and a row-start flag)
// is staged through LDS 64 k at a time and read as broadcasts; at a row start
// the lane seeds exp(i k.r) with one sincospi, then steps it along g0.
template <int PREC, bool GROUPS>
__global__ __launch_bounds__(64) void k_ewald_force(const ynb_sys *sys, const int32_t *wave_sys, ynb_ewald ew,
                                                    int64_t npad) {
#pragma clang fp contract(fast)
  __shared__ double l_k[3][64];
  __shared__ double l_c[6][64];
  __shared__ int4 l_g[64];
  const int64_t wave = blockIdx.x;
  const int kc = blockIdx.y;
  const int lane = threadIdx.x;
  const int sid = uniform_int(wave_sys[wave]);
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (S.nk == 0) return;
  const int64_t g = wave * 64 + lane;
  const int li = (int)(g - S.atom0);
  const bool active = li < S.natom;
  const double f0 = ew.frac[g], f1 = ew.frac[npad + g], f2 = ew.frac[2 * npad + g];
  const double q = ew.frac[3 * npad + g];
  const double cr = ew.step[g], ci = ew.step[npad + g];
  double wl = 0.0, wh = 0.0;
  if (GROUPS) {
    wh = (li < S.nhigh) ? 1.0 : 0.0;
    wl = (li >= S.nlow) ? 1.0 : 0.0;
  }
  const int rb = S.row0 + (int)(((int64_t)S.nrow * kc) / ew.kchunks);
  const int re = S.row0 + (int)(((int64_t)S.nrow * (kc + 1)) / ew.kchunks);
  double G0 = 0.0, G1 = 0.0, G2 = 0.0;
  if (re > rb) {
    const int kbeg = ew.krow_k[rb];
    const int kend = ew.krow_k[re - 1] + ew.krow[4 * (re - 1) + 3];
    double zr = 1.0, zi = 0.0;
    for (int k0 = kbeg; k0 < kend; k0 += 64) {
      const int kk = k0 + lane;
      __syncthreads();
      if (kk < kend) {
        const double4 kv = *(const double4 *)(ew.kvec + 4 * (int64_t)kk);
        l_k[0][lane] = kv.x;
        l_k[1][lane] = kv.y;
        l_k[2][lane] = kv.z;
        const double *cf = ew.coef + 6 * (int64_t)kk;
        l_c[0][lane] = cf[0];
        l_c[1][lane] = cf[1];
        if (GROUPS) {
          l_c[2][lane] = cf[2];
          l_c[3][lane] = cf[3];
          l_c[4][lane] = cf[4];
          l_c[5][lane] = cf[5];
        }
        l_g[lane] = *(const int4 *)(ew.kg + 4 * (int64_t)kk);
      }
      __syncthreads();
      const int nt = min(64, kend - k0);
      for (int t = 0; t < nt; ++t) {
        const int4 gg = l_g[t];
        if (uniform_int(gg.w)) {  // row start (same k for the whole wave): seed exp(2 pi i g.f)
          double s, c;
          seed_sincospi<PREC>(2.0 * reduced_phase(gg.x, gg.y, gg.z, f0, f1, f2), &s, &c);
          zr = c;
          zi = s;
        }
        double ac = l_c[0][t], as = l_c[1][t];
        if (GROUPS) {
          ac += wh * l_c[4][t] + wl * l_c[2][t];
          as += wh * l_c[5][t] + wl * l_c[3][t];
        }
        // ewald.c:104-111: x = 2c (S_c (-q sin) + S_s (q cos)); q applied below.
        const double x = as * zr - ac * zi;
        G0 = fma(l_k[0][t], x, G0);
        G1 = fma(l_k[1][t], x, G1);
        G2 = fma(l_k[2][t], x, G2);
        const double nr = zr * cr - zi * ci;
        zi = zr * ci + zi * cr;
        zr = nr;
      }
    }
  }
  const double scale = active ? q / S.dielectric : 0.0;
  double *o = ew.gpart + ((int64_t)kc * npad + g) * 3;
  o[0] = G0 * scale;
  o[1] = G1 * scale;
  o[2] = G2 * scale;
}

// Without nlow/nhigh groups: two atoms per lane (atoms g and g + 64 of a
// 128-atom block; systems start at multiples of 128 atoms), so each k's data,
// read from LDS as broadcasts, serve two atoms.  k data are packed as
// (kx, ky, kz, 2c S_s) and (2c S_c, row start) per k.
template <int PREC>
__global__ __launch_bounds__(64) void k_ewald_force2(const ynb_sys *sys, const int32_t *wave_sys, ynb_ewald ew,
                                                     int64_t npad) {
#pragma clang fp contract(fast)
  __shared__ double4 l_ka[64];
  __shared__ double2 l_kb[64];
  __shared__ int4 l_g[64];
  const int64_t blk = blockIdx.x;
  const int kc = blockIdx.y;
  const int lane = threadIdx.x;
  const int sid = uniform_int(wave_sys[2 * blk]);
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (S.nk == 0) return;
  const int64_t ga = blk * 128 + lane, gb = ga + 64;
  const bool act_a = ga - S.atom0 < S.natom, act_b = gb - S.atom0 < S.natom;
  const double fa0 = ew.frac[ga], fa1 = ew.frac[npad + ga], fa2 = ew.frac[2 * npad + ga];
  const double fb0 = ew.frac[gb], fb1 = ew.frac[npad + gb], fb2 = ew.frac[2 * npad + gb];
  const double cra = ew.step[ga], cia = ew.step[npad + ga];
  const double crb = ew.step[gb], cib = ew.step[npad + gb];
  const int rb = S.row0 + (int)(((int64_t)S.nrow * kc) / ew.kchunks);
  const int re = S.row0 + (int)(((int64_t)S.nrow * (kc + 1)) / ew.kchunks);
  double A0 = 0.0, A1 = 0.0, A2 = 0.0, B0 = 0.0, B1 = 0.0, B2 = 0.0;
  if (re > rb) {
    const int kbeg = ew.krow_k[rb];
    const int kend = ew.krow_k[re - 1] + ew.krow[4 * (re - 1) + 3];
    double zar = 1.0, zai = 0.0, zbr = 1.0, zbi = 0.0;
    for (int k0 = kbeg; k0 < kend; k0 += 64) {
      const int kk = k0 + lane;
      __syncthreads();
      if (kk < kend) {
        const double4 kv = *(const double4 *)(ew.kvec + 4 * (int64_t)kk);
        const double *cf = ew.coef + 6 * (int64_t)kk;
        const int4 gg = *(const int4 *)(ew.kg + 4 * (int64_t)kk);
        l_ka[lane] = make_double4(kv.x, kv.y, kv.z, cf[1]);
        l_kb[lane] = make_double2(cf[0], gg.w ? 1.0 : 0.0);
        l_g[lane] = gg;
      }
      __syncthreads();
      const int nt = min(64, kend - k0);
      for (int t = 0; t < nt; ++t) {
        const double4 ka = l_ka[t];
        const double2 kb = l_kb[t];
        if (kb.y != 0.0) {  // row start (the same k for the whole wave): seed exp(2 pi i g.f)
          const int4 gg = l_g[t];
          double s, c;
          seed_sincospi<PREC>(2.0 * reduced_phase(gg.x, gg.y, gg.z, fa0, fa1, fa2), &s, &c);
          zar = c;
          zai = s;
          seed_sincospi<PREC>(2.0 * reduced_phase(gg.x, gg.y, gg.z, fb0, fb1, fb2), &s, &c);
          zbr = c;
          zbi = s;
        }
        // ewald.c:104-111: x = 2c (S_c (-q sin) + S_s (q cos)); q applied below.
        const double xa = ka.w * zar - kb.x * zai;
        const double xb = ka.w * zbr - kb.x * zbi;
        A0 = fma(ka.x, xa, A0);
        A1 = fma(ka.y, xa, A1);
        A2 = fma(ka.z, xa, A2);
        B0 = fma(ka.x, xb, B0);
        B1 = fma(ka.y, xb, B1);
        B2 = fma(ka.z, xb, B2);
        const double nar = zar * cra - zai * cia;
        zai = zar * cia + zai * cra;
        zar = nar;
        const double nbr = zbr * crb - zbi * cib;
        zbi = zbr * cib + zbi * crb;
        zbr = nbr;
      }
    }
  }
  const double sa = act_a ? ew.frac[3 * npad + ga] / S.dielectric : 0.0;
  const double sb = act_b ? ew.frac[3 * npad + gb] / S.dielectric : 0.0;
  double *oa = ew.gpart + ((int64_t)kc * npad + ga) * 3;
  double *ob = ew.gpart + ((int64_t)kc * npad + gb) * 3;
  oa[0] = A0 * sa;
  oa[1] = A1 * sa;
  oa[2] = A2 * sa;
  ob[0] = B0 * sb;
  ob[1] = B1 * sb;
  ob[2] = B2 * sb;
}

// Gradient: sum of the kchunks partial gradients, one thread per (atom,
// component), four interleaved accumulators combined in a fixed order.
__global__ void k_ewald_gsum(const ynb_sys *sys, const int32_t *wave_sys, ynb_ewald ew, int64_t npad, double *gpos,
                             int accumulate) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3 * npad) return;
  const int64_t g = t / 3;
  const int sid = wave_sys[g >> 6];
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (S.nk == 0 || g - S.atom0 >= S.natom) return;
  const int kcn = ew.kchunks;
  const double *src = ew.gpart + t;  // gpart[(kc * npad + g) * 3 + c] = src[kc * 3 * npad]
  const int64_t stride = 3 * npad;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int kc = 0;
  for (; kc + 4 <= kcn; kc += 4) {
    a0 += src[(int64_t)kc * stride];
    a1 += src[(int64_t)(kc + 1) * stride];
    a2 += src[(int64_t)(kc + 2) * stride];
    a3 += src[(int64_t)(kc + 3) * stride];
  }
  for (; kc < kcn; ++kc) a0 += src[(int64_t)kc * stride];
  const double G = (a0 + a1) + (a2 + a3);
  if (accumulate)
    gpos[t] += G;
  else
    gpos[t] = G;
}

// out[s] = {E, V (3x3)}; ewald.c:128-149 (virial diagonal -= E, then 1/eps_r).
// One block per system: 256 threads take the k blocks t, t + 256, ... and a
// fixed LDS tree adds them.
__global__ __launch_bounds__(256) void k_ewald_sysreduce(const ynb_sys *sys, int nsys, ynb_ewald ew, double *out) {
  const int s = blockIdx.x;
  const int t = threadIdx.x;
  const ynb_sys &S = sys[s];
  __shared__ double red[7][256];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  const int b0 = S.k0 / 32, b1 = (S.k0 + S.nk) / 32;
  for (int b = b0 + t; b < b1; b += 256)
#pragma unroll
    for (int c = 0; c < 7; ++c) acc[c] += ew.kpart[8 * (int64_t)b + c];
#pragma unroll
  for (int c = 0; c < 7; ++c) red[c][t] = acc[c];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off)
#pragma unroll
      for (int c = 0; c < 7; ++c) red[c][t] += red[c][t + off];
    __syncthreads();
  }
  if (t == 0) {
    const double e = red[0][0];
    const double f = (S.nk > 0) ? 1.0 / S.dielectric : 0.0;
    double *o = out + 10 * (int64_t)s;
    o[0] = e * f;
    o[1] = (red[1][0] - e) * f; o[2] = red[4][0] * f; o[3] = red[5][0] * f;
    o[4] = red[4][0] * f; o[5] = (red[2][0] - e) * f; o[6] = red[6][0] * f;
    o[7] = red[5][0] * f; o[8] = red[6][0] * f; o[9] = (red[3][0] - e) * f;
  }
}

// ewald.c:278-349 for one system, rows processed in order by lane 0 of one
// wavefront; the self term is a wave-strided sum in a fixed order.
__global__ void k_ewald_corr(const ynb_sys *sysd, const double *pos, const double *q, const ynb_scaling_row *stab,
                             int64_t nstab, double *gpos, int accumulate, double *out) {
  const ynb_sys &S = *sysd;
  const int lane = threadIdx.x;
  const double alpha = S.alpha;
  const double x = alpha / 1.7724538509055160273;  // M_SQRT_PI
  double self = 0.0;
  for (int i = S.nlow + lane; i < S.nhigh; i += 64) self += x * q[S.atom0 + i] * q[S.atom0 + i];
  self = wave_sum(self);
  const double f = 1.0 / S.dielectric;
  if (!accumulate && gpos)
    for (int i = lane; i < S.natom; i += 64)
      for (int c = 0; c < 3; ++c) gpos[3 * (int64_t)(S.atom0 + i) + c] = 0.0;
  __syncthreads();
  if (lane != 0) return;
  double energy = -self;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t r = 0; r < nstab; ++r) {
    const int64_t a = stab[r].a, b = stab[r].b;
    if (a < S.nlow && b < S.nlow) continue;
    if (a >= S.nhigh && b >= S.nhigh) continue;
    const int64_t ga = S.atom0 + a, gb = S.atom0 + b;
    double dl[3] = {pos[3 * gb] - pos[3 * ga], pos[3 * gb + 1] - pos[3 * ga + 1], pos[3 * gb + 2] - pos[3 * ga + 2]};
    int m[3];
    mic_exact(dl, S.rvecs, S.gvecs, S.nvec, m);
    const double d = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
    const double xx = alpha * d;
    const double e = erf(xx) / d;
    const double fac = (1 - stab[r].scale) * q[gb] * q[ga];
    const double gg = -fac * (1.1283791670955125739 * alpha * exp(-xx * xx) - e) / d / d;
    if (gpos) {
      for (int c = 0; c < 3; ++c) {
        gpos[3 * gb + c] += dl[c] * gg * f;
        gpos[3 * ga + c] -= dl[c] * gg * f;
      }
    }
    v[0] += dl[0] * dl[0] * gg;
    v[4] += dl[1] * dl[1] * gg;
    v[8] += dl[2] * dl[2] * gg;
    const double v01 = dl[1] * dl[0] * gg, v02 = dl[2] * dl[0] * gg, v12 = dl[2] * dl[1] * gg;
    v[1] += v01; v[3] += v01;
    v[2] += v02; v[6] += v02;
    v[5] += v12; v[7] += v12;
    energy -= fac * e;
  }
  out[0] = energy * f;
  for (int c = 0; c < 9; ++c) out[1 + c] = v[c] * f;
}

}  // namespace
}  // namespace ynb

using namespace ynb;

extern "C" int ynb_ewald_reci(const ynb_batch *b, const ynb_ewald *ew, const double *pos, const double *charges,
                              int32_t precision, double *gpos, int32_t accumulate, double *out, ynb_stream stream) {
  if (!b || !ew || !pos || !charges || !out || b->npad % 128 != 0 || ew->nk_tot % 256 != 0 || ew->kchunks < 1 ||
      (ew->seglen != 4 && ew->seglen != 8 && ew->seglen != 16) ||
      (precision != YNB_FP64 && precision != YNB_MIXED)) {
    set_error("ynb_ewald_reci: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t npad = b->npad;
  hipLaunchKernelGGL(k_ewald_prep, dim3((unsigned)ceil_div(npad, 256)), dim3(256), 0, st, b->sys, b->wave_sys, npad,
                     pos, charges, ew->frac, ew->step);
  YNB_LAUNCHED();
  if (ew->nwork > 0) {
    const dim3 grid((unsigned)(ew->nwork < 65535 ? ew->nwork : 65535));
#define YNB_SK(P, LL)                                                                                          \
  YNB_LAUNCH("ewald_sk", (k_ewald_sk<P, LL>), grid, dim3(64), 0, st, b->sys, ew->work, ew->chunk, ew->seg,       \
             ew->seg_len, ew->frac, ew->step, ew->spart, ew->nwork, npad)
    if (precision == YNB_FP64) {
      if (ew->seglen == 4) YNB_SK(YNB_FP64, 4);
      else if (ew->seglen == 8) YNB_SK(YNB_FP64, 8);
      else YNB_SK(YNB_FP64, 16);
    } else {
      if (ew->seglen == 4) YNB_SK(YNB_MIXED, 4);
      else if (ew->seglen == 8) YNB_SK(YNB_MIXED, 8);
      else YNB_SK(YNB_MIXED, 16);
    }
#undef YNB_SK
    YNB_LAUNCHED();
  }
  if (ew->nk_tot > 0) {
    YNB_LAUNCH("ewald_kreduce", k_ewald_kreduce, dim3(ew->nk_tot / 32), dim3(256), 0, st, b->sys, *ew);
    YNB_LAUNCHED();
  }
  const dim3 fgrid((unsigned)(npad / 64), (unsigned)ew->kchunks);
  const dim3 fgrid2((unsigned)(npad / 128), (unsigned)ew->kchunks);
  if (precision == YNB_FP64) {
    if (ew->groups)
      YNB_LAUNCH("ewald_force", (k_ewald_force<YNB_FP64, true>), fgrid, dim3(64), 0, st, b->sys, b->wave_sys, *ew, npad);
    else
      YNB_LAUNCH("ewald_force", (k_ewald_force2<YNB_FP64>), fgrid2, dim3(64), 0, st, b->sys, b->wave_sys, *ew, npad);
  } else {
    if (ew->groups)
      YNB_LAUNCH("ewald_force", (k_ewald_force<YNB_MIXED, true>), fgrid, dim3(64), 0, st, b->sys, b->wave_sys, *ew, npad);
    else
      YNB_LAUNCH("ewald_force", (k_ewald_force2<YNB_MIXED>), fgrid2, dim3(64), 0, st, b->sys, b->wave_sys, *ew, npad);
  }
  YNB_LAUNCHED();
  if (gpos) {
    hipLaunchKernelGGL(k_ewald_gsum, dim3((unsigned)ceil_div(3 * npad, 256)), dim3(256), 0, st, b->sys, b->wave_sys, *ew,
                       npad, gpos, accumulate);
    YNB_LAUNCHED();
  }
  hipLaunchKernelGGL(k_ewald_sysreduce, dim3((unsigned)b->nsys), dim3(256), 0, st, b->sys, b->nsys, *ew,
                     out);
  YNB_LAUNCHED();
  return YNB_OK;
}

extern "C" int ynb_ewald_corr(const ynb_sys *h_sys, const ynb_sys *d_sys, const double *pos, const double *charges,
                              const ynb_scaling_row *stab, int64_t nstab, double *gpos, int32_t accumulate, double *out,
                              ynb_stream stream) {
  if (!h_sys || !d_sys || !pos || !charges || !out || nstab < 0 || (nstab > 0 && !stab)) {
    set_error("ynb_ewald_corr: invalid arguments");
    return YNB_EINVAL;
  }
  hipLaunchKernelGGL(k_ewald_corr, dim3(1), dim3(64), 0, (hipStream_t)stream, d_sys, pos, charges, stab, nstab, gpos,
                     accumulate, out);
  YNB_LAUNCHED();
  return YNB_OK;
}
