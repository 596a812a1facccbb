// Pair potentials over the full neighbour list, all potential slots of a
// system in one pass (Yaff makes one pass per ForcePartPair,
// yaff/pes/pair_pot.c:84-176).
//
// One lane per atom i walks its ELL entries; both directions of a pair are in
// the list, so the lane accumulates its own gradient in registers without
// atomics and energy/virial get a factor 1/2.  Per-block sums of energy and
// virial go to `partial`; the last block of each system to finish adds the
// system's block sums in a fixed order (pair_system_reduce), so results are
// deterministic and independent of the other systems of the batch.
//
// FP64 mode rebuilds every row vector exactly as nlist_build_low does
// (yaff_row_vector), so d, the cutoff test d < rcut and the scaling lookup on
// r == 0 rows see Yaff's bits.  Mixed mode takes the image vector
// pos_j - pos_i + n.R in FP64 and evaluates the potentials in FP32,
// accumulating gradient, energy and virial in FP64.
#include <mutex>

#include "common.h"
#include "pair_eval.h"

namespace ynb {
namespace {

// pair_pot.c:65-81 as a binary search over the system's sorted rows.
__device__ inline double stab_scale(const ynb_scaling_row *st, int n, int64_t a, int64_t b) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const ynb_scaling_row &r = st[mid];
    if (r.a < a || (r.a == a && r.b < b))
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo < n && st[lo].a == a && st[lo].b == b) return st[lo].scale;
  return 1.0;
}

template <int G, typename T>
__device__ inline T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

// G lanes per atom i walk the atom's row (entries s, s+G, ...); both
// directions of every pair are in the list, so each lane accumulates the
// gradient of its own atom and energy/virial get a factor 1/2.
//
// Image vectors: u = pos_j - pos_i + n.R in FP64.  Where Yaff's bits matter
// -- a distance within a relative 1e-12 of a cutoff, or the image index of a
// scaled (bonded) pair -- the vector is rebuilt exactly as nlist_build_low
// does (yaff_row_vector), so the cutoff tests d < rcut and the scaling
// lookups agree with the reference.  FP64 mode evaluates the potentials in
// FP64; mixed mode takes exp/erfc in FP32 on FP64-reduced arguments,
// accumulating gradient, energy and virial in FP64.
//
// A system goes to its specialised kernel when it has a form and the slot
// mask keeps all of its potentials.
__device__ inline bool form_applies(const ynb_sys &S, uint32_t mask) {
  const uint32_t all = (1u << S.npot) - 1u;
  return S.pair_form != YNB_FORM_GENERIC && (mask & all) == all;
}

// SLOTS: per-potential energy, virial and gradient (the drop-in parts need
// them): one pass over the row per potential.  Otherwise one pass in which
// the potentials of a pair are summed first, so that only totals are
// accumulated (one gradient/virial update per pair); the totals go to
// potential slot 0 and the other slots get zeros.
// The system's block partials summed by its last block (256 threads): nval/8
// value groups (nval = 8: totals in slot 0; nval = 8 npot: per-potential
// slots) x 256/nval strided parts, four interleaved accumulators, then a
// fixed LDS tree over the parts.  out[s][p] = {E, V00 V01 V02 V10 V11 V12 V20
// V21 V22} with the 1/2 of the full list applied; slots past nval/8 get
// zeros.  apb: atoms per block of the pair kernel.
__device__ void pair_system_reduce(const ynb_sys &S, int s, const double *partial, double *out, int apb, int nval) {
  __shared__ double l_red[256];
  const int t = threadIdx.x;
  const int nval_p2 = nval <= 8 ? 8 : (nval <= 16 ? 16 : 32);
  const int nparts = 256 / nval_p2;
  const int q = t % nval_p2, part = t / nval_p2;
  const int64_t b0 = S.atom0 / apb, b1 = (S.atom0 + S.natom + apb - 1) / apb;
  double a[4] = {0, 0, 0, 0};
  if (q < nval) {
    // 16 loads in flight per thread (clamped addresses, masked adds): the
    // whole system's partials arrive in one or two memory round trips
    for (int64_t b = b0 + part; b < b1; b += 16 * nparts) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int64_t bu = min(b + (int64_t)nparts * u, b1 - 1);
        v[u] = gather(partial + bu * (YNB_MAX_POT * 8) + q);
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (b + (int64_t)nparts * u < b1) a[u & 3] += v[u];
    }
  }
  l_red[t] = (a[0] + a[1]) + (a[2] + a[3]);
  __syncthreads();
  for (int off = nparts / 2; off > 0; off >>= 1) {
    if (part < off) l_red[t] += l_red[t + off * nval_p2];
    __syncthreads();
  }
  if (t < S.npot) {
    double r[8];
    for (int c = 0; c < 8; ++c) r[c] = (t * 8 < nval) ? 0.5 * l_red[t * 8 + c] : 0.0;
    double *o = out + ((int64_t)s * YNB_MAX_POT + t) * 10;
    o[0] = r[0];
    o[1] = r[1]; o[2] = r[4]; o[3] = r[5];
    o[4] = r[4]; o[5] = r[2]; o[6] = r[6];
    o[7] = r[5]; o[8] = r[6]; o[9] = r[3];
  }
}

// Blocks of a system in the pair kernels: the system's waves (at least one)
// in blocks of apb atoms.
__device__ inline int pair_system_blocks(const ynb_sys &S, int apb) {
  const int n = S.natom > 0 ? S.natom : 1;
  return ((n + 63) / 64) * 64 / apb;
}

template <int PREC, bool SLOTS, int G>
__global__ __launch_bounds__(256) void k_pair(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad,
                                              ynb_pairpar pp, const double *pos, ynb_nlist nl, uint32_t mask,
                                              double *gpos, int accumulate, double *gpos_slots,
                                              double *partial, double *out) {
  using PE = PotEval<PREC>;
  using PF = PotFast<PREC>;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t gi = t / G;
  const int s = (int)(t % G);
  const int lane = threadIdx.x & 63;
  if (gi >= npad) return;  // whole waves: npad * G is a multiple of 64
  const int sid = uniform_int(wave_sys[gi >> 6]);
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (!SLOTS && form_applies(S, mask)) return;  // a specialised kernel handles this system
  const int li = (int)(gi - S.atom0);
  const bool active = li < S.natom;
  const int npot = S.npot;
  double rcmax = 0.0;
  for (int p = 0; p < npot; ++p) rcmax = fmax(rcmax, S.pot[p].rcut);
  const int ngroup = SLOTS ? npot : 1;

  __shared__ double red[4][YNB_MAX_POT * 8];
  const int wib = threadIdx.x >> 6;
  double G3[3] = {0.0, 0.0, 0.0};
  const double pi[3] = {active ? pos[3 * gi] : 0.0, active ? pos[3 * gi + 1] : 0.0, active ? pos[3 * gi + 2] : 0.0};
  const int n = active ? min(nl.cnt[gi], nl.kcap) : 0;
  const int32_t *rowj = nl.j + gi * nl.kcap;
  const int32_t *rowi = nl.img + gi * nl.kcap;

  for (int grp = 0; grp < ngroup; ++grp) {
    const int p_lo = SLOTS ? grp : 0, p_hi = SLOTS ? grp + 1 : npot;
    double acc_e = 0.0, acc_v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, acc_g[3] = {0.0, 0.0, 0.0};
    for (int k = s; k < n; k += G) {
      const int64_t gj = rowj[k];
      int nn[3];
      unpack_img(rowi[k], nn);
      const int lj = (int)(gj - S.atom0);
      const bool self = (gj == gi);
      const double pj[3] = {pos[3 * gj], pos[3 * gj + 1], pos[3 * gj + 2]};
      double u[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      for (int a = 0; a < 3; ++a) {
        const double na = (double)nn[a];
        u[0] = fma(na, S.rvecs[3 * a], u[0]);
        u[1] = fma(na, S.rvecs[3 * a + 1], u[1]);
        u[2] = fma(na, S.rvecs[3 * a + 2], u[2]);
      }
      double dd = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      bool near_cut = false;
      for (int p = 0; p < npot; ++p) near_cut |= fabs(dd - S.pot[p].rcut) <= 1e-12 * S.pot[p].rcut;
      bool r_zero = false;
      const int A = li > lj ? li : lj, B = li > lj ? lj : li;
      if (near_cut || (S.any_stab && !self)) {
        const bool i_is_a = (li == A);
        const int nab[3] = {i_is_a ? nn[0] : -nn[0], i_is_a ? nn[1] : -nn[1], i_is_a ? nn[2] : -nn[2]};
        double v[3];
        int r[3];
        yaff_row_vector(i_is_a ? pi : pj, i_is_a ? pj : pi, self, nab, S.rvecs, S.gvecs, S.nvec, S.rmax, v, r);
        r_zero = !self && r[0] == 0 && r[1] == 0 && r[2] == 0;
        if (near_cut) {
          u[0] = i_is_a ? v[0] : -v[0];
          u[1] = i_is_a ? v[1] : -v[1];
          u[2] = i_is_a ? v[2] : -v[2];
          dd = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
      }
      if (!(dd < rcmax)) continue;
      const double d = dd;
      const double inv = 1.0 / d;
      const double inv2 = inv * inv;
      double vg_sum = 0.0, e_sum = 0.0;
#pragma unroll 1
      for (int p = p_lo; p < p_hi; ++p) {
        if (!((mask >> p) & 1u)) continue;
        const ynb_pot &P = S.pot[p];
        if (!(d < P.rcut)) continue;
        double sc = 1.0;
        if (P.nstab > 0 && r_zero) sc = stab_scale(pp.stab + P.stab0, P.nstab, A, B);
        if (sc == 0.0) continue;
        const double *par0 = pp.par + (int64_t)p * 3 * npad;
        const double *par1 = par0 + npad;
        const double *par2 = par1 + npad;
        double g = 0.0, e;
        switch (P.kind) {
          case YNB_POT_LJ:  // par2 = sqrt(epsilon)
            e = PF::lj(par0[gi], par0[gj], par2[gi] * par2[gj], inv, inv2, g);
            break;
          case YNB_POT_MM3:
            e = PE::mm3(par0[gi], par0[gj], par1[gi], par1[gj], (int)par2[gi] + (int)par2[gj], d, g);
            break;
          case YNB_POT_EXPREP: {
            const int nf = P.nffatype;
            const int64_t ci = P.cross0 + (int64_t)((int)par0[gi] * nf + (int)par0[gj]);  // symmetric tables
            e = PF::exprep(pp.cross[ci], pp.cross[ci + (int64_t)nf * nf], d, inv, g);
            break;
          }
          default: {
            const double ra = par1[gi], rb = par1[gj];
            if (P.alpha > 0.0 && ra == 0.0 && rb == 0.0) {
              e = PF::ei_point((par0[gi] * par0[gj]) / P.dielectric, P.alpha, d, inv, inv2, g);
            } else {
              // Gaussian charges or alpha = 0: the reference expression
              e = PE::ei(par0[gi], par0[gj], ra, rb, P.alpha, P.dielectric, d, g);
            }
            break;
          }
        }
        if (P.trunc != YNB_TR_NONE && (e != 0.0 || g != 0.0)) {  // pair_pot.c:130-139
          double hg;
          const double h = (P.trunc == YNB_TR_SWITCH3) ? PE::switch3(d, P.rcut, P.trunc_par, hg)
                                                        : PE::hammer(d, P.rcut, P.trunc_par, hg);
          g = g * h + e * hg * inv;
          e *= h;
        }
        vg_sum += g * sc;
        e_sum += sc * e;
      }
      acc_e += e_sum;
      acc_g[0] -= u[0] * vg_sum;
      acc_g[1] -= u[1] * vg_sum;
      acc_g[2] -= u[2] * vg_sum;
      acc_v[0] += u[0] * (u[0] * vg_sum);
      acc_v[1] += u[1] * (u[1] * vg_sum);
      acc_v[2] += u[2] * (u[2] * vg_sum);
      acc_v[3] += u[0] * (u[1] * vg_sum);
      acc_v[4] += u[0] * (u[2] * vg_sum);
      acc_v[5] += u[1] * (u[2] * vg_sum);
    }
    // gradient: sum over the atom's G lanes; lane 0 of the group stores
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      acc_g[c] = group_sum<G>(acc_g[c]);
      G3[c] += acc_g[c];
    }
    if (SLOTS && gpos_slots && active && s == 0)
#pragma unroll
      for (int c = 0; c < 3; ++c) gpos_slots[((int64_t)grp * npad + gi) * 3 + c] = acc_g[c];
    // wave sums of energy and virial of this slot
    const double e = wave_sum(acc_e);
    double v[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = wave_sum(acc_v[c]);
    if (lane == 0) {
      red[wib][grp * 8] = e;
#pragma unroll
      for (int c = 0; c < 6; ++c) red[wib][grp * 8 + 1 + c] = v[c];
      red[wib][grp * 8 + 7] = 0.0;
    }
  }
  if (active && s == 0 && gpos) {
    for (int c = 0; c < 3; ++c) {
      if (accumulate)
        gpos[3 * gi + c] += G3[c];
      else
        gpos[3 * gi + c] = G3[c];
    }
  }
  // Block sums (the block's atoms belong to one system; lanes past natom
  // contribute zeros): the 4 waves of the block in order; the system's last
  // block adds the system's block sums.
  __syncthreads();
  if (threadIdx.x < npot * 8) {
    const int q = threadIdx.x;
    publish(partial + (int64_t)blockIdx.x * (YNB_MAX_POT * 8) + q,
            (q < ngroup * 8) ? ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q] : 0.0);
  }
  if (last_block(pp.ctr + sid, pair_system_blocks(S, 256 / G)))
    pair_system_reduce(S, sid, partial, out, 256 / G, SLOTS ? 8 * npot : 8);
}


// Specialised totals kernels for the common potential sets (YNB_FORM_*): the
// same lane layout, image vectors, exact near-cutoff treatment and block
// partials as k_pair, with the potentials inlined and the per-atom
// parameters read as one 32-byte record per atom.
//   FORM_LJ:    pair_fn_lj (+ switch3), pair_pot.c:277-301, truncation.c:59-76
//   FORM_IONIC: pair_fn_exprep + pair_fn_lj + pair_fn_ei (alpha > 0, point
//               charges), pair_pot.c:473-489, 277-301, 711-751
template <int PREC, int FORM, int G>
__global__ __launch_bounds__(256) void k_pair_form(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad,
                                                   ynb_pairpar pp, const double *pos, ynb_nlist nl, uint32_t mask,
                                                   double *gpos, int accumulate, double *partial, double *out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t gi = t / G;
  const int s = (int)(t % G);
  const int lane = threadIdx.x & 63;
  if (gi >= npad) return;
  const int sid = uniform_int(wave_sys[gi >> 6]);
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (S.pair_form != FORM || !form_applies(S, mask)) return;
  const int li = (int)(gi - S.atom0);
  const bool active = li < S.natom;
  const int npot = S.npot;
  FormConst C = form_const<FORM>(S, pp.cross);
  // ExpRep cross tables of up to 16 atom types go to LDS (a block's atoms
  // belong to one system, so the condition is the same for the block)
  __shared__ double l_amp[256], l_bex[256];
  __shared__ float l_ampf[256], l_bexf[256];
  if (FORM == YNB_FORM_IONIC && C.nf <= 16) {
    for (int q = threadIdx.x; q < C.nf * C.nf; q += blockDim.x) {
      l_amp[q] = C.amp[q];
      l_bex[q] = C.bex[q];
      l_ampf[q] = (float)C.amp[q];
      l_bexf[q] = (float)C.bex[q];
    }
    __syncthreads();
    C.lds_cross = true;
    C.l_amp = l_amp;
    C.l_bex = l_bex;
    C.l_ampf = l_ampf;
    C.l_bexf = l_bexf;
  }

  double acc_e = 0.0, acc_v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, acc_g[3] = {0.0, 0.0, 0.0};
  if (active) {
    const double pi[3] = {pos[3 * gi], pos[3 * gi + 1], pos[3 * gi + 2]};
    const double4 ri = *(const double4 *)(pp.rec + 4 * gi);
    const int ti = FORM == YNB_FORM_IONIC ? (int)ri.w : 0;
    const int n = min(nl.cnt[gi], nl.kcap);
    const int32_t *rowj = nl.j + gi * nl.kcap;
    const int32_t *rowi = nl.img + gi * nl.kcap;
    if constexpr (PREC == YNB_FP64) {
      // Software pipeline over the lane's entries k = s, s + G, ...: the
      // partner's position and record are requested one entry ahead and its
      // index two entries ahead, so each entry's loads overlap the previous
      // entry's arithmetic.  Past the row end the atom itself stands in (a
      // valid address; its data are never used).
      const int32_t gi32 = (int32_t)gi;
      int32_t jA = s < n ? rowj[s] : gi32;
      int32_t iA = s < n ? rowi[s] : 0;
      int32_t jB = s + G < n ? rowj[s + G] : gi32;
      int32_t iB = s + G < n ? rowi[s + G] : 0;
      double pA[3] = {pos[3 * (int64_t)jA], pos[3 * (int64_t)jA + 1], pos[3 * (int64_t)jA + 2]};
      double4 rA = *(const double4 *)(pp.rec + 4 * (int64_t)jA);
      for (int k = s; k < n; k += G) {
        const int64_t gj = jA;
        const int32_t imgk = iA;
        const double pj[3] = {pA[0], pA[1], pA[2]};
        const double4 rj = rA;
        jA = jB;
        iA = iB;
        pA[0] = pos[3 * (int64_t)jB];
        pA[1] = pos[3 * (int64_t)jB + 1];
        pA[2] = pos[3 * (int64_t)jB + 2];
        rA = *(const double4 *)(pp.rec + 4 * (int64_t)jB);
        // the index two entries ahead: an unconditional load of a valid entry
        // (k itself past the row end), selected afterwards, so no branch makes
        // the loads in flight wait
        const int k2 = k + 2 * G;
        const int kk = k2 < n ? k2 : k;
        const int32_t jt = rowj[kk], it = rowi[kk];
        jB = k2 < n ? jt : gi32;
        iB = k2 < n ? it : 0;
        int nn[3];
        unpack_img(imgk, nn);
        double u[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
        for (int a = 0; a < 3; ++a) {
          const double na = (double)nn[a];
          u[0] = fma(na, S.rvecs[3 * a], u[0]);
          u[1] = fma(na, S.rvecs[3 * a + 1], u[1]);
          u[2] = fma(na, S.rvecs[3 * a + 2], u[2]);
        }
        double e_sum, vg_sum;
        if (!form_pair<PREC, FORM>(S, C, li, gi, gj, pi, pos, nn, ri, rj, ti, u, e_sum, vg_sum)) continue;
        acc_e += e_sum;
        acc_g[0] -= u[0] * vg_sum;
        acc_g[1] -= u[1] * vg_sum;
        acc_g[2] -= u[2] * vg_sum;
        acc_v[0] += u[0] * (u[0] * vg_sum);
        acc_v[1] += u[1] * (u[1] * vg_sum);
        acc_v[2] += u[2] * (u[2] * vg_sum);
        acc_v[3] += u[0] * (u[1] * vg_sum);
        acc_v[4] += u[0] * (u[2] * vg_sum);
        acc_v[5] += u[1] * (u[2] * vg_sum);
      }
    } else {
      // Mixed precision: NE = 2 entries per trip, k and k + G (the lane's
      // entries are k = s, s + G, ...), whose arithmetic interleaves: their
      // geometry first (form_geom_mixed, with the rare exact path), then the
      // potential sums as straight-line code (form_pot_mixed).  (In FP64 two
      // entries per trip need more registers than the interleaving saves:
      // 409 against 392 us on the 64-system batch; mixed 332 against 349.)
      // Software pipeline: the next trip's partner positions and records are
      // requested one trip ahead, its indices two trips ahead.  Past the row
      // end the atom itself stands in (a valid address); such an entry's
      // results are selected away, as are those of an entry beyond every
      // cutoff, so the sums are the ones of the entries in row order.
      constexpr int NE = 2;
      const int32_t gi32 = (int32_t)gi;
      auto idx = [&](int k, int32_t &j, int32_t &im) {
        // an unconditional load of a valid entry (the lane's first one past
        // the row end), selected afterwards: no branch makes loads wait
        const int kk = k < n ? k : (s < n ? s : 0);
        const int32_t jt = rowj[kk], it = rowi[kk];
        j = k < n ? jt : gi32;
        im = k < n ? it : 0;
      };
      auto data = [&](int32_t j, double p[3], double4 &r) {
        p[0] = pos[3 * (int64_t)j];
        p[1] = pos[3 * (int64_t)j + 1];
        p[2] = pos[3 * (int64_t)j + 2];
        r = *(const double4 *)(pp.rec + 4 * (int64_t)j);
      };
      int32_t jA[NE], iA[NE], jB[NE], iB[NE];
      double pA[NE][3];
      double4 rA[NE];
  #pragma unroll
      for (int e = 0; e < NE; ++e) {
        idx(s + e * G, jA[e], iA[e]);
        idx(s + (NE + e) * G, jB[e], iB[e]);
        data(jA[e], pA[e], rA[e]);
      }
      for (int k = s; k < n; k += NE * G) {
        int64_t gj[NE];
        int nn[NE][3];
        double u[NE][3];
        double4 rj[NE];
  #pragma unroll
        for (int e = 0; e < NE; ++e) {
          gj[e] = jA[e];
          unpack_img(iA[e], nn[e]);
          u[e][0] = pA[e][0] - pi[0];
          u[e][1] = pA[e][1] - pi[1];
          u[e][2] = pA[e][2] - pi[2];
          rj[e] = rA[e];
          jA[e] = jB[e];
          iA[e] = iB[e];
          data(jA[e], pA[e], rA[e]);
          idx(k + (2 * NE + e) * G, jB[e], iB[e]);
        }
        bool ok[NE];
        PairGeom gm[NE];
  #pragma unroll
        for (int e = 0; e < NE; ++e) {
          for (int a = 0; a < 3; ++a) {
            const double na = (double)nn[e][a];
            u[e][0] = fma(na, S.rvecs[3 * a], u[e][0]);
            u[e][1] = fma(na, S.rvecs[3 * a + 1], u[e][1]);
            u[e][2] = fma(na, S.rvecs[3 * a + 2], u[e][2]);
          }
          ok[e] = form_geom_mixed<FORM>(S, C, li, gi, gj[e], pi, pos, nn[e], u[e], gm[e]) && k + e * G < n;
        }
        double es[NE], vgs[NE];
  #pragma unroll
        for (int e = 0; e < NE; ++e) form_pot_mixed<FORM>(C, gm[e], ri, rj[e], ti, es[e], vgs[e]);
  #pragma unroll
        for (int e = 0; e < NE; ++e) {
          if (!ok[e]) continue;
          const double vg = vgs[e];
          acc_e += es[e];
          acc_g[0] -= u[e][0] * vg;
          acc_g[1] -= u[e][1] * vg;
          acc_g[2] -= u[e][2] * vg;
          acc_v[0] += u[e][0] * (u[e][0] * vg);
          acc_v[1] += u[e][1] * (u[e][1] * vg);
          acc_v[2] += u[e][2] * (u[e][2] * vg);
          acc_v[3] += u[e][0] * (u[e][1] * vg);
          acc_v[4] += u[e][0] * (u[e][2] * vg);
          acc_v[5] += u[e][1] * (u[e][2] * vg);
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) acc_g[c] = group_sum<G>(acc_g[c]);
  if (active && s == 0 && gpos) {
    for (int c = 0; c < 3; ++c) {
      if (accumulate)
        gpos[3 * gi + c] += acc_g[c];
      else
        gpos[3 * gi + c] = acc_g[c];
    }
  }
  __shared__ double red[4][8];
  const int wib = threadIdx.x >> 6;
  const double e = wave_sum(acc_e);
  double v[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) v[c] = wave_sum(acc_v[c]);
  if (lane == 0) {
    red[wib][0] = e;
#pragma unroll
    for (int c = 0; c < 6; ++c) red[wib][1 + c] = v[c];
    red[wib][7] = 0.0;
  }
  __syncthreads();
  if (threadIdx.x < npot * 8) {
    const int q = threadIdx.x;
    publish(partial + (int64_t)blockIdx.x * (YNB_MAX_POT * 8) + q,
            (q < 8) ? ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q] : 0.0);
  }
  if (last_block(pp.ctr + sid, pair_system_blocks(S, 256 / G))) pair_system_reduce(S, sid, partial, out, 256 / G, 8);
}

// pair_pot_compute over explicit Yaff rows (yaff/pes/pair_pot.c:84-176):
// one thread per row, rows in any order.  With derivs, the row's
// delta * vg goes to rowh (when gradients are wanted) and its virial terms
// to the block sums; the energy always.  No atomics: each atom's gradient
// is summed afterwards by k_pair_rows_gather, and the block sums by
// k_pair_rows_reduce, both in a fixed order.
__global__ __launch_bounds__(256) void k_pair_rows(const ynb_neigh_row *rows, int64_t nrow,
                                                   const ynb_scaling_row *stab, int64_t nstab, ynb_pot P,
                                                   const double *p0, const double *p1, const double *p2,
                                                   const double *cross, int derivs, double *rowh, double *partial) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double e = 0.0, v[6] = {0, 0, 0, 0, 0, 0}, h3[3] = {0.0, 0.0, 0.0};
  if (q < nrow) {
    const ynb_neigh_row w = rows[q];
    if (w.d < P.rcut) {
      double s = 1.0;
      if (w.r0 == 0 && w.r1 == 0 && w.r2 == 0 && nstab > 0) s = stab_scale(stab, (int)nstab, w.a, w.b);
      if (s != 0.0) {
        double g = 0.0, pe;
        const double d = w.d;
        switch (P.kind) {
          case YNB_POT_LJ: pe = PotEval<YNB_FP64>::lj(p0[w.a], p0[w.b], p1[w.a], p1[w.b], d, g); break;
          case YNB_POT_MM3:
            pe = PotEval<YNB_FP64>::mm3(p0[w.a], p0[w.b], p1[w.a], p1[w.b], (int)p2[w.a] + (int)p2[w.b], d, g);
            break;
          case YNB_POT_EXPREP: {
            const int nf = P.nffatype;
            const int64_t ci = P.cross0 + (int64_t)((int)p0[w.a] * nf + (int)p0[w.b]);
            pe = PotEval<YNB_FP64>::exprep(cross[ci], cross[ci + (int64_t)nf * nf], d, g);
            break;
          }
          default: pe = PotEval<YNB_FP64>::ei(p0[w.a], p0[w.b], p1[w.a], p1[w.b], P.alpha, P.dielectric, d, g); break;
        }
        if (!derivs) {
          if (P.trunc != YNB_TR_NONE && pe != 0.0) {
            double hg;
            pe *= (P.trunc == YNB_TR_SWITCH3) ? PotEval<YNB_FP64>::switch3(d, P.rcut, P.trunc_par, hg)
                                              : PotEval<YNB_FP64>::hammer(d, P.rcut, P.trunc_par, hg);
          }
        } else {
          if (P.trunc != YNB_TR_NONE && (pe != 0.0 || g != 0.0)) {
            double hg;
            const double h = (P.trunc == YNB_TR_SWITCH3) ? PotEval<YNB_FP64>::switch3(d, P.rcut, P.trunc_par, hg)
                                                          : PotEval<YNB_FP64>::hammer(d, P.rcut, P.trunc_par, hg);
            g = g * h + pe * hg / d;
            pe *= h;
          }
          g *= s;
          h3[0] = w.dx * g;
          h3[1] = w.dy * g;
          h3[2] = w.dz * g;
          v[0] = w.dx * (w.dx * g);
          v[1] = w.dy * (w.dy * g);
          v[2] = w.dz * (w.dz * g);
          v[3] = w.dx * (w.dy * g);
          v[4] = w.dx * (w.dz * g);
          v[5] = w.dy * (w.dz * g);
        }
        e = s * pe;
      }
    }
    if (rowh) {
      rowh[3 * q] = h3[0];
      rowh[3 * q + 1] = h3[1];
      rowh[3 * q + 2] = h3[2];
    }
  }
  __shared__ double red[7][256];
  const int t = threadIdx.x;
  red[0][t] = e;
  for (int c = 0; c < 6; ++c) red[1 + c][t] = v[c];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off)
      for (int c = 0; c < 7; ++c) red[c][t] += red[c][t + off];
    __syncthreads();
  }
  if (t < 8) partial[8 * (int64_t)blockIdx.x + t] = t < 7 ? red[t][0] : 0.0;
}

// The block sums of k_pair_rows in a fixed order by one block of 256
// threads: thread (value c, part p) adds blocks p, p + 32, ... in order, then
// a tree over the 32 parts; out = {E, V00 V01 V02 V10 V11 V12 V20 V21 V22}.
__global__ __launch_bounds__(256) void k_pair_rows_reduce(int64_t nblk, const double *partial, double *out) {
  __shared__ double l_red[256];
  const int t = threadIdx.x;
  const int c = t & 7, part = t >> 3;
  double a = 0.0;
  if (c < 7)
    for (int64_t b = part; b < nblk; b += 32) a += partial[8 * b + c];
  l_red[t] = a;
  __syncthreads();
  for (int off = 16; off > 0; off >>= 1) {
    if (part < off) l_red[t] += l_red[t + 8 * off];
    __syncthreads();
  }
  if (t == 0) {
    const double *r = l_red;
    out[0] = r[0];
    out[1] = r[1]; out[2] = r[4]; out[3] = r[5];
    out[4] = r[4]; out[5] = r[2]; out[6] = r[6];
    out[7] = r[5]; out[8] = r[6]; out[9] = r[3];
  }
}

// gpos[x] += the atom's rows in row order: + delta vg where x is the row's
// b, - where it is a (atom_ref: 2 r and 2 r + 1), the sequence in which
// pair_pot_compute accumulates them (pair_pot.c:147-157).  One thread per
// (atom, component).
__global__ __launch_bounds__(256) void k_pair_rows_gather(int64_t natom, const int32_t *atom_ref0,
                                                          const int32_t *atom_ref, const double *rowh,
                                                          double *gpos) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 3 * natom) return;
  const int64_t x = e / 3;
  const int comp = (int)(e - 3 * x);
  const int k0 = atom_ref0[x], k1 = atom_ref0[x + 1];
  double acc = 0.0;
  for (int k = k0; k < k1; ++k) {
    const int ref = atom_ref[k];
    const double h = rowh[3 * (int64_t)(ref >> 1) + comp];
    acc = (ref & 1) ? acc - h : acc + h;
  }
  gpos[e] += acc;
}

}  // namespace
}  // namespace ynb

using namespace ynb;

extern "C" int ynb_pair_compute(const ynb_batch *b, const ynb_pairpar *pp, const double *pos, const ynb_nlist *nl,
                                uint32_t slot_mask, int32_t precision, double *gpos, int32_t accumulate,
                                double *gpos_slots, double *partial, double *out, ynb_stream stream) {
  if (!b || !pp || !pp->ctr || !nl || !pos || !partial || !out || b->npad % 64 != 0 ||
      (precision != YNB_FP64 && precision != YNB_MIXED)) {
    set_error("ynb_pair_compute: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int G = pp->lanes == 0 ? 16 : pp->lanes;
  if (G != 16 && G != 32 && G != 64) {
    set_error("ynb_pair_compute: lanes must be 16, 32 or 64");
    return YNB_EINVAL;
  }
  const dim3 grid((unsigned)ceil_div(b->npad * G, 256)), block(256);
  const bool slots = gpos_slots != nullptr;
#define YNB_PAIR(P, SL, GG)                                                                                     \
  YNB_LAUNCH("pair", (k_pair<P, SL, GG>), grid, block, 0, st, b->sys, b->wave_sys, b->npad, *pp, pos, *nl,      \
             slot_mask, gpos, accumulate, gpos_slots, partial, out)
#define YNB_PAIR_G(P, SL)        \
  if (G == 16) YNB_PAIR(P, SL, 16); \
  else if (G == 32) YNB_PAIR(P, SL, 32); \
  else YNB_PAIR(P, SL, 64)
  if (precision == YNB_FP64) {
    if (slots) { YNB_PAIR_G(YNB_FP64, true); } else if (b->flags & YNB_BATCH_FORM_GENERIC) { YNB_PAIR_G(YNB_FP64, false); }
  } else {
    if (slots) { YNB_PAIR_G(YNB_MIXED, true); } else if (b->flags & YNB_BATCH_FORM_GENERIC) { YNB_PAIR_G(YNB_MIXED, false); }
  }
  if (!slots) {
#define YNB_FORMK(P, F, GG)                                                                                    \
  YNB_LAUNCH("pair", (k_pair_form<P, F, GG>), grid, block, 0, st, b->sys, b->wave_sys, b->npad, *pp, pos, *nl,  \
             slot_mask, gpos, accumulate, partial, out)
#define YNB_FORM_G(P, F)                 \
  if (G == 16) YNB_FORMK(P, F, 16);      \
  else if (G == 32) YNB_FORMK(P, F, 32); \
  else YNB_FORMK(P, F, 64)
    if (b->flags & YNB_BATCH_FORM_LJ) {
      if (precision == YNB_FP64) { YNB_FORM_G(YNB_FP64, YNB_FORM_LJ); } else { YNB_FORM_G(YNB_MIXED, YNB_FORM_LJ); }
    }
    if (b->flags & YNB_BATCH_FORM_IONIC) {
      if (precision == YNB_FP64) { YNB_FORM_G(YNB_FP64, YNB_FORM_IONIC); } else { YNB_FORM_G(YNB_MIXED, YNB_FORM_IONIC); }
    }
#undef YNB_FORM_G
#undef YNB_FORMK
  }
#undef YNB_PAIR_G
#undef YNB_PAIR
  YNB_LAUNCHED();
  return YNB_OK;
}

namespace ynb {
namespace {

// The three launches of ynb_pair_rows on the given work buffers.
int pair_rows_launch(const ynb_neigh_row *rows, int64_t nrow, const ynb_scaling_row *stab, int64_t nstab,
                     const ynb_pot *pot, const double *p0, const double *p1, const double *p2, const double *cross,
                     int32_t derivs, int64_t natom, const int32_t *atom_ref0, const int32_t *atom_ref, double *gpos,
                     double *rowh, double *partial, double *out, hipStream_t st) {
  const int64_t nblk = nrow > 0 ? ceil_div(nrow, 256) : 1;
  const bool grad = derivs && gpos != nullptr;
  YNB_LAUNCH("pair_rows", k_pair_rows, dim3((unsigned)nblk), dim3(256), 0, st, rows, nrow, stab, nstab, *pot, p0,
             p1 ? p1 : p0, p2 ? p2 : p0, cross ? cross : p0, derivs ? 1 : 0, grad ? rowh : nullptr, partial);
  YNB_LAUNCHED();
  YNB_LAUNCH("pair_rows_reduce", k_pair_rows_reduce, dim3(1), dim3(256), 0, st, nblk, partial, out);
  YNB_LAUNCHED();
  if (grad && natom > 0) {
    YNB_LAUNCH("pair_rows_gather", k_pair_rows_gather, dim3((unsigned)ceil_div(3 * natom, 256)), dim3(256), 0, st,
               natom, atom_ref0, atom_ref, rowh, gpos);
    YNB_LAUNCHED();
  }
  return YNB_OK;
}

bool pair_rows_args_ok(const ynb_neigh_row *rows, int64_t nrow, const ynb_pot *pot, const double *p0,
                       int32_t derivs, int64_t natom, const int32_t *atom_ref0, const int32_t *atom_ref,
                       double *gpos, const double *out) {
  if (!pot || nrow < 0 || nrow > (int64_t)1 << 30 || !out || (nrow > 0 && !rows) || !p0 || natom < 0) return false;
  if (derivs && gpos && (!atom_ref0 || (nrow > 0 && !atom_ref))) return false;
  return true;
}

}  // namespace
}  // namespace ynb

extern "C" int ynb_pair_rows(const ynb_neigh_row *rows, int64_t nrow, const ynb_scaling_row *stab, int64_t nstab,
                             const ynb_pot *pot, const double *p0, const double *p1, const double *p2,
                             const double *cross, int32_t derivs, int64_t natom, const int32_t *atom_ref0,
                             const int32_t *atom_ref, double *gpos, double *rowh, double *partial, double *out,
                             ynb_stream stream) {
  if (!pair_rows_args_ok(rows, nrow, pot, p0, derivs, natom, atom_ref0, atom_ref, gpos, out) || !partial ||
      (derivs && gpos && nrow > 0 && !rowh)) {
    set_error("ynb_pair_rows: invalid arguments");
    return YNB_EINVAL;
  }
  return pair_rows_launch(rows, nrow, stab, nstab, pot, p0, p1, p2, cross, derivs, natom, atom_ref0, atom_ref, gpos,
                          rowh, partial, out, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// ynb_pair_rows_static: ynb_pair_rows with its work buffers (rowh, partial) in
// a file-scope device array instead of caller memory, for callers that cannot
// allocate.  The array's host-side name is no device address: its address on
// each device comes from hipGetSymbolAddress, made once per device (with that
// device current) on the first call there, and only that address is passed to
// the kernels.  Every call on one device shares the array; nothing in it needs
// clearing (each launch writes what it reads).
static __device__ __align__(256) double scratch[3 * YNB_PAIR_ROWS_STATIC_MAX_ROWS +
                                                8 * (YNB_PAIR_ROWS_STATIC_MAX_ROWS / 256)];

namespace ynb {
namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(double **ptr) {
  int dev = -1;
  YNB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
    set_error("ynb_pair_rows_static: device %d beyond the %d devices the library tracks", dev, SCRATCH_MAX_DEVICES);
    return YNB_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("ynb_pair_rows_static: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                hipGetErrorString(e));
      return YNB_EHIP;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return YNB_OK;
}

}  // namespace
}  // namespace ynb

extern "C" int ynb_pair_rows_static(const ynb_neigh_row *rows, int64_t nrow, const ynb_scaling_row *stab,
                                    int64_t nstab, const ynb_pot *pot, const double *p0, const double *p1,
                                    const double *p2, const double *cross, int32_t derivs, int64_t natom,
                                    const int32_t *atom_ref0, const int32_t *atom_ref, double *gpos, double *out,
                                    ynb_stream stream) {
  if (!pair_rows_args_ok(rows, nrow, pot, p0, derivs, natom, atom_ref0, atom_ref, gpos, out)) {
    set_error("ynb_pair_rows_static: invalid arguments");
    return YNB_EINVAL;
  }
  if (nrow > YNB_PAIR_ROWS_STATIC_MAX_ROWS) {
    set_error("ynb_pair_rows_static: %lld rows exceed the static work array of %d rows (use ynb_pair_rows)",
              (long long)nrow, YNB_PAIR_ROWS_STATIC_MAX_ROWS);
    return YNB_ECAPACITY;
  }
  double *base = nullptr;
  const int rc = scratch_address(&base);
  if (rc != YNB_OK) return rc;
  return pair_rows_launch(rows, nrow, stab, nstab, pot, p0, p1, p2, cross, derivs, natom, atom_ref0, atom_ref, gpos,
                          base, base + 3 * (int64_t)YNB_PAIR_ROWS_STATIC_MAX_ROWS, out, (hipStream_t)stream);
}
