// Pair potentials evaluated per pair, for the list-driven pair kernels
// (pair.hip):
// Yaff's pair functions (yaff/pes/pair_pot.c:277-352, 473-489, 711-751) and
// truncations (yaff/pes/truncation.c:30-95) in FP64, their FP32 forms for
// mixed precision, and one entry of a specialised potential set
// (YNB_FORM_LJ, YNB_FORM_IONIC) with its exact treatment near the cutoffs.
#pragma once

#include "common.h"
#include "erfcx_table.h"

namespace ynb {
namespace {

constexpr double TWO_DIV_SQRT_PI = 1.1283791670955125739;  // yaff/pes/constants.h:31

// Transcendentals of the two precisions.
//
// FP64: exp_fp64 (Cody-Waite, Taylor to r^13); erfc(x) = exp(-x^2) erfcx(x) with erfcx
// a degree-24 polynomial in y = 4/(2 + x) - 1 (relative error < 3e-14 on
// [0, 27], table generated by gen_erfcx_table.py), so that erfc and the EI
// gradient share one exp(-x^2).
//
// Mixed: every argument reduction in FP64, only a well-conditioned
// remainder in FP32:
//   exp(a)   = 2^k * exp2f(f),  k = rint(a log2 e), f = a log2 e - k in FP64;
//   erfc(x)  = exp(-x^2) * erfcxf(x), erfcx = exp(x^2) erfc(x) being smooth,
//              so rounding x to FP32 barely moves it.
__device__ inline double erfcx_poly(double x) {  // x >= 0
  // y = 4 / (2 + x) - 1; p(y) = E(y^2) + y O(y^2), the even and odd parts by
  // Horner's rule in y^2 (two independent chains of 12 and 11 steps)
  const double y = fma(4.0, rcp_fp64(2.0 + x), -1.0);
  const double z = y * y;
  double e = ERFCX_MONO[ERFCX_NCOEF - 1];
  double o = ERFCX_MONO[ERFCX_NCOEF - 2];
#pragma unroll
  for (int j = (ERFCX_NCOEF - 3) / 2; j >= 0; --j) {
    e = fma3(e, z, ERFCX_MONO[2 * j]);
    if (2 * j + 1 < ERFCX_NCOEF - 2) o = fma3(o, z, ERFCX_MONO[2 * j + 1]);
  }
  return fma(y, o, e);
}

template <int PREC>
struct Tr {
  __device__ static double ex(double a) { return exp_fp64(a); }
  // erfc(x) given e2 = exp(-x^2)
  __device__ static double erfc_e2(double x, double e2) {
    return x >= 0.0 ? e2 * erfcx_poly(x) : 2.0 - e2 * erfcx_poly(-x);
  }
  __device__ static double erfc_(double x) { return erfc_e2(x, exp_fp64(-x * x)); }
  __device__ static double erf_(double x) { return 1.0 - erfc_(x); }
  // erfc(x) for x >= 0 given e2 = exp(-x^2)
  __device__ static double erfc_pos(double x, double e2) { return e2 * erfcx_poly(x); }
};
template <>
struct Tr<YNB_MIXED> {
  __device__ static double ex(double a) {
    const double t = a * 1.4426950408889634074;  // log2(e)
    const double k = rint(t);
    return ldexp((double)exp2f((float)(t - k)), (int)k);
  }
  __device__ static double erfc_(double x) { return ex(-x * x) * (double)erfcxf((float)x); }
  __device__ static double erf_(double x) { return 1.0 - erfc_(x); }
  __device__ static double erfc_e2(double x, double e2) { return e2 * (double)erfcxf((float)x); }
  __device__ static double erfc_pos(double x, double e2) { return e2 * (double)erfcxf((float)x); }
};

template <int PREC>
struct PotEval {
  using R = Tr<PREC>;
  // Returns the energy; g receives dE/dd / d (Yaff's vg).
  __device__ static double lj(double sa, double sb, double ea, double eb, double d, double &g) {  // pair_pot.c:277-301
    const double sigma = 0.5 * (sa + sb);
    const double eps = sqrt(ea * eb);
    double x = sigma / d;
    x *= x;
    x *= x * x;
    g = 24.0 * eps / d / d * x * (1.0 - 2.0 * x);
    return 4.0 * eps * (x * (x - 1.0));
  }
  __device__ static double mm3(double sa, double sb, double ea, double eb, int pauli, double d,
                               double &g) {  // pair_pot.c:321-352
    const double sigma = sa + sb;
    const double eps = sqrt(ea * eb);
    double x = sigma / d;
    const double e = 1.84e5 * R::ex(-12.0 / x);
    if (pauli != 0) {
      g = eps / d * (-12.0 / sigma * e);
      return eps * e;
    }
    x *= x;
    x *= 2.25 * x * x;
    g = eps / d * (-12.0 / sigma * e + 6.0 / d * x);
    return eps * (e - x);
  }
  __device__ static double exprep(double amp, double b, double d, double &g) {  // pair_pot.c:473-489
    if (amp == 0.0 || b == 0.0) {
      g = 0.0;
      return 0.0;
    }
    const double e = amp * R::ex(-b * d);
    g = -e * b / d;
    return e;
  }
  __device__ static double ei(double qa, double qb, double ra, double rb, double alpha, double diel, double d,
                              double &g) {  // pair_pot.c:711-751
    const double qq = (qa * qb) / diel;
    const double rab = sqrt(ra * ra + rb * rb);
    double e;
    if (alpha > 0.0) {
      const double x = alpha * d;
      if (rab > 0.0) {
        const double y = d / rab;
        e = qq / d * (R::erfc_(x) - R::erfc_(y));
        g = (TWO_DIV_SQRT_PI * (R::ex(-y * y) / rab - R::ex(-x * x) * alpha) * qq - e) / (d * d);
      } else {
        e = qq * R::erfc_(x) / d;
        g = (-TWO_DIV_SQRT_PI * alpha * R::ex(-x * x) * qq - e) / (d * d);
      }
    } else {
      if (rab > 0.0) {
        const double y = d / rab;
        e = qq / d * R::erf_(y);
        g = (TWO_DIV_SQRT_PI / rab * R::ex(-y * y) * qq - e) / (d * d);
      } else {
        e = qq / d;
        g = -e / (d * d);
      }
    }
    return e;
  }
  __device__ static double switch3(double d, double rcut, double w, double &hg) {  // truncation.c:59-76
    if (d < rcut) {
      double x = rcut - d;
      if (x > w) {
        hg = 0.0;
        return 1.0;
      }
      x /= w;
      hg = -6 * x * (1 - x) / w;
      return (3 - 2 * x) * x * x;
    }
    hg = 0.0;
    return 0.0;
  }
  __device__ static double hammer(double d, double rcut, double tau, double &hg) {  // truncation.c:30-41
    if (d < rcut) {
      const double x = d - rcut;
      const double h = R::ex(tau / x);
      hg = -h * tau / x / x;
      return h;
    }
    hg = 0.0;
    return 0.0;
  }
};

// The same potentials for the neighbour-list kernel, written around one
// reciprocal distance inv = 1/d per pair (the reference divides by d in every
// term).  LJ takes sqrt(epsilon) per atom (precomputed on the host), so
// eps_ab = se_a * se_b instead of sqrt(eps_a * eps_b); EI takes 1/dielectric
// and shares exp(-alpha^2 d^2) between erfc and the gradient.  These change
// roundings only (relative 1e-16 per term).
template <int PREC>
struct PotFast {
  using R = Tr<PREC>;
  __device__ static double lj(double sa, double sb, double eps, double inv, double inv2, double &g) {
    const double sigma = 0.5 * (sa + sb);
    double x = sigma * inv;
    x *= x;
    x *= x * x;
    g = 24.0 * eps * inv2 * x * (1.0 - 2.0 * x);
    return 4.0 * eps * (x * (x - 1.0));
  }
  __device__ static double exprep(double amp, double b, double d, double inv, double &g) {
    if (amp == 0.0 || b == 0.0) {
      g = 0.0;
      return 0.0;
    }
    const double e = amp * R::ex(-b * d);
    g = -e * b * inv;
    return e;
  }
  // alpha > 0, point charges (the common case); qq = qa qb / dielectric.
  // x = alpha d > 0, so erfc needs only its non-negative branch.
  __device__ static double ei_point(double qq, double alpha, double d, double inv, double inv2, double &g) {
    const double x = alpha * d;
    const double e2 = R::ex(-x * x);
    const double e = qq * R::erfc_pos(x, e2) * inv;
    g = (-TWO_DIV_SQRT_PI * alpha * e2 * qq - e) * inv2;
    return e;
  }
};

// FP32 pieces of the mixed-precision forms.
// LJ with eps_ab = sqrt(eps_a) sqrt(eps_b) and inv = 1/d (pair_pot.c:277-301).
__device__ inline float lj32(float sa, float sb, float eps, float inv, float inv2, float &g) {
  const float x = 0.5f * (sa + sb) * inv;
  const float x2 = x * x;
  const float x6 = x2 * x2 * x2;
  g = 24.0f * eps * inv2 * x6 * (1.0f - 2.0f * x6);
  return 4.0f * eps * (x6 * (x6 - 1.0f));
}
// Switch3 (truncation.c:59-76) for d < rcut.
__device__ inline float switch3_32(float d, float rcut, float w, float &hg) {
  float x = rcut - d;
  if (x > w) {
    hg = 0.0f;
    return 1.0f;
  }
  const float rw = __builtin_amdgcn_rcpf(w);
  x *= rw;
  hg = -6.0f * x * (1.0f - x) * rw;
  return (3.0f - 2.0f * x) * x * x;
}
// erfcx(x) for 0 <= x <= ERFCX32_XMAX: polynomial in z = (t - 5/8) 8/3,
// t = 2 / (2 + x) (relative error below 1e-6, erfcx_table.h).
__device__ inline float erfcx32(float x) {
  const float t = 2.0f * __builtin_amdgcn_rcpf(2.0f + x);
  const float z = (t - 0.625f) * 2.66666666666666667f;
  float r = ERFCX32_MONO[ERFCX32_NCOEF - 1];
#pragma unroll
  for (int k = ERFCX32_NCOEF - 2; k >= 0; --k) r = fmaf(r, z, ERFCX32_MONO[k]);
  return r;
}

// Per-system constants of a specialised potential set (uniform per block).
struct FormConst {
  double rc0, rc1, rc2, rcmax, tol0, tol1, tol2, alpha, w0;
  double rc0sq, rc1sq, rc2sq, tol0sq, tol1sq, tol2sq, alpha2;  // mixed precision: squared cutoffs and bands
  float rc0f, w0f, alphaf;
  int trunc0, nf;
  const double *amp, *bex;        // ExpRep tables (global)
  const double *l_amp, *l_bex;    // the same in LDS when lds_cross
  const float *l_ampf, *l_bexf;
  bool lds_cross;
};

template <int FORM>
__device__ inline FormConst form_const(const ynb_sys &S, const double *cross) {
  FormConst C;
  C.rc0 = S.pot[0].rcut;
  C.rc1 = FORM == YNB_FORM_IONIC ? S.pot[1].rcut : 0.0;
  C.rc2 = FORM == YNB_FORM_IONIC ? S.pot[2].rcut : 0.0;
  C.rcmax = fmax(C.rc0, fmax(C.rc1, C.rc2));
  C.tol0 = 1e-12 * C.rc0;
  C.tol1 = 1e-12 * C.rc1;
  C.tol2 = 1e-12 * C.rc2;
  C.alpha = FORM == YNB_FORM_IONIC ? S.pot[2].alpha : 0.0;
  C.trunc0 = S.pot[0].trunc;
  C.w0 = S.pot[0].trunc_par;
  C.rc0sq = C.rc0 * C.rc0;
  C.rc1sq = C.rc1 * C.rc1;
  C.rc2sq = C.rc2 * C.rc2;
  C.tol0sq = 4e-12 * C.rc0sq;
  C.tol1sq = 4e-12 * C.rc1sq;
  C.tol2sq = 4e-12 * C.rc2sq;
  C.alpha2 = C.alpha * C.alpha;
  C.rc0f = (float)C.rc0;
  C.w0f = (float)C.w0;
  C.alphaf = (float)C.alpha;
  C.nf = FORM == YNB_FORM_IONIC ? S.pot[0].nffatype : 0;
  C.amp = cross + (FORM == YNB_FORM_IONIC ? S.pot[0].cross0 : 0);
  C.bex = C.amp + (int64_t)C.nf * C.nf;
  C.l_amp = C.l_bex = nullptr;
  C.l_ampf = C.l_bexf = nullptr;
  C.lds_cross = false;
  return C;
}

// One entry (i, j, image n) of a specialised set.  u holds pos_j - pos_i +
// n.R (FP64); where a distance lies within 1e-12 of a cutoff the vector is
// rebuilt exactly as nlist_build_low does (yaff_row_vector, positions from
// pos), so the cutoff tests d < rcut see Yaff's bits, and u is replaced by
// it.  Returns false when the pair is beyond every cutoff; otherwise the
// energy and Yaff's vg (dE/dd / d) summed over the set's potentials, in
// FP64: one reciprocal distance per pair (device rsqrt, a few ulp); ExpRep,
// LJ (eps from sqrt(eps) per atom) and EI (alpha > 0, point charges,
// q / sqrt(dielectric) per atom) as PotFast.  ri/rj: per-atom records (LJ
// sigma, sqrt(LJ eps), charge / sqrt(dielectric), ExpRep type).  Mixed
// precision: form_geom_mixed and form_pot_mixed below.
template <int PREC, int FORM>
__device__ inline bool form_pair(const ynb_sys &S, const FormConst &C, int li, int64_t gi, int64_t gj,
                                 const double pi[3], const double *pos, const int nn[3], const double4 &ri,
                                 const double4 &rj, int ti, double u[3], double &e_sum, double &vg_sum) {
  using PE = PotEval<PREC>;
  using PF = PotFast<PREC>;
  auto exact = [&](double v[3]) {  // Yaff's own row vector (no scaling rows in a form)
    const int lj = (int)(gj - S.atom0);
    const bool self = (gj == gi);
    const bool i_is_a = li > lj || (li == lj);
    const int nab[3] = {i_is_a ? nn[0] : -nn[0], i_is_a ? nn[1] : -nn[1], i_is_a ? nn[2] : -nn[2]};
    const double pj[3] = {pos[3 * gj], pos[3 * gj + 1], pos[3 * gj + 2]};
    int r[3];
    yaff_row_vector(i_is_a ? pi : pj, i_is_a ? pj : pi, self, nab, S.rvecs, S.gvecs, S.nvec, S.rmax, v, r);
    u[0] = i_is_a ? v[0] : -v[0];
    u[1] = i_is_a ? v[1] : -v[1];
    u[2] = i_is_a ? v[2] : -v[2];
  };
  e_sum = 0.0;
  vg_sum = 0.0;
  static_assert(PREC == YNB_FP64, "form_pair evaluates FP64 entries");
  {
    // 1/d from the device rsqrt (a few ulp) and d = r2 / d: every decision
    // that depends on the last bits of d is made on the exact path
    const double r2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    double inv = rsqrt(r2);
    double dd = r2 * inv;
    bool near_cut = fabs(dd - C.rc0) <= C.tol0;
    if (FORM == YNB_FORM_IONIC) near_cut = near_cut || fabs(dd - C.rc1) <= C.tol1 || fabs(dd - C.rc2) <= C.tol2;
    if (near_cut) {
      double v[3];
      exact(v);
      dd = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      inv = 1.0 / dd;
    }
    if (!(dd < C.rcmax)) return false;
    const double d = dd;
    const double inv2 = inv * inv;
    if (FORM == YNB_FORM_LJ) {
      if (d < C.rc0) {
        double g;
        double e = PF::lj(ri.x, rj.x, ri.y * rj.y, inv, inv2, g);
        if (C.trunc0 == YNB_TR_SWITCH3 && (e != 0.0 || g != 0.0)) {
          double hg;
          const double h = PE::switch3(d, C.rc0, C.w0, hg);
          g = g * h + e * hg * inv;
          e *= h;
        }
        e_sum = e;
        vg_sum = g;
      }
    } else {
      if (d < C.rc0) {  // ExpRep
        const int ci = ti * C.nf + (int)rj.w;
        double g;
        e_sum += PF::exprep(C.lds_cross ? C.l_amp[ci] : C.amp[ci], C.lds_cross ? C.l_bex[ci] : C.bex[ci], d, inv, g);
        vg_sum += g;
      }
      if (d < C.rc1) {  // LJ
        double g;
        e_sum += PF::lj(ri.x, rj.x, ri.y * rj.y, inv, inv2, g);
        vg_sum += g;
      }
      if (d < C.rc2) {  // EI
        double g;
        e_sum += PF::ei_point(ri.z * rj.z, C.alpha, d, inv, inv2, g);
        vg_sum += g;
      }
    }
    return true;
  }
}

// One mixed-precision entry of a specialised set in two phases, so that a
// lane can evaluate two entries with their arithmetic interleaved (the FP64
// kernels use form_pair, one entry at a time):
//
// form_geom_mixed: u holds pos_j - pos_i + n.R (FP64); where a squared
// distance lies within a relative 4e-12 of a squared cutoff the vector is
// rebuilt exactly as nlist_build_low does (yaff_row_vector) and u replaced
// by it.  Fills r^2 (FP64) and the flags of the set's cutoffs, decided in
// FP64 as form_pair decides them; returns false when the pair is beyond
// every cutoff.
//
// form_pot_mixed: the energy and Yaff's vg (dE/dd / d) summed over the set's
// potentials in FP32 (the arithmetic of form_pair's mixed branch), as
// straight-line code: every potential is evaluated and those beyond their
// cutoff (or with a zero ExpRep parameter) are selected away, which gives
// form_pair's sums bit for bit.
struct PairGeom {
  double r2;  // squared distance (FP64)
  bool in0, in1, in2;
};

template <int FORM>
__device__ inline bool form_geom_mixed(const ynb_sys &S, const FormConst &C, int li, int64_t gi, int64_t gj,
                                       const double pi[3], const double *pos, const int nn[3], double u[3],
                                       PairGeom &Gm) {
  double r2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  bool near_cut = fabs(r2 - C.rc0sq) <= C.tol0sq;
  if (FORM == YNB_FORM_IONIC)
    near_cut = near_cut || fabs(r2 - C.rc1sq) <= C.tol1sq || fabs(r2 - C.rc2sq) <= C.tol2sq;
  if (near_cut) {  // Yaff's own row vector (no scaling rows in a form)
    const int lj = (int)(gj - S.atom0);
    const bool self = (gj == gi);
    const bool i_is_a = li > lj || (li == lj);
    const int nab[3] = {i_is_a ? nn[0] : -nn[0], i_is_a ? nn[1] : -nn[1], i_is_a ? nn[2] : -nn[2]};
    const double pj[3] = {pos[3 * gj], pos[3 * gj + 1], pos[3 * gj + 2]};
    double v[3];
    int r[3];
    yaff_row_vector(i_is_a ? pi : pj, i_is_a ? pj : pi, self, nab, S.rvecs, S.gvecs, S.nvec, S.rmax, v, r);
    u[0] = i_is_a ? v[0] : -v[0];
    u[1] = i_is_a ? v[1] : -v[1];
    u[2] = i_is_a ? v[2] : -v[2];
    r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const double dd = sqrt(r2);
    Gm.in0 = dd < C.rc0;
    Gm.in1 = dd < C.rc1;
    Gm.in2 = dd < C.rc2;
  } else {
    Gm.in0 = r2 < C.rc0sq;
    Gm.in1 = r2 < C.rc1sq;
    Gm.in2 = r2 < C.rc2sq;
  }
  Gm.r2 = r2;
  return Gm.in0 || (FORM == YNB_FORM_IONIC && (Gm.in1 || Gm.in2));
}

template <int FORM>
__device__ inline void form_pot_mixed(const FormConst &C, const PairGeom &Gm, const double4 &ri, const double4 &rj,
                                      int ti, double &e_sum, double &vg_sum) {
  const double r2 = Gm.r2;
  const float r2f = (float)r2;
  const float inv = __builtin_amdgcn_rsqf(r2f);
  const float inv2 = inv * inv;
  const float df = r2f * inv;
  float ef = 0.0f, gf = 0.0f;
  if (FORM == YNB_FORM_LJ) {
    float g;
    float e = lj32((float)ri.x, (float)rj.x, (float)(ri.y * rj.y), inv, inv2, g);
    if (C.trunc0 == YNB_TR_SWITCH3) {  // switch3_32 as selects
      const float xr = C.rc0f - df;
      const float rw = __builtin_amdgcn_rcpf(C.w0f);
      const float x = xr * rw;
      const float hgp = -6.0f * x * (1.0f - x) * rw;
      const float hp = (3.0f - 2.0f * x) * x * x;
      const bool flat = xr > C.w0f;
      const float h = flat ? 1.0f : hp, hg = flat ? 0.0f : hgp;
      g = g * h + e * hg * inv;
      e *= h;
    }
    ef = e;
    gf = g;
  } else {
    {  // ExpRep
      const int ci = ti * C.nf + (int)rj.w;
      const float a = C.lds_cross ? C.l_ampf[ci] : (float)C.amp[ci];
      const float b = C.lds_cross ? C.l_bexf[ci] : (float)C.bex[ci];
      const float e = a * __builtin_amdgcn_exp2f(-b * df * 1.44269504088896341f);
      const bool on = Gm.in0 && a != 0.0f && b != 0.0f;
      ef += on ? e : 0.0f;
      gf -= on ? e * b * inv : 0.0f;
    }
    {  // LJ
      float g;
      const float e = lj32((float)ri.x, (float)rj.x, (float)(ri.y * rj.y), inv, inv2, g);
      ef += Gm.in1 ? e : 0.0f;
      gf += Gm.in1 ? g : 0.0f;
    }
    {  // EI, point charges, alpha > 0: exp(-x^2) with x^2 = alpha^2 r^2 reduced in FP64
      const double a2r2 = C.alpha2 * r2;
      const double tq = a2r2 * -1.4426950408889634074;
      const double kq = rint(tq);
      const float e2 = __builtin_amdgcn_ldexpf(__builtin_amdgcn_exp2f((float)(tq - kq)), (int)kq);
      const float x = fminf(__builtin_amdgcn_sqrtf((float)a2r2), ERFCX32_XMAX);
      const float qq = (float)(ri.z * rj.z);
      const float e = qq * (e2 * erfcx32(x)) * inv;
      ef += Gm.in2 ? e : 0.0f;
      gf += Gm.in2 ? (-1.12837916709551257f * C.alphaf * e2 * qq - e) * inv2 : 0.0f;
    }
  }
  e_sum = (double)ef;
  vg_sum = (double)gf;
}

}  // namespace
}  // namespace ynb
