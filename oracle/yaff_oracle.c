/*
 * yaff_oracle.c -- CPU restatement of Yaff's FP64 nonbonded path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine in
 * yaff_amd/ is compared against; it is never linked into, loaded by or called
 * from the product.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function restates one routine of molmod/yaff (yaff/pes/*.c) with the
 * same floating-point operation order, so that built with
 * `gcc -O2 -ffp-contract=off` (x86-64 SSE2, no FMA) it reproduces the
 * reference C bit for bit.  The citation above each function names the
 * reference lines it follows.  Parity pins (tests/test_oracle.py):
 *   - bit-exact against the reference C compiled from its own sources into
 *     oracle/_ref (rows, energies, gpos, vtens) on water32, quartz, graphene8,
 *     polyethylene4, an isolated water molecule, an LJ fluid and rocksalt;
 *   - the golden numbers of the reference tests (cell mic/add_vec vectors,
 *     Switch3 exact values, the REAL/RECI/CORR Ewald tables for water32 and
 *     quartz 2x2x2, committed under tests/golden/);
 *   - mathematics: Madelung constant of rocksalt, alpha independence of the
 *     Ewald total, brute-force double loops, finite-difference gradients.
 */
#include <math.h>
#include <stddef.h>

#define ORC_TWO_PI 6.2831853071795864769
#define ORC_FOUR_PI 12.5663706143591729538
#define ORC_SQRT_PI 1.7724538509055160273
#define ORC_TWO_DIV_SQRT_PI 1.1283791670955125739

/* Same memory layout as cell_type (yaff/pes/cell.h:28-33). */
typedef struct {
  double rvecs[9], gvecs[9];
  double rspacings[3], gspacings[3];
  double volume;
  int nvec;
} orc_cell;

/* Same memory layout as neigh_row_type (yaff/pes/nlist.h:30-35), 72 bytes. */
typedef struct {
  long a, b;
  double d, dx, dy, dz;
  long r0, r1, r2;
} orc_row;

/* Same memory layout as scaling_row_type (yaff/pes/pair_pot.h:46-50). */
typedef struct {
  long a, b;
  double scale;
  long nbond;
} orc_srow;

/* Potential kinds and truncation kinds understood by orc_pair_compute. */
enum { ORC_LJ = 0, ORC_MM3 = 1, ORC_EXPREP = 2, ORC_EI = 3 };
enum { ORC_TR_NONE = 0, ORC_TR_SWITCH3 = 1, ORC_TR_HAMMER = 2 };

typedef struct {
  int kind;
  int trunc;
  double rcut;
  double trunc_par;
  /* LJ/MM3: p0 = sigmas, p1 = epsilons; EI: p0 = charges, p1 = radii */
  const double *p0, *p1;
  const int *onlypauli;        /* MM3 only */
  const long *ffatype_ids;     /* ExpRep only */
  const double *amp_cross, *b_cross;
  long nffatype;
  double alpha, dielectric;    /* EI only */
} orc_pot;

/* cell.c:37-79 */
void orc_cell_update(orc_cell *c, const double *rvecs, const double *gvecs, int nvec) {
  int i;
  double t;
  const double *r = rvecs, *g = gvecs;
  c->nvec = nvec;
  for (i = 0; i < 9; i++) {
    c->rvecs[i] = r[i];
    c->gvecs[i] = g[i];
  }
  for (i = 0; i < 3; i++) {
    c->rspacings[i] = 1.0 / sqrt(g[3*i]*g[3*i] + g[3*i+1]*g[3*i+1] + g[3*i+2]*g[3*i+2]);
    c->gspacings[i] = 1.0 / sqrt(r[3*i]*r[3*i] + r[3*i+1]*r[3*i+1] + r[3*i+2]*r[3*i+2]);
  }
  if (nvec == 0) {
    c->volume = 0.0;
  } else if (nvec == 1) {
    c->volume = sqrt(r[0]*r[0] + r[1]*r[1] + r[2]*r[2]);
  } else if (nvec == 2) {
    t = r[0]*r[3] + r[1]*r[4] + r[2]*r[5];
    t = (r[0]*r[0] + r[1]*r[1] + r[2]*r[2]) * (r[3]*r[3] + r[4]*r[4] + r[5]*r[5]) - t*t;
    c->volume = (t > 0) ? sqrt(t) : 0.0;
  } else {
    c->volume = fabs(r[0]*(r[4]*r[8] - r[5]*r[7]) +
                     r[1]*(r[5]*r[6] - r[3]*r[8]) +
                     r[2]*(r[3]*r[7] - r[4]*r[6]));
  }
}

/* cell.c:81-106: sequential per-axis wrap x = ceil(g_i.d - 1/2); d -= x r_i. */
void orc_mic(double *d, const orc_cell *c) {
  int i;
  double x;
  for (i = 0; i < c->nvec; i++) {
    const double *g = c->gvecs + 3*i, *r = c->rvecs + 3*i;
    x = ceil(g[0]*d[0] + g[1]*d[1] + g[2]*d[2] - 0.5);
    d[0] -= x*r[0];
    d[1] -= x*r[1];
    d[2] -= x*r[2];
  }
}

/* cell.c:109-121 */
void orc_to_center(const double *v, const orc_cell *c, long *center) {
  int i;
  for (i = 0; i < c->nvec; i++) {
    const double *g = c->gvecs + 3*i;
    center[i] = -ceil(g[0]*v[0] + g[1]*v[1] + g[2]*v[2] - 0.5);
  }
}

/* cell.c:124-143: d += sum_i r[i] * rvec_i, axis by axis. */
void orc_add_vec(double *d, const orc_cell *c, const long *r) {
  int i;
  for (i = 0; i < c->nvec; i++) {
    const double *rv = c->rvecs + 3*i;
    d[0] += r[i]*rv[0];
    d[1] += r[i]*rv[1];
    d[2] += r[i]*rv[2];
  }
}

/* Image odometer of nlist.c:137-180: advance r to the next image of the half
 * space; returns 0 (and resets r to the origin) once every image was seen. */
static int orc_next_image(const orc_cell *c, long *r, const long *rmax) {
  int i;
  if (c->nvec == 0) return 0;
  for (i = 0; i < c->nvec; i++) {
    r[i]++;
    if (r[i] <= rmax[i]) return 1;
    r[i] = -rmax[i];
  }
  for (i = 0; i < c->nvec; i++) r[i] = 0;
  return 0;
}

/* nlist.c:30-131 (whole build in one call, same row order).  Rows beyond
 * `cap` are counted but not written.  Returns the total number of rows. */
long orc_nlist_build(const double *pos, long natom, const orc_cell *c, double rcut,
                     const long *rmax, long nlow, long nhigh, orc_row *rows, long cap) {
  long a, b, n = 0;
  long r[3] = {0, 0, 0};
  double rc2 = rcut*rcut;
  for (a = nlow; a < natom; a++) {
    for (b = 0; b <= a; b++) {
      double d0[3];
      int more, sign;
      if (!((a >= nlow && b < nhigh) || (b >= nlow && a < nhigh))) continue;
      d0[0] = pos[3*b] - pos[3*a];
      d0[1] = pos[3*b+1] - pos[3*a+1];
      d0[2] = pos[3*b+2] - pos[3*a+2];
      orc_mic(d0, c);
      r[0] = r[1] = r[2] = 0;
      more = 1;
      while (more) {
        int image = (r[0] != 0) || (r[1] != 0) || (r[2] != 0);
        for (sign = 1; sign >= -1; sign -= 2) {
          double v[3], d2;
          if (sign < 0 && !(image && a != b)) break;
          if (!(b < a || image)) continue;
          v[0] = sign*d0[0];
          v[1] = sign*d0[1];
          v[2] = sign*d0[2];
          orc_add_vec(v, c, r);
          d2 = v[0]*v[0] + v[1]*v[1] + v[2]*v[2];
          if (d2 < rc2) {
            if (n < cap) {
              orc_row *w = rows + n;
              w->a = (sign > 0) ? a : b;
              w->b = (sign > 0) ? b : a;
              w->d = sqrt(d2);
              w->dx = v[0];
              w->dy = v[1];
              w->dz = v[2];
              w->r0 = r[0];
              w->r1 = r[1];
              w->r2 = r[2];
            }
            n++;
          }
        }
        more = orc_next_image(c, r, rmax);
      }
    }
  }
  return n;
}

/* nlist.c:183-232 (skin path: recompute the deltas of known rows). */
void orc_nlist_recompute(const double *pos, const double *pos_old, const orc_cell *c,
                         orc_row *rows, long n) {
  long i, a = -1, b = -1;
  long center[3] = {0, 0, 0};
  double d0[3] = {0, 0, 0};
  for (i = 0; i < n; i++) {
    orc_row *w = rows + i;
    double v[3];
    if (w->a != a || w->b != b) {
      a = w->a;
      b = w->b;
      d0[0] = pos_old[3*b] - pos_old[3*a];
      d0[1] = pos_old[3*b+1] - pos_old[3*a+1];
      d0[2] = pos_old[3*b+2] - pos_old[3*a+2];
      orc_to_center(d0, c, center);
      d0[0] = pos[3*b] - pos[3*a];
      d0[1] = pos[3*b+1] - pos[3*a+1];
      d0[2] = pos[3*b+2] - pos[3*a+2];
      orc_add_vec(d0, c, center);
    }
    v[0] = d0[0];
    v[1] = d0[1];
    v[2] = d0[2];
    orc_add_vec(v, c, &w->r0);
    w->d = sqrt(v[0]*v[0] + v[1]*v[1] + v[2]*v[2]);
    w->dx = v[0];
    w->dy = v[1];
    w->dz = v[2];
  }
}

/* truncation.c:59-76 */
double orc_switch3(double d, double rcut, double width, double *g) {
  double x, h;
  if (d < rcut) {
    x = rcut - d;
    if (x > width) {
      h = 1.0;
      if (g) *g = 0.0;
    } else {
      x /= width;
      h = (3 - 2*x)*x*x;
      if (g) *g = -6*x*(1-x)/width;
    }
  } else {
    h = 0.0;
    if (g) *g = 0.0;
  }
  return h;
}

/* truncation.c:30-41 */
double orc_hammer(double d, double rcut, double tau, double *g) {
  double x, h;
  if (d < rcut) {
    x = d - rcut;
    h = exp(tau/x);
    if (g) *g = -h*tau/x/x;
  } else {
    h = 0.0;
    if (g) *g = 0.0;
  }
  return h;
}

/* pair_pot.c:277-301 */
static double orc_fn_lj(const orc_pot *p, long a, long b, double d, double *g) {
  double sig = 0.5*(p->p0[a] + p->p0[b]);
  double eps = sqrt(p->p1[a]*p->p1[b]);
  double x = sig/d;
  x *= x;
  x *= x*x;
  if (g) *g = 24.0*eps/d/d*x*(1.0 - 2.0*x);
  return 4.0*eps*(x*(x - 1.0));
}

/* pair_pot.c:321-352 */
static double orc_fn_mm3(const orc_pot *p, long a, long b, double d, double *g) {
  double sig = p->p0[a] + p->p0[b];
  double eps = sqrt(p->p1[a]*p->p1[b]);
  int pauli = p->onlypauli[a] + p->onlypauli[b];
  double x = sig/d;
  double e = 1.84e5*exp(-12.0/x);
  if (pauli != 0) {
    if (g) *g = eps/d*(-12.0/sig*e);
    return eps*e;
  }
  x *= x;
  x *= 2.25*x*x;
  if (g) *g = eps/d*(-12.0/sig*e + 6.0/d*x);
  return eps*(e - x);
}

/* pair_pot.c:473-489 */
static double orc_fn_exprep(const orc_pot *p, long a, long b, double d, double *g) {
  long i = p->ffatype_ids[a]*p->nffatype + p->ffatype_ids[b];
  double amp = p->amp_cross[i], bb, e;
  if (amp == 0.0 || (bb = p->b_cross[i]) == 0.0) {
    if (g) *g = 0.0;
    return 0.0;
  }
  e = amp*exp(-bb*d);
  if (g) *g = -e*bb/d;
  return e;
}

/* pair_pot.c:711-751 */
static double orc_fn_ei(const orc_pot *p, long a, long b, double d, double *g) {
  const double *q = p->p0, *rad = p->p1;
  double qq = (q[a]*q[b])/p->dielectric;
  double rab = sqrt(rad[a]*rad[a] + rad[b]*rad[b]);
  double alpha = p->alpha, x, y, e;
  if (alpha > 0) {
    x = alpha*d;
    if (rab > 0) {
      y = d/rab;
      e = qq/d*(erfc(x) - erfc(y));
      if (g) *g = (ORC_TWO_DIV_SQRT_PI*(exp(-y*y)/rab - exp(-x*x)*alpha)*qq - e)/(d*d);
    } else {
      e = qq*erfc(x)/d;
      if (g) *g = (-ORC_TWO_DIV_SQRT_PI*alpha*exp(-x*x)*qq - e)/(d*d);
    }
  } else {
    if (rab > 0) {
      y = d/rab;
      e = qq/d*erf(y);
      if (g) *g = (ORC_TWO_DIV_SQRT_PI/rab*exp(-y*y)*qq - e)/(d*d);
    } else {
      e = qq/d;
      if (g) *g = -e/(d*d);
    }
  }
  return e;
}

static double orc_pair_fn(const orc_pot *p, long a, long b, double d, double *g) {
  switch (p->kind) {
    case ORC_LJ: return orc_fn_lj(p, a, b, d, g);
    case ORC_MM3: return orc_fn_mm3(p, a, b, d, g);
    case ORC_EXPREP: return orc_fn_exprep(p, a, b, d, g);
    default: return orc_fn_ei(p, a, b, d, g);
  }
}

static double orc_trunc(const orc_pot *p, double d, double *g) {
  if (p->trunc == ORC_TR_SWITCH3) return orc_switch3(d, p->rcut, p->trunc_par, g);
  return orc_hammer(d, p->rcut, p->trunc_par, g);
}

/* pair_pot.c:65-81: merge walk through the sorted scaling table. */
static double orc_scale_lookup(const orc_srow *st, long nst, long a, long b, long *k) {
  while (*k < nst && st[*k].a < a) (*k)++;
  if (*k >= nst || st[*k].a != a) return 1.0;
  while (st[*k].b < b) {
    (*k)++;
    if (*k >= nst || st[*k].a != a) return 1.0;
  }
  return (st[*k].b == b) ? st[*k].scale : 1.0;
}

/* pair_pot.c:84-176.  The cartesian part of the gradient (vg_cart) is
 * identically zero for LJ, MM3, ExpRep and EI; it is still carried through
 * the chain rule and the scaling so that signed zeros match the reference. */
double orc_pair_compute(const orc_row *rows, long n, const orc_srow *st, long nst,
                        const orc_pot *p, double *gpos, double *vtens) {
  long i, k = 0;
  double energy = 0.0;
  for (i = 0; i < n; i++) {
    const orc_row *w = rows + i;
    double s, v, vg, h, hg, vc0, vc1, vc2;
    if (!(w->d < p->rcut)) continue;
    if (w->r0 == 0 && w->r1 == 0 && w->r2 == 0)
      s = orc_scale_lookup(st, nst, w->a, w->b, &k);
    else
      s = 1.0;
    if (s == 0.0) continue;
    if (gpos == NULL && vtens == NULL) {
      v = orc_pair_fn(p, w->a, w->b, w->d, NULL);
      if (p->trunc != ORC_TR_NONE && v != 0.0) v *= orc_trunc(p, w->d, NULL);
    } else {
      vc0 = 0.0;
      vc1 = 0.0;
      vc2 = 0.0;
      v = orc_pair_fn(p, w->a, w->b, w->d, &vg);
      if (p->trunc != ORC_TR_NONE && (v != 0.0 || vg != 0.0)) {
        h = orc_trunc(p, w->d, &hg);
        vg = vg*h + v*hg/w->d;
        vc0 = vc0*h;
        vc1 = vc1*h;
        vc2 = vc2*h;
        v *= h;
      }
      vg *= s;
      vc0 *= s;
      vc1 *= s;
      vc2 *= s;
      if (gpos) {
        h = w->dx*vg;
        gpos[3*w->b] += h + vc0;
        gpos[3*w->a] -= h + vc0;
        h = w->dy*vg;
        gpos[3*w->b+1] += h + vc1;
        gpos[3*w->a+1] -= h + vc1;
        h = w->dz*vg;
        gpos[3*w->b+2] += h + vc2;
        gpos[3*w->a+2] -= h + vc2;
      }
      if (vtens) {
        vtens[0] += w->dx*(w->dx*vg + vc0);
        vtens[4] += w->dy*(w->dy*vg + vc1);
        vtens[8] += w->dz*(w->dz*vg + vc2);
        vtens[1] += w->dx*(w->dy*vg + vc1);
        vtens[3] += w->dy*(w->dx*vg + vc0);
        vtens[2] += w->dx*(w->dz*vg + vc2);
        vtens[6] += w->dz*(w->dx*vg + vc0);
        vtens[5] += w->dy*(w->dz*vg + vc2);
        vtens[7] += w->dz*(w->dy*vg + vc1);
      }
    }
    energy += s*v;
  }
  return energy;
}

/* ewald.c:32-150.  `work` holds 2*natom doubles when gpos != NULL. */
double orc_ewald_reci(const double *pos, long natom, long nlow, long nhigh,
                      const double *q, const orc_cell *c, double alpha, const long *gmax,
                      double gcut, double dielectric, double *gpos, double *work,
                      double *vtens) {
  long g0, g1, g2, i;
  double kv[9], energy = 0.0, fac1, fac2, dfac;
  for (i = 0; i < 9; i++) kv[i] = ORC_TWO_PI*c->gvecs[i];
  fac1 = ORC_FOUR_PI/c->volume;
  fac2 = 0.25/alpha/alpha;
  gcut *= ORC_TWO_PI;
  gcut *= gcut;
  for (g0 = -gmax[0]; g0 <= gmax[0]; g0++) {
    for (g1 = -gmax[1]; g1 <= gmax[1]; g1++) {
      for (g2 = 0; g2 <= gmax[2]; g2++) {
        double k[3], ksq, cs, sn, csl = 0, snl = 0, csh = 0, snh = 0, cc, ss, x, pre;
        if (g2 == 0 && (g1 < 0 || (g1 == 0 && g0 <= 0))) continue;
        k[0] = (g0*kv[0] + g1*kv[3] + g2*kv[6]);
        k[1] = (g0*kv[1] + g1*kv[4] + g2*kv[7]);
        k[2] = (g0*kv[2] + g1*kv[5] + g2*kv[8]);
        ksq = k[0]*k[0] + k[1]*k[1] + k[2]*k[2];
        if (ksq > gcut) continue;
        cs = 0.0;
        sn = 0.0;
        for (i = 0; i < natom; i++) {
          x = k[0]*pos[3*i] + k[1]*pos[3*i+1] + k[2]*pos[3*i+2];
          cc = q[i]*cos(x);
          ss = q[i]*sin(x);
          if (i < nlow) {
            csl += cc;
            snl += ss;
          } else if (i >= nhigh) {
            csh += cc;
            snh += ss;
          } else {
            cs += cc;
            sn += ss;
          }
          if (gpos) {
            work[2*i] = cc;
            work[2*i+1] = -ss;
          }
        }
        pre = fac1*exp(-ksq*fac2)/ksq;
        ss = cs*(cs + 2.0*(csl + csh)) + sn*(sn + 2.0*(snl + snh));
        ss += 2.0*(csl*csh + snl*snh);
        energy += pre*ss;
        if (gpos) {
          x = 2.0*pre;
          cs *= x;
          sn *= x;
          csl *= x;
          snl *= x;
          csh *= x;
          snh *= x;
          for (i = 0; i < natom; i++) {
            x = cs*work[2*i+1] + sn*work[2*i];
            if (i < nhigh) x += csh*work[2*i+1] + snh*work[2*i];
            if (i >= nlow) x += csl*work[2*i+1] + snl*work[2*i];
            gpos[3*i] += k[0]*x;
            gpos[3*i+1] += k[1]*x;
            gpos[3*i+2] += k[2]*x;
          }
        }
        if (vtens) {
          pre *= 2.0*(1.0/ksq + fac2)*ss;
          vtens[0] += pre*k[0]*k[0];
          vtens[4] += pre*k[1]*k[1];
          vtens[8] += pre*k[2]*k[2];
          x = pre*k[1]*k[0];
          vtens[1] += x;
          vtens[3] += x;
          x = pre*k[2]*k[0];
          vtens[2] += x;
          vtens[6] += x;
          x = pre*k[2]*k[1];
          vtens[5] += x;
          vtens[7] += x;
        }
      }
    }
  }
  if (vtens) {
    vtens[0] -= energy;
    vtens[4] -= energy;
    vtens[8] -= energy;
  }
  dfac = 1.0/dielectric;
  if (gpos)
    for (i = 0; i < 3*natom; i++) gpos[i] *= dfac;
  if (vtens)
    for (i = 0; i < 9; i++) vtens[i] *= dfac;
  return energy*dfac;
}

/* ewald.c:278-349 */
double orc_ewald_corr(const double *pos, const double *q, const orc_cell *c, double alpha,
                      const orc_srow *st, long nst, double dielectric, double *gpos,
                      double *vtens, long natom, long nlow, long nhigh) {
  long i;
  double energy = 0.0, x, g = 0.0, dfac;
  x = alpha/ORC_SQRT_PI;
  for (i = nlow; i < nhigh; i++) energy -= x*q[i]*q[i];
  for (i = 0; i < nst; i++) {
    long a = st[i].a, b = st[i].b;
    double dl[3], d, e, f;
    if (a < nlow && b < nlow) continue;
    if (a >= nhigh && b >= nhigh) continue;
    dl[0] = pos[3*b] - pos[3*a];
    dl[1] = pos[3*b+1] - pos[3*a+1];
    dl[2] = pos[3*b+2] - pos[3*a+2];
    orc_mic(dl, c);
    d = sqrt(dl[0]*dl[0] + dl[1]*dl[1] + dl[2]*dl[2]);
    x = alpha*d;
    e = erf(x)/d;
    f = (1 - st[i].scale)*q[b]*q[a];
    if (gpos || vtens) g = -f*(ORC_TWO_DIV_SQRT_PI*alpha*exp(-x*x) - e)/d/d;
    if (gpos) {
      x = dl[0]*g;
      gpos[3*b] += x;
      gpos[3*a] -= x;
      x = dl[1]*g;
      gpos[3*b+1] += x;
      gpos[3*a+1] -= x;
      x = dl[2]*g;
      gpos[3*b+2] += x;
      gpos[3*a+2] -= x;
    }
    if (vtens) {
      vtens[0] += dl[0]*dl[0]*g;
      vtens[4] += dl[1]*dl[1]*g;
      vtens[8] += dl[2]*dl[2]*g;
      x = dl[1]*dl[0]*g;
      vtens[1] += x;
      vtens[3] += x;
      x = dl[2]*dl[0]*g;
      vtens[2] += x;
      vtens[6] += x;
      x = dl[2]*dl[1]*g;
      vtens[5] += x;
      vtens[7] += x;
    }
    energy -= f*e;
  }
  dfac = 1.0/dielectric;
  if (gpos)
    for (i = 0; i < 3*natom; i++) gpos[i] *= dfac;
  if (vtens)
    for (i = 0; i < 9; i++) vtens[i] *= dfac;
  return energy*dfac;
}
