// Velocity Verlet halves on device (yaff/sampling/verlet.py:339-367), totals
// of a force-field evaluation, and the library's error/version plumbing.
#include <stdarg.h>
#include <string.h>

#include "common.h"
#include "ewald_prep.h"
#include "nlist_work.h"

namespace ynb {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

namespace {

constexpr int PROF_MAX = 8192;
struct Prof {
  bool on = false;
  int n = 0;
  const char *name[PROF_MAX];
  hipEvent_t ev[2 * PROF_MAX];
  bool created = false;
} g_prof;

}  // namespace

void prof_before(const char *name, hipStream_t st) {
  if (!g_prof.on || g_prof.n >= PROF_MAX) return;
  g_prof.name[g_prof.n] = name;
  hipEventRecord(g_prof.ev[2 * g_prof.n], st);
}

void prof_after(hipStream_t st) {
  if (!g_prof.on || g_prof.n >= PROF_MAX) return;
  hipEventRecord(g_prof.ev[2 * g_prof.n + 1], st);
  ++g_prof.n;
}

namespace {

// acc = -gpos/m; vel += dt/2 acc; pos += dt vel  (verlet.py:344-347); then,
// with frac, the Ewald set-up of the atom at its new position, and with
// bin_count its cell binning for the next neighbour-list build.
__global__ __launch_bounds__(256) void k_kick_drift(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad,
                                                    double *pos, double *vel, const double *gpos, const double *minv,
                                                    const double *dt, const double *charges, double *frac,
                                                    double *step, int32_t cb, int32_t *bin_count, int32_t *slot_idx,
                                                    double4 *xw, int4 *wr, int32_t *status) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= npad) return;  // whole waves: npad is a multiple of 64
  const int sid = uniform_int(wave_sys[g >> 6]);
  const ynb_sys *S = sid >= 0 ? &sys[sid] : nullptr;
  const bool active = S && g - S->atom0 < S->natom;
  double p[3] = {0.0, 0.0, 0.0};
  if (active) {
    const double h = dt[sid];
    const double mi = minv[g];
    for (int c = 0; c < 3; ++c) {
      const double a = -gpos[3 * g + c] * mi;
      const double v = vel[3 * g + c] + 0.5 * a * h;
      vel[3 * g + c] = v;
      p[c] = pos[3 * g + c] + h * v;
      pos[3 * g + c] = p[c];
    }
  }
  if (frac) ewald_prep_atom(S, active, g, npad, p[0], p[1], p[2], charges, frac, step);
  if (bin_count && S) bin_atom(*S, active, g, p[0], p[1], p[2], cb, bin_count, slot_idx, xw, wr, status);
}

// Sum of ns > 0 gradient slots src[0], src[stride], ...: slot l goes to
// accumulator l mod 4, the accumulators are added in a fixed order.  Eight
// loads per round (slots past ns read slot ns - 1, a valid address, and add
// nothing), so a partial last round costs one round trip, not one per slot.
__device__ inline double sum_slots(const double *src, int64_t stride, int ns) {
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int sl = 0; sl < ns; sl += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[(int64_t)min(sl + u, ns - 1) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (sl + u < ns) a[u & 3] += v[u];
  }
  return (a[0] + a[1]) + (a[2] + a[3]);
}

// Term c of system s's total: its potentials' pair terms in order, then the
// Ewald reciprocal term.
__device__ inline void system_total(const ynb_sys &S, int s, int c, const double *pair_out, const double *ewald_out,
                                    double *total) {
  double acc = 0.0;
  if (pair_out)
    for (int p = 0; p < S.npot; ++p) acc += pair_out[((int64_t)s * YNB_MAX_POT + p) * 10 + c];
  if (ewald_out) acc += ewald_out[(int64_t)s * 10 + c];
  total[(int64_t)s * 10 + c] = acc;
}

// vel += dt/2 acc (verlet.py:354-355) and 2 Ekin of each system.  Blocks of
// 64 atoms x 3 components (one thread per component; systems start at
// multiples of 64 atoms, so a block belongs to the system of its wave entry);
// each block's sum of m v^2 goes to `partial` and the system's last block
// adds them in a fixed order.  With gpart, the gradient is gpos + the
// system's nslot Ewald partial gradients (slots of 3 npad doubles,
// sum_slots) and is written
// back to gpos.  With total, the last block also totals the system's energy
// and virial terms (system_total).
__global__ __launch_bounds__(192) void k_kick(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad,
                                              double *vel, double *gpos, const double *gpart,
                                              const double *minv, const double *masses, const double *dt,
                                              double *partial, int32_t *ctr, double *ekin2,
                                              const double *pair_out, const double *ewald_out, double *total) {
  const int64_t blk = blockIdx.x;
  const int sid = wave_sys[blk];
  if (sid < 0) return;  // the same for the whole block
  const ynb_sys &S = sys[sid];
  const int t = threadIdx.x;
  const int64_t g = blk * 64 + t / 3;
  const int64_t e = blk * 192 + t;  // = 3 g + c
  double e2 = 0.0;
  if (g - S.atom0 < S.natom) {
    double gr = gpos[e];
    const int ns = S.nslot;
    if (gpart && ns > 0) {
      const double *src = gpart + e;
      const int64_t stride = 3 * npad;
      gr += sum_slots(src, stride, ns);
      gpos[e] = gr;
    }
    const double v = vel[e] + 0.5 * (-gr * minv[g]) * dt[sid];
    vel[e] = v;
    e2 = masses[g] * v * v;
  }
  __shared__ double l_w[3];
  e2 = wave_sum(e2);
  if ((t & 63) == 0) l_w[t >> 6] = e2;
  __syncthreads();
  if (t == 0) publish(partial + blk, (l_w[0] + l_w[1]) + l_w[2]);
  const int nb = S.natom > 0 ? (S.natom + 63) / 64 : 1;
  if (last_block(ctr + sid, nb)) {
    const int64_t b0 = S.atom0 / 64;
    double acc = 0.0;
    for (int w = t; w < nb; w += 192) acc += gather(partial + b0 + w);
    acc = wave_sum(acc);
    __syncthreads();
    if ((t & 63) == 0) l_w[t >> 6] = acc;
    __syncthreads();
    if (t == 0) ekin2[sid] = (l_w[0] + l_w[1]) + l_w[2];
    if (total && t < 10) system_total(S, sid, t, pair_out, ewald_out, total);
  }
}

// The kick that ends one step and the kick and drift that start the next
// (verlet.py:354-355, then 344-347), one kernel: the k_kick block layout and
// arithmetic (gradient slots summed into gpos, vel += dt/2 acc, 2 Ekin by the
// system's last block, energy/virial totals), then vel += dt/2 acc once more
// and pos += dt vel with the k_kick_drift expressions, so every value is
// bitwise the one the two kernels give; then, per atom (the block's new
// positions through LDS), the Ewald set-up and the binning of k_kick_drift.
__global__ __launch_bounds__(192) void k_kick_kick_drift(
    const ynb_sys *sys, const int32_t *wave_sys, int64_t npad, double *pos, double *vel, double *gpos,
    const double *gpart, const double *minv, const double *masses, const double *dt, double *partial, int32_t *ctr,
    double *ekin2, const double *pair_out, const double *ewald_out, double *total, const double *charges, double *frac,
    double *step, int32_t cb, int32_t *bin_count, int32_t *slot_idx, double4 *xw, int4 *wr, int32_t *status) {
  const int64_t blk = blockIdx.x;
  const int sid = wave_sys[blk];  // the same for the whole block
  const ynb_sys *S = sid >= 0 ? &sys[sid] : nullptr;
  const int t = threadIdx.x;
  const int64_t g = blk * 64 + t / 3;
  const int64_t e = blk * 192 + t;  // = 3 g + c
  __shared__ double l_p[192];
  __shared__ double l_w[3];
  l_p[t] = 0.0;
  if (S) {
    double e2 = 0.0;
    if (g - S->atom0 < S->natom) {
      double gr = gpos[e];
      const int ns = S->nslot;
      if (gpart && ns > 0) {
        const double *src = gpart + e;
        const int64_t stride = 3 * npad;
        gr += sum_slots(src, stride, ns);
        gpos[e] = gr;
      }
      const double h = dt[sid];
      const double mi = minv[g];
      const double v = vel[e] + 0.5 * (-gr * mi) * h;  // end of this step
      e2 = masses[g] * v * v;
      const double a = -gr * mi;
      const double v2 = v + 0.5 * a * h;  // start of the next
      vel[e] = v2;
      const double p = pos[e] + h * v2;
      pos[e] = p;
      l_p[t] = p;
    }
    e2 = wave_sum(e2);
    if ((t & 63) == 0) l_w[t >> 6] = e2;
    __syncthreads();
    if (t == 0) publish(partial + blk, (l_w[0] + l_w[1]) + l_w[2]);
    const int nb = S->natom > 0 ? (S->natom + 63) / 64 : 1;
    if (last_block(ctr + sid, nb)) {
      const int64_t b0 = S->atom0 / 64;
      double acc = 0.0;
      for (int w = t; w < nb; w += 192) acc += gather(partial + b0 + w);
      acc = wave_sum(acc);
      __syncthreads();
      if ((t & 63) == 0) l_w[t >> 6] = acc;
      __syncthreads();
      if (t == 0) ekin2[sid] = (l_w[0] + l_w[1]) + l_w[2];
      if (total && t < 10) system_total(*S, sid, t, pair_out, ewald_out, total);
    }
  }
  __syncthreads();
  if (t < 64) {
    const int64_t ga = blk * 64 + t;
    const bool active = S && ga - S->atom0 < S->natom;
    const double p0 = l_p[3 * t], p1 = l_p[3 * t + 1], p2 = l_p[3 * t + 2];
    if (frac) ewald_prep_atom(S, active, ga, npad, p0, p1, p2, charges, frac, step);
    if (bin_count && S) bin_atom(*S, active, ga, p0, p1, p2, cb, bin_count, slot_idx, xw, wr, status);
  }
}

__global__ void k_total(const ynb_sys *sys, int nsys, const double *pair_out, const double *ewald_out,
                        double *total) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  for (int c = 0; c < 10; ++c) system_total(sys[s], s, c, pair_out, ewald_out, total);
}

}  // namespace
}  // namespace ynb

using namespace ynb;

extern "C" {

const char *ynb_version(void) { return "yaff_nb 0.1 gfx950"; }

const char *ynb_last_error(void) { return g_err; }

int ynb_profile_enable(int on) {
  if (on && !g_prof.created) {
    for (int i = 0; i < 2 * PROF_MAX; ++i) YNB_HIP(hipEventCreate(&g_prof.ev[i]));
    g_prof.created = true;
  }
  g_prof.on = on != 0;
  g_prof.n = 0;
  return YNB_OK;
}

int ynb_profile_count(void) { return g_prof.n; }

int ynb_profile_read(int i, const char **name, float *ms) {
  if (i < 0 || i >= g_prof.n) {
    set_error("ynb_profile_read: index %d out of range", i);
    return YNB_EINVAL;
  }
  YNB_HIP(hipEventSynchronize(g_prof.ev[2 * i + 1]));
  YNB_HIP(hipEventElapsedTime(ms, g_prof.ev[2 * i], g_prof.ev[2 * i + 1]));
  *name = g_prof.name[i];
  return YNB_OK;
}

int ynb_verlet_kick_drift(const ynb_batch *b, double *pos, double *vel, const double *gpos, const double *minv,
                          const double *dt, const ynb_ewald *ew, const double *charges, void *nl_work,
                          size_t nl_work_bytes, const ynb_nlist *nl, ynb_stream stream) {
  if (!b || !pos || !vel || !gpos || !minv || !dt || b->npad % 64 != 0 || (ew && (!charges || !ew->frac || !ew->step)) ||
      (nl_work && (!nl || nl->rebuild || b->max_bin_atoms < 1 || b->max_bin_atoms > 1024))) {
    set_error("ynb_verlet_kick_drift: invalid arguments");
    return YNB_EINVAL;
  }
  Work w{};
  const bool bin = nl_work && b->nbin > 0;
  if (bin) {
    w = carve(nl_work, b->npad, b->nbin, b->max_bin_atoms);
    if (nl_work_bytes < w.total) {
      set_error("ynb_verlet_kick_drift: workspace of %zu bytes, need %zu", nl_work_bytes, w.total);
      return YNB_EWORKSPACE;
    }
  }
  YNB_LAUNCH("verlet_kick_drift", k_kick_drift, dim3((unsigned)ceil_div(b->npad, 256)), dim3(256), 0,
             (hipStream_t)stream, b->sys, b->wave_sys, b->npad, pos, vel, gpos, minv, dt, charges,
             ew ? ew->frac : nullptr, ew ? ew->step : nullptr, b->max_bin_atoms, bin ? w.bin_count : nullptr,
             w.slot_idx, w.xw, w.wr, bin ? nl->status : nullptr);
  YNB_LAUNCHED();
  return YNB_OK;
}

int ynb_verlet_kick(const ynb_batch *b, double *vel, double *gpos, const double *gpart, const double *minv,
                    const double *masses, const double *dt, double *partial, int32_t *ctr, double *ekin2,
                    const double *pair_out, const double *ewald_out, double *total, ynb_stream stream) {
  if (!b || !vel || !gpos || !minv || !masses || !dt || !partial || !ctr || !ekin2 || b->npad % 256 != 0) {
    set_error("ynb_verlet_kick: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  YNB_LAUNCH("verlet_kick", k_kick, dim3((unsigned)(b->npad / 64)), dim3(192), 0, st, b->sys, b->wave_sys, b->npad,
             vel, gpos, gpart, minv, masses, dt, partial, ctr, ekin2, pair_out, ewald_out, total);
  YNB_LAUNCHED();
  return YNB_OK;
}

int ynb_verlet_kick_kick_drift(const ynb_batch *b, double *pos, double *vel, double *gpos, const double *gpart,
                               const double *minv, const double *masses, const double *dt, double *partial,
                               int32_t *ctr, double *ekin2, const double *pair_out, const double *ewald_out,
                               double *total, const ynb_ewald *ew, const double *charges, void *nl_work,
                               size_t nl_work_bytes, const ynb_nlist *nl, ynb_stream stream) {
  if (!b || !pos || !vel || !gpos || !minv || !masses || !dt || !partial || !ctr || !ekin2 || b->npad % 256 != 0 ||
      (ew && (!charges || !ew->frac || !ew->step)) ||
      (nl_work && (!nl || nl->rebuild || b->max_bin_atoms < 1 || b->max_bin_atoms > 1024))) {
    set_error("ynb_verlet_kick_kick_drift: invalid arguments");
    return YNB_EINVAL;
  }
  Work w{};
  const bool bin = nl_work && b->nbin > 0;
  if (bin) {
    w = carve(nl_work, b->npad, b->nbin, b->max_bin_atoms);
    if (nl_work_bytes < w.total) {
      set_error("ynb_verlet_kick_kick_drift: workspace of %zu bytes, need %zu", nl_work_bytes, w.total);
      return YNB_EWORKSPACE;
    }
  }
  YNB_LAUNCH("verlet_kick_kick_drift", k_kick_kick_drift, dim3((unsigned)(b->npad / 64)), dim3(192), 0,
             (hipStream_t)stream, b->sys, b->wave_sys, b->npad, pos, vel, gpos, gpart, minv, masses, dt, partial, ctr,
             ekin2, pair_out, ewald_out, total, charges, ew ? ew->frac : nullptr, ew ? ew->step : nullptr,
             b->max_bin_atoms, bin ? w.bin_count : nullptr, w.slot_idx, w.xw, w.wr, bin ? nl->status : nullptr);
  YNB_LAUNCHED();
  return YNB_OK;
}

int ynb_total(const ynb_batch *b, const double *pair_out, const double *ewald_out, double *total,
              ynb_stream stream) {
  if (!b || !total) {
    set_error("ynb_total: invalid arguments");
    return YNB_EINVAL;
  }
  YNB_LAUNCH("total", k_total, dim3((unsigned)ceil_div(b->nsys, 64)), dim3(64), 0, (hipStream_t)stream, b->sys,
                     b->nsys, pair_out, ewald_out, total);
  YNB_LAUNCHED();
  return YNB_OK;
}

}  // extern "C"
