// Velocity Verlet halves on device (yaff/sampling/verlet.py:339-367), totals
// of a force-field evaluation, and the library's error/version plumbing.
#include <stdarg.h>
#include <string.h>

#include "common.h"

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

// acc = -gpos/m; vel += dt/2 acc; pos += dt vel  (verlet.py:344-347)
__global__ void k_kick_drift(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad, double *pos, double *vel,
                             const double *gpos, const double *minv, const double *dt) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= npad) return;
  const int sid = wave_sys[g >> 6];
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  if (g - S.atom0 >= S.natom) return;
  const double h = dt[sid];
  const double mi = minv[g];
  for (int c = 0; c < 3; ++c) {
    const double a = -gpos[3 * g + c] * mi;
    const double v = vel[3 * g + c] + 0.5 * a * h;
    vel[3 * g + c] = v;
    pos[3 * g + c] += h * v;
  }
}

// vel += dt/2 acc (verlet.py:354-355) and per-wave sums of m v^2.
__global__ void k_kick(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad, double *vel, const double *gpos,
                       const double *minv, const double *masses, const double *dt, double *partial) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= npad) return;  // npad is a multiple of 64: whole waves leave
  const int sid = uniform_int(wave_sys[g >> 6]);
  if (sid < 0) return;
  const ynb_sys &S = sys[sid];
  double e2 = 0.0;
  if (g - S.atom0 < S.natom) {
    const double h = dt[sid];
    const double mi = minv[g];
    for (int c = 0; c < 3; ++c) {
      const double a = -gpos[3 * g + c] * mi;
      const double v = vel[3 * g + c] + 0.5 * a * h;
      vel[3 * g + c] = v;
      e2 += masses[g] * v * v;
    }
  }
  e2 = wave_sum(e2);
  if ((threadIdx.x & 63) == 0) partial[g >> 6] = e2;
}

// One block per system: fixed strided sums over the system's waves, then a
// fixed LDS tree.
__global__ __launch_bounds__(256) void k_wave_to_sys(const ynb_sys *sys, const double *partial, double *out) {
  const ynb_sys &S = sys[blockIdx.x];
  const int t = threadIdx.x;
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t w = S.atom0 / 64 + t; w < (S.atom0 + S.natom + 63) / 64; w += 256) acc += partial[w];
  red[t] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = red[0];
}

__global__ void k_total(const ynb_sys *sys, int nsys, const double *pair_out, const double *ewald_out,
                        double *total) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const ynb_sys &S = sys[s];
  for (int c = 0; c < 10; ++c) {
    double acc = 0.0;
    if (pair_out)
      for (int p = 0; p < S.npot; ++p) acc += pair_out[((int64_t)s * YNB_MAX_POT + p) * 10 + c];
    if (ewald_out) acc += ewald_out[(int64_t)s * 10 + c];
    total[(int64_t)s * 10 + c] = acc;
  }
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
                          const double *dt, ynb_stream stream) {
  if (!b || !pos || !vel || !gpos || !minv || !dt) {
    set_error("ynb_verlet_kick_drift: invalid arguments");
    return YNB_EINVAL;
  }
  YNB_LAUNCH("verlet_kick_drift", k_kick_drift, dim3((unsigned)ceil_div(b->npad, 256)), dim3(256), 0,
             (hipStream_t)stream, b->sys, b->wave_sys, b->npad, pos, vel, gpos, minv, dt);
  YNB_LAUNCHED();
  return YNB_OK;
}

int ynb_verlet_kick(const ynb_batch *b, double *vel, const double *gpos, const double *minv, const double *masses,
                    const double *dt, double *partial, double *ekin2, ynb_stream stream) {
  if (!b || !vel || !gpos || !minv || !masses || !dt || !partial || !ekin2 || b->npad % 64 != 0) {
    set_error("ynb_verlet_kick: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  YNB_LAUNCH("verlet_kick", k_kick, dim3((unsigned)ceil_div(b->npad, 256)), dim3(256), 0, st, b->sys, b->wave_sys,
             b->npad, vel, gpos, minv, masses, dt, partial);
  YNB_LAUNCHED();
  YNB_LAUNCH("verlet_ekin", k_wave_to_sys, dim3((unsigned)b->nsys), dim3(256), 0, st, b->sys, partial, ekin2);
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
