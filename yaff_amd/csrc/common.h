// Shared device helpers of libyaff_nb.so.
//
// This library is compiled with -ffp-contract=off: every FP64 expression that
// restates Yaff's C (yaff/pes/cell.c, nlist.c, pair_pot.c) is evaluated with
// the same operations in the same order and without fused multiply-adds, so
// the device reproduces the reference bits (pair membership, d, dx, dy, dz).
// Code that may contract writes fma() explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/yaff_nb.h"

namespace ynb {

void set_error(const char *fmt, ...);

#define YNB_HIP(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::ynb::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                       __FILE__, __LINE__);                                    \
      return YNB_EHIP;                                                         \
    }                                                                          \
  } while (0)

#define YNB_LAUNCHED()                                                         \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      ::ynb::set_error("kernel launch failed: %s (%s:%d)",                     \
                       hipGetErrorString(e_), __FILE__, __LINE__);             \
      return YNB_EHIP;                                                         \
    }                                                                          \
  } while (0)

// Kernel timing with HIP events (ynb_profile_*): when enabled, every launch
// made through YNB_LAUNCH is bracketed by two pre-created events recorded on
// the launch's own stream.
void prof_before(const char *name, hipStream_t st);
void prof_after(hipStream_t st);

#define YNB_LAUNCH(name, kernel, grid, block, shmem, stream, ...)               \
  do {                                                                         \
    ::ynb::prof_before(name, stream);                                          \
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);       \
    ::ynb::prof_after(stream);                                                 \
  } while (0)

// Image translations are packed 10 bits per axis with an offset of 512.
__host__ __device__ inline int32_t pack_img(int n0, int n1, int n2) {
  return (n0 + 512) | ((n1 + 512) << 10) | ((n2 + 512) << 20);
}
__host__ __device__ inline void unpack_img(int32_t c, int n[3]) {
  n[0] = (c & 1023) - 512;
  n[1] = ((c >> 10) & 1023) - 512;
  n[2] = ((c >> 20) & 1023) - 512;
}

// cell.c:81-106 -- sequential per-axis wrap.  m receives the integer
// subtracted along each axis (0 beyond nvec).
__device__ inline void mic_exact(double d[3], const double *rv, const double *gv, int nvec,
                                 int m[3]) {
  m[0] = m[1] = m[2] = 0;
  for (int i = 0; i < 3; ++i) {
    if (i >= nvec) break;
    const double x = ceil(gv[3 * i] * d[0] + gv[3 * i + 1] * d[1] + gv[3 * i + 2] * d[2] - 0.5);
    d[0] -= x * rv[3 * i];
    d[1] -= x * rv[3 * i + 1];
    d[2] -= x * rv[3 * i + 2];
    m[i] = (int)x;
  }
}

// cell.c:124-143 -- d += r.R axis by axis.
__device__ inline void add_vec_exact(double d[3], const double *rv, int nvec, const int r[3]) {
  for (int i = 0; i < 3; ++i) {
    if (i >= nvec) break;
    const double ri = (double)r[i];
    d[0] += ri * rv[3 * i];
    d[1] += ri * rv[3 * i + 1];
    d[2] += ri * rv[3 * i + 2];
  }
}

// Membership of an image index in the half space Yaff enumerates
// (nlist.c:137-180): the origin, then r2 > 0, or r2 == 0 and r1 > 0, or
// r2 == r1 == 0 and r0 > 0.
__host__ __device__ inline bool in_half_space(const int r[3]) {
  if (r[2] != 0) return r[2] > 0;
  if (r[1] != 0) return r[1] > 0;
  return r[0] >= 0;
}

// Yaff's half row for the unordered pair (A >= B) whose partner image is
// shifted by n_ab cells: delta0 = mic(pos_B - pos_A) (nlist.c:108-113),
// r = n_ab + m and v = delta0 + r.R (nlist.c:120-123).  v and r are always
// computed; the return value is false when r lies outside the image box
// [-rmax, rmax] that the reference scans (such a row is not Yaff's).
__device__ inline bool yaff_row_vector(const double *pA, const double *pB, bool self,
                                       const int nab[3], const double *rv, const double *gv,
                                       int nvec, const int *rmax, double v[3], int r[3]) {
  int m[3] = {0, 0, 0};
  if (self) {
    v[0] = 0.0; v[1] = 0.0; v[2] = 0.0;
  } else {
    v[0] = pB[0] - pA[0];
    v[1] = pB[1] - pA[1];
    v[2] = pB[2] - pA[2];
    mic_exact(v, rv, gv, nvec, m);
  }
  bool inside = true;
  for (int i = 0; i < 3; ++i) {
    r[i] = nab[i] + m[i];
    if (i < nvec) {
      if (r[i] > rmax[i] || r[i] < -rmax[i]) inside = false;
    } else if (r[i] != 0) {
      inside = false;
    }
  }
  add_vec_exact(v, rv, nvec, r);
  return inside;
}

// Home bins per block of the cell-binned build: runs of 2 bins along the
// first cell vector when the union of their search boxes (range[0..2] bins
// on each side) fits YNB_BUILD_NB_SMALL partner bins, else single bins.  A
// function of the system alone.
__host__ __device__ inline int build_group_len(const int32_t range[3]) {
  const int n12 = (2 * range[1] + 1) * (2 * range[2] + 1), w0 = 2 * range[0] + 1;
  return n12 * (w0 + 1) <= YNB_BUILD_NB_SMALL ? 2 : 1;
}

// Partner bins the build stages for one home group of the system.
__host__ __device__ inline int build_search_bins(const int32_t range[3]) {
  const int n12 = (2 * range[1] + 1) * (2 * range[2] + 1), w0 = 2 * range[0] + 1;
  return n12 * (w0 + build_group_len(range) - 1);
}

__device__ inline bool nlow_nhigh_ok(int a, int b, int nlow, int nhigh) {
  return (a >= nlow && b < nhigh) || (b >= nlow && a < nhigh);  // nlist.c:57
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ inline int wave_max(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// Status counters written by many threads: one atomic per wavefront (device
// atomics on one address from every lane serialise across the chip).
__device__ inline void wave_status_max(int32_t *addr, int v) {
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0 && v > 0) atomicMax(addr, v);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Partial results handed from one block to another within a kernel are
// written with publish() and read with gather(): relaxed atomic store and
// load at device scope, which go through the XCD's L2 to the device-coherent
// level (the sc1 bit), so no L2 write-back or invalidate is needed to pass
// them between XCDs.
__device__ inline void publish(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double gather(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Completion of a group of nblocks blocks (the blocks of one system, say):
// each block calls this, uniformly in the block, after publishing its
// partial results; exactly one block -- the last to arrive -- gets true, and
// it may then gather every partial of the group, reduce them in a fixed order
// and write the group's result, so a separate reduction kernel is not needed.
// The counter is reset by the last block, ready for the next launch (and for
// graph replays).  Counters start at zero.
//
// Ordering: each wave waits until its publish() stores are acknowledged
// (device-coherent), the block meets at a barrier, then one thread counts the
// block in with a device-scope atomic; the last block's gathers are issued
// after it has read the final count.  Data written with plain stores is not
// covered (that would need a device-scope release fence, an L2 write-back
// per block; __threadfence in every wave also invalidates the L2 the kernel
// still reads from, which cost 30-50% of the fused kernels' time).
__device__ inline bool last_block(int32_t *ctr, int nblocks) {
  __shared__ int l_last;
  __builtin_amdgcn_s_waitcnt(0);  // vmcnt = expcnt = lgkmcnt = 0
  __syncthreads();
  if (threadIdx.x == 0) {
    const int prev = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int last = (prev == nblocks - 1);
    if (last) __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    l_last = last;
  }
  __syncthreads();
  return l_last != 0;
}

// a * b + c as the three-operand v_fma_f64 with the addend c a scalar
// register (c: a constant coefficient).  For an addend that stays live (a
// polynomial coefficient held in a vector register across a loop) the
// compiler otherwise picks the two-operand v_fmac_f64, whose destination is
// the addend, and copies the coefficient into it first: two instructions per
// Horner step instead of one, and a vector register pair per coefficient.
__device__ inline double fma3(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}

// 1/x in FP64: the hardware reciprocal estimate refined by two Newton steps
// (relative error about 1 ulp; no special handling of 0, inf or denormals --
// callers pass finite normal arguments).
__device__ inline double rcp_fp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// exp(a) in FP64 by Cody-Waite reduction: k = rint(a / ln 2),
// r = a - k ln2_hi - k ln2_lo with ln2_hi holding the leading 32 bits of ln 2
// (k ln2_hi is exact for |k| <= 1100), |r| <= ln 2 / 2; e^r from its Taylor
// series to r^13 (truncation < 5e-18 relative), evaluated as
// E(r^2) + r O(r^2) with its even and odd parts by Horner's rule in r^2 (two
// independent chains of 6 steps); then 2^k.  Relative error about 1 ulp;
// 0 below a = -760, inf above 709.8, NaN kept.
__device__ inline double exp_fp64(double a) {
  const double k = rint(a * 1.4426950408889634);
  const double kc = fmax(fmin(k, 1100.0), -1100.0);
  double r = fma(-kc, 0.6931471803691238, a);
  r = fma(-kc, 1.9082149292705877e-10, r);
  const double z = r * r;
  double e = fma3(z, 2.08767569878681e-09, 2.755731922398589e-07);  // 1/12!, 1/10!
  double o = fma3(z, 1.6059043836821613e-10, 2.505210838544172e-08);  // 1/13!, 1/11!
  e = fma3(e, z, 2.48015873015873e-05);                               // 1/8!
  o = fma3(o, z, 2.7557319223985893e-06);                             // 1/9!
  e = fma3(e, z, 0.001388888888888889);                               // 1/6!
  o = fma3(o, z, 0.0001984126984126984);                              // 1/7!
  e = fma3(e, z, 0.041666666666666664);                               // 1/4!
  o = fma3(o, z, 0.008333333333333333);                               // 1/5!
  e = fma3(e, z, 0.5);                                                // 1/2!
  o = fma3(o, z, 0.16666666666666666);                                // 1/3!
  e = fma3(e, z, 1.0);                                                // 1/0!
  o = fma3(o, z, 1.0);                                                // 1/1!
  return ldexp(fma(r, o, e), (int)kc);
}

}  // namespace ynb
