// Batched table updates of a persistent batch: one launch applies every copy
// and fill of a membership change (new systems' slices of the batch tables,
// owner entries of freed ranges), described by a segment list that travels
// to the device together with the data in one staging buffer.
#include "common.h"

namespace ynb {
namespace {

// One block per segment; 4- or 8-byte elements, vector stores.
__global__ __launch_bounds__(256) void k_scatter(const ynb_segment *__restrict__ seg,
                                                 const uint8_t *__restrict__ src) {
  const ynb_segment s = seg[blockIdx.x];
  if (s.elem == 8) {
    uint64_t *d = (uint64_t *)s.dst;
    const int64_t n = s.nbytes / 8;
    if (s.src < 0) {
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = s.fill;
    } else {
      const uint64_t *q = (const uint64_t *)(src + s.src);
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = q[i];
    }
  } else {
    uint32_t *d = (uint32_t *)s.dst;
    const int64_t n = s.nbytes / 4;
    const uint32_t f = (uint32_t)s.fill;
    if (s.src < 0) {
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = f;
    } else {
      const uint32_t *q = (const uint32_t *)(src + s.src);
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d[i] = q[i];
    }
  }
}

// Structure-factor work items of one system (as ynb_ewald_work).
__global__ __launch_bounds__(256) void k_ewald_work(int32_t *work, int64_t nblk, int64_t nchunk, int32_t sys,
                                                    int32_t blk0, int32_t chunk0) {
  const int64_t n = nblk * nchunk;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < n; w += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = w / nchunk;
    *(int4 *)(work + 4 * w) = make_int4(sys, (int32_t)(blk0 + b), (int32_t)(chunk0 + (w - b * nchunk)), 0);
  }
}

}  // namespace
}  // namespace ynb

using namespace ynb;

extern "C" int ynb_scatter(const ynb_segment *segs, int32_t nseg, const void *src, ynb_stream stream) {
  if (nseg < 0 || (nseg > 0 && (!segs || !src))) {
    set_error("ynb_scatter: invalid arguments");
    return YNB_EINVAL;
  }
  if (nseg == 0) return YNB_OK;
  YNB_LAUNCH("scatter", k_scatter, dim3((unsigned)nseg), dim3(256), 0, (hipStream_t)stream, segs,
             (const uint8_t *)src);
  YNB_LAUNCHED();
  return YNB_OK;
}

extern "C" int ynb_ewald_work_device(int32_t *work, int64_t nblk, int64_t nchunk, int32_t sys, int32_t blk0,
                                     int32_t chunk0, ynb_stream stream) {
  if (!work || nblk < 0 || nchunk < 0) {
    set_error("ynb_ewald_work_device: invalid arguments");
    return YNB_EINVAL;
  }
  const int64_t n = nblk * nchunk;
  if (n == 0) return YNB_OK;
  const int64_t nb = (n + 255) / 256;
  YNB_LAUNCH("ewald_work", k_ewald_work, dim3((unsigned)(nb < 4096 ? nb : 4096)), dim3(256), 0,
             (hipStream_t)stream, work, nblk, nchunk, sys, blk0, chunk0);
  YNB_LAUNCHED();
  return YNB_OK;
}
