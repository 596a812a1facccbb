// Neighbour lists: cell-binned build for 3-D periodic systems, all-pairs build
// for 0/1/2-D systems, Yaff-ordered row materialisation and the skin-path
// recompute.
//
// Pair set.  Yaff keeps the half row (A >= B, image r in its half space) when
// |v|^2 < rc^2 with v = mic(pos_B - pos_A) + r.R evaluated exactly as
// nlist_build_low does (yaff/pes/nlist.c:30-131).  The binned search below is
// only a conservative candidate filter (padded radius); every candidate is
// then re-tested with that exact expression (yaff_row_vector), so the
// accepted set is the reference's bit for bit, including the image box
// |r_i| <= rmax_i that limits Yaff in skewed cells.
#include <hipcub/hipcub.hpp>

#include <type_traits>

#include "common.h"
#include "nlist_work.h"

namespace ynb {

namespace {

__device__ inline int floor_div(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

// Fractional wrap and bin of every atom of the binned systems; each atom
// takes a slot of its bin (arrival order, made deterministic by k_bin_order).
__global__ void k_bin(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad, const double *pos, int32_t cb,
                      int32_t *bin_count, int32_t *slot_idx, double4 *xw, int4 *wr, int32_t *status,
                      const int32_t *rebuild) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= npad) return;  // whole waves: npad is a multiple of 64
  const int sid = uniform_int(wave_sys[g >> 6]);
  if (sid < 0 || (rebuild && !rebuild[sid])) return;
  const ynb_sys &S = sys[sid];
  const bool active = g - S.atom0 < S.natom;
  bin_atom(S, active, g, active ? pos[3 * g] : 0.0, active ? pos[3 * g + 1] : 0.0, active ? pos[3 * g + 2] : 0.0,
           cb, bin_count, slot_idx, xw, wr, status);
}

// One block per bin: orders the bin's atoms by index (rank among the bin's
// atoms) and writes their wrapped data to the bin's slots, so the layout
// does not depend on the order of the atomics in k_bin.
__global__ __launch_bounds__(64) void k_bin_order(int32_t cb, int32_t *bin_count, const int32_t *slot_idx,
                                                   const double4 *xw, const int4 *wr, double4 *sxw, int4 *sw,
                                                   int32_t *bin_end, const int32_t *bin_sys,
                                                   const int32_t *rebuild) {
  __shared__ int32_t l_idx[1024];
  const int b = blockIdx.x;
  const int sid = bin_sys[b];
  if (sid < 0 || (rebuild && !rebuild[sid])) return;  // a bin of no system holds no atoms
  const int n = min(bin_count[b], cb);
  for (int e = threadIdx.x; e < n; e += blockDim.x) l_idx[e] = slot_idx[(int64_t)b * cb + e];
  __syncthreads();
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int g = l_idx[e];
    int rank = 0;
    for (int q = 0; q < n; ++q) rank += l_idx[q] < g;
    const int64_t o = (int64_t)b * cb + rank;
    sxw[o] = xw[g];
    sw[o] = wr[g];
  }
  if (threadIdx.x == 0) {
    bin_end[b] = b * cb + n;
    bin_count[b] = 0;  // every thread has read the count (barrier above): ready for the next build
  }
}

__device__ inline void store_entry(const ynb_nlist &nl, int64_t npad, int64_t gi, int cnt, int32_t gj,
                                   int32_t img) {
  if (cnt < nl.kcap) {
    nl.j[gi * nl.kcap + cnt] = gj;
    nl.img[gi * nl.kcap + cnt] = img;
  }
}

constexpr int BW = 8;         // atoms (wavefronts) per block of the binned build
constexpr int CAP = 1536;     // partner candidates staged in LDS at once
constexpr int MAXNB = 512;    // partner bins per home group (ynb_sys_plan keeps (2R+1)^3 below)
constexpr int NB_SMALL = YNB_BUILD_NB_SMALL;  // partner-bin tables of the common search boxes

// Lanes of the wave below this one that hold a set bit of mask.
__device__ inline int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

// Block (home group of bin b, part y of its atoms), one atom per wavefront,
// the atoms BW gridDim.y apart.  A home group is a run of up to
// build_group_len(S) bins consecutive along the first cell vector (the run
// starts at a first-axis bin index divisible by the length), so
// neighbouring home bins share most partner bins: the union of their search
// boxes is staged once.  The blocks of the other bins of a run exit at
// once.  Partner bins are staged in (first, second, third) axis order, the
// first axis slowest, so the search box of home bin h of the run is one
// contiguous range of the staged candidates, in the order a single home
// bin would stage them; as FP32 coordinates relative to the run's centre
// with the image shift applied.  The wave scans 64 candidates per step;
// accepted candidates are appended to the atom's row with a ballot (stream
// compaction), so rows are written coalesced and in a fixed order (partner
// bins in offset order, partner atoms by index) that depends on the system
// alone.
//
// Acceptance.  Relative coordinates stay below ~150 bohr, so the FP32
// distance squared is within 1e-5 relative of the exact one; below a band of
// -2e-4 relative around rc^2 it accepts, above +2e-4 it rejects.  Inside the
// band: with EXACT (some system of the batch is not box_safe), the exact
// Yaff expression on the FP64 positions decides here; without it the
// candidate is kept tentatively (the image box cannot bind when box_safe, see
// ynb_sys_plan) and the exact test is made by whoever needs it: the pair
// kernel re-tests every entry within 1e-12 of a cutoff exactly and skips
// d >= rc, and the row materialisation drops entries with exact d >= rc.
// NB: capacity of the partner-bin tables (NB_SMALL or MAXNB).
template <bool EXACT, int NB>
__global__ __launch_bounds__(BW * 64) void k_build_binned(const ynb_sys *__restrict__ sys,
                                                          const int32_t *__restrict__ bin_sys, int64_t npad,
                                                          const double *__restrict__ pos,
                                                          const int32_t *__restrict__ bin_end, int32_t cb,
                                                          const double4 *__restrict__ sxw,
                                                          const int4 *__restrict__ sw, ynb_nlist nl) {
  __shared__ float l_x[CAP], l_y[CAP], l_z[CAP];
  __shared__ int32_t l_m[CAP], l_g[CAP];
  __shared__ int32_t l_pre[NB + 1];
  __shared__ int32_t l_bs[NB];       // slot start of each partner bin
  __shared__ double l_sh[3][NB];     // image shift of each partner bin minus the run's centre
  __shared__ int32_t l_s[NB];        // packed shift (s0, s1, s2)
  __shared__ int16_t l_bin[CAP];     // partner bin of each staged candidate
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = uniform_int(tid >> 6);
  const int sid = bin_sys[b];
  if (sid < 0 || (nl.rebuild && !nl.rebuild[sid])) return;
  const ynb_sys &S = sys[sid];
  const int hstride = S.grid[1] * S.grid[2];
  const int LG = build_group_len(S.range);
  int lb = b - S.bin0;
  const int c2 = lb % S.grid[2], c1 = (lb / S.grid[2]) % S.grid[1];
  int c0 = lb / hstride;
  // The run's blocks: with one block row (gridDim.y == 1, many bins) the
  // run's first bin builds it alone and the other bin's block exits; with
  // several rows the blocks of both bins split the run's atoms (parts
  // 2 y + 0 and 2 y + 1), so the first row holds the work of every run.
  const int off = c0 % LG;
  c0 -= off;
  lb -= off * hstride;
  const int nh = min(LG, S.grid[0] - c0);  // home bins bh + h hstride, h < nh
  const int bh = S.bin0 + lb;               // the run's first bin
  const bool split = nh > 1 && gridDim.y > 1;
  if (off != 0 && !split) return;
  const int part = split ? 2 * (int)blockIdx.y + off : (int)blockIdx.y;
  const int nparts = split ? 2 * (int)gridDim.y : (int)gridDim.y;
  const int R0 = S.range[0], R1 = S.range[1], R2 = S.range[2];
  // atoms of the run: home bin h holds n_h of them from slot (bh + h hstride) cb
  int hend[2];  // prefix of the home bins' atom counts (uniform)
  int natoms = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h < nh) natoms += bin_end[bh + h * hstride] - (bh + h * hstride) * cb;
    hend[h] = natoms;
  }
  if (part * BW >= natoms) return;

  const int n1 = 2 * R1 + 1, n2 = 2 * R2 + 1, n12 = n1 * n2;
  const int w0 = 2 * R0 + 1;             // first-axis slices searched per home bin
  const int nnb = n12 * (w0 + nh - 1);   // partner bins of the run
  if (nnb > NB) {  // a batch flag that does not match its systems (ynb_build_search_bins)
    if (tid == 0) atomicOr(nl.status, 8);
    return;
  }
  const double f0 = (c0 + 0.5 * nh) / S.grid[0], f1 = (c1 + 0.5) / S.grid[1], f2 = (c2 + 0.5) / S.grid[2];
  const double X0 = f0 * S.rvecs[0] + f1 * S.rvecs[3] + f2 * S.rvecs[6];
  const double X1 = f0 * S.rvecs[1] + f1 * S.rvecs[4] + f2 * S.rvecs[7];
  const double X2 = f0 * S.rvecs[2] + f1 * S.rvecs[5] + f2 * S.rvecs[8];

  // Partner-bin table: one thread per partner bin, all bins at once (one
  // round trip for their sizes), then the exclusive prefix of the sizes
  // (64 bins per step in wave 0).  Bin k = (i0 n1 + i1) n2 + i2.
  for (int k = tid; k < nnb; k += BW * 64) {
    const int i2 = k % n2, i1 = (k / n2) % n1, i0 = k / n12;
    int nb0 = c0 + i0 - R0, nb1 = c1 + i1 - R1, nb2 = c2 + i2 - R2;
    const int s0 = floor_div(nb0, S.grid[0]), s1 = floor_div(nb1, S.grid[1]), s2 = floor_div(nb2, S.grid[2]);
    nb0 -= s0 * S.grid[0];
    nb1 -= s1 * S.grid[1];
    nb2 -= s2 * S.grid[2];
    const int nbin = S.bin0 + (nb0 * S.grid[1] + nb1) * S.grid[2] + nb2;
    const int bs = nbin * cb;
    l_pre[k + 1] = bin_end[nbin] - bs;  // the size, made a prefix below
    l_bs[k] = bs;
    l_sh[0][k] = s0 * S.rvecs[0] + s1 * S.rvecs[3] + s2 * S.rvecs[6] - X0;
    l_sh[1][k] = s0 * S.rvecs[1] + s1 * S.rvecs[4] + s2 * S.rvecs[7] - X1;
    l_sh[2][k] = s0 * S.rvecs[2] + s1 * S.rvecs[5] + s2 * S.rvecs[8] - X2;
    const bool far = s0 < -511 || s0 > 511 || s1 < -511 || s1 > 511 || s2 < -511 || s2 > 511;
    l_s[k] = far ? INT32_MIN : pack_img(s0, s1, s2);
  }
  __syncthreads();
  if (wv == 0) {
    int carry = 0;
    for (int k0 = 0; k0 < nnb; k0 += 64) {
      const int k = k0 + lane;
      const int c = k < nnb ? l_pre[k + 1] : 0;  // each lane reads its own entry before writing it
      int incl = c;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
      }
      if (k < nnb) l_pre[k + 1] = carry + incl;
      carry += __shfl(incl, 63, 64);
    }
    if (lane == 0) l_pre[0] = 0;
  }
  __syncthreads();
  // Candidates staged: the search boxes of the home bins this block's atoms
  // belong to (with one pass of the block over the run's atoms, its parts
  // stage only what their atoms search), from sb on.
  int h_lo = 0, h_hi = nh - 1;
  if (nparts * BW >= natoms) {
    const int a0 = part * BW, a1 = min(a0 + BW, natoms) - 1;
    h_lo = a0 >= hend[0] ? 1 : 0;
    h_hi = a1 >= hend[0] ? 1 : 0;
  }
  const int sb = l_pre[h_lo * n12];
  const int T = l_pre[(h_hi + w0) * n12] - sb;
  const double rc2 = S.rc * S.rc;
  // per-system values used in the scan, held in registers: the scan stores
  // rows through pointers the compiler cannot separate from sys, so reads of
  // S inside it would be repeated memory loads
  const int atom0 = S.atom0, nlow = S.nlow, nhigh = S.nhigh;
  const float lo32 = (float)(rc2 * (1.0 - 2e-4)), hi32 = (float)(rc2 * (1.0 + 2e-4));
  bool bad = false;
  bool wide = false;  // a staged image shift outside [-255, 255] (this thread)

  // Stage the candidates sb + [base, base + CAP): candidate f belongs to
  // the partner bin k with l_pre[k] <= f < l_pre[k + 1] (a table written bin
  // by bin, then one lookup per candidate).
  auto stage = [&](int base) {
    const int lim = min(T - base, CAP);
    const int ab = sb + base;
    // partner bin of each staged slot: thread k writes bin k over its range
    for (int k = tid; k < nnb; k += BW * 64) {
      const int f1 = min(l_pre[k + 1] - ab, lim);
      for (int f = max(l_pre[k] - ab, 0); f < f1; ++f) l_bin[f] = (int16_t)k;
    }
    __syncthreads();
    for (int q = tid; q < lim; q += BW * 64) {
      const int f = ab + q;
      const int lo = l_bin[q];
      const int j = l_bs[lo] + (f - l_pre[lo]);
      const double4 x = sxw[j];
      const int4 w = sw[j];
      l_x[q] = (float)(x.x + l_sh[0][lo]);
      l_y[q] = (float)(x.y + l_sh[1][lo]);
      l_z[q] = (float)(x.z + l_sh[2][lo]);
      const int sp = l_s[lo];
      int s3[3];
      unpack_img(sp, s3);
      const int m0 = s3[0] - w.x, m1 = s3[1] - w.y, m2 = s3[2] - w.z;
      if (sp == INT32_MIN || m0 < -511 || m0 > 511 || m1 < -511 || m1 > 511 || m2 < -511 || m2 > 511) bad = true;
      wide |= m0 < -255 || m0 > 255 || m1 < -255 || m1 > 255 || m2 < -255 || m2 > 255;
      l_m[q] = pack_img(m0, m1, m2);
      l_g[q] = w.w;
    }
  };

  const bool one_stage = T <= CAP;
  int any_wide = 0;  // uniform: some staged shift needs the unpacked image test
  if (one_stage) {
    stage(0);
    any_wide = __syncthreads_or(wide);
  }
  // Packed images add field by field while every field stays in range:
  // with all staged shifts and the atom's wrap counts within [-255, 255],
  // code(m + w) = code(m) + w0 + w1 2^10 + w2 2^20 and no sum leaves
  // [-511, 511].  The unpacked test covers the rest.
  const bool no_groups = nlow <= 0 && nhigh >= S.natom;
  const int32_t zero_code = pack_img(0, 0, 0);
  for (int ia = part * BW; ia < natoms; ia += nparts * BW) {
    const int a = ia + wv;
    const bool active = a < natoms;  // uniform in the wave
    // home bin h of the atom and its slot p
    const int h = a >= hend[0] ? 1 : 0;
    const int p = (bh + h * hstride) * cb + (a - (h > 0 ? hend[0] : 0));
    // its search box: first-axis slices h .. h + 2 R0 of the run
    const int rlo = l_pre[h * n12] - sb, rhi = l_pre[(h + w0) * n12] - sb;
    double4 xi = make_double4(0, 0, 0, 0);
    int4 wi = make_int4(0, 0, 0, 0);
    if (active) {
      xi = sxw[p];
      wi = sw[p];
    }
    const int64_t gi = active ? wi.w : 0;
    const int li = (int)(gi - atom0);
    const float xr = (float)(xi.x - X0), yr = (float)(xi.y - X1), zr = (float)(xi.z - X2);
    double pi[3] = {0, 0, 0};
    if (EXACT && active) {
      pi[0] = pos[3 * gi];
      pi[1] = pos[3 * gi + 1];
      pi[2] = pos[3 * gi + 2];
    }
    const bool fast_ok = !EXACT || S.box_safe != 0;
    int32_t *rowj = nl.j + gi * nl.kcap;
    int32_t *rowi = nl.img + gi * nl.kcap;
    int cnt = 0;  // uniform in the wave
    for (int base = 0; base < T; base += CAP) {
      if (!one_stage) {
        __syncthreads();
        stage(base);
        any_wide = __syncthreads_or(wide);
      }
      if (!active) continue;
      const int lim = min(T - base, CAP);
      // the atom's candidates in this stage: [q_lo, q_hi)
      const int q_lo = max(rlo - base, 0), q_hi = min(rhi - base, lim);
      if (q_lo >= q_hi) continue;
      const bool wi_small = wi.x >= -255 && wi.x <= 255 && wi.y >= -255 && wi.y <= 255 && wi.z >= -255 && wi.z <= 255;
      if (!EXACT && !any_wide && wi_small) {  // uniform in the wave
        const int32_t wlin = wi.x + (wi.y << 10) + (wi.z << 20);
        // the atom index and its row as wave-uniform values (the row is
        // written with a scalar base and 32-bit offsets)
        const int32_t gi_u = __builtin_amdgcn_readfirstlane((int32_t)gi);
        int32_t *const rj = nl.j + (int64_t)gi_u * nl.kcap;
        int32_t *const ri = nl.img + (int64_t)gi_u * nl.kcap;
        // the test of candidate q (qc: the slot read, q clamped for a tail)
        auto test = [&](int q, int qc, auto tail_c, auto groups_c, int32_t &gj, int32_t &code) {
          constexpr bool TAIL = decltype(tail_c)::value, GROUPS = decltype(groups_c)::value;
          const float dx = l_x[qc] - xr, dy = l_y[qc] - yr, dz = l_z[qc] - zr;
          const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          gj = l_g[qc];
          code = l_m[qc] + wlin;
          bool acc = (d2 < hi32) & ((gj != gi_u) | (code != zero_code));
          if (TAIL) acc = acc & (q < q_hi);
          if (GROUPS) {
            const int lj = gj - atom0;
            acc = acc & nlow_nhigh_ok(li > lj ? li : lj, li > lj ? lj : li, nlow, nhigh);
          }
          return acc;
        };
        // appends the accepted candidates of one ballot at offset base of the
        // row; ROOM: the row has room for 64 more entries (uniform), so no
        // lane can pass the capacity
        auto append = [&](bool acc, uint64_t mask, int base, int32_t gj, int32_t code, auto room_c) {
          const uint32_t o = (uint32_t)(base + lanes_below(mask));
          if (acc & (decltype(room_c)::value || o < (uint32_t)nl.kcap)) {
            const uint32_t ob = o * 4u;  // bytes; a row is far below 4 GiB
            *(int32_t *)((char *)rj + ob) = gj;
            *(int32_t *)((char *)ri + ob) = code;
          }
        };
        // one step of 64 candidates; TAIL: the last, partial step; GROUPS:
        // the nlow/nhigh test is needed
        auto scan_step = [&](int q0, auto tail_c, auto groups_c) {
          const int q = q0 + lane;
          const int qc = decltype(tail_c)::value ? min(q, q_hi - 1) : q;
          int32_t gj, code;
          const bool acc = test(q, qc, tail_c, groups_c, gj, code);
          const uint64_t mask = __ballot(acc);
          if (cnt + 64 <= nl.kcap)
            append(acc, mask, cnt, gj, code, std::true_type{});
          else
            append(acc, mask, cnt, gj, code, std::false_type{});
          cnt += __popcll(mask);
        };
        // two full steps at once (128 candidates, entries in candidate order):
        // the loop and count bookkeeping once per 128
        auto scan_step2 = [&](int q0, auto groups_c) {
          const int qa = q0 + lane, qb = q0 + 64 + lane;
          int32_t gja, codea, gjb, codeb;
          const bool acca = test(qa, qa, std::false_type{}, groups_c, gja, codea);
          const bool accb = test(qb, qb, std::false_type{}, groups_c, gjb, codeb);
          const uint64_t ma = __ballot(acca), mb = __ballot(accb);
          const int cb2 = cnt + __popcll(ma);
          if (cnt + 128 <= nl.kcap) {  // uniform: the common case
            append(acca, ma, cnt, gja, codea, std::true_type{});
            append(accb, mb, cb2, gjb, codeb, std::true_type{});
          } else {
            append(acca, ma, cnt, gja, codea, std::false_type{});
            append(accb, mb, cb2, gjb, codeb, std::false_type{});
          }
          cnt = cb2 + __popcll(mb);
        };
        const int n = q_hi - q_lo;
        const int n128 = q_lo + (n & ~127), nfull = q_lo + (n & ~63);
        if (no_groups) {
          for (int q0 = q_lo; q0 < n128; q0 += 128) scan_step2(q0, std::false_type{});
          if (n128 < nfull) scan_step(n128, std::false_type{}, std::false_type{});
          if (nfull < q_hi) scan_step(nfull, std::true_type{}, std::false_type{});
        } else {
          for (int q0 = q_lo; q0 < n128; q0 += 128) scan_step2(q0, std::true_type{});
          if (n128 < nfull) scan_step(n128, std::false_type{}, std::true_type{});
          if (nfull < q_hi) scan_step(nfull, std::true_type{}, std::true_type{});
        }
        continue;
      }
      for (int q0 = q_lo; q0 < q_hi; q0 += 64) {
        const int q = q0 + lane;
        const int qc = q < q_hi ? q : q_hi - 1;
        // branch-free test of the candidate (all lanes read a valid slot)
        const float dx = l_x[qc] - xr, dy = l_y[qc] - yr, dz = l_z[qc] - zr;
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const int32_t gj = l_g[qc];
        int n[3];
        unpack_img(l_m[qc], n);
        n[0] += wi.x;
        n[1] += wi.y;
        n[2] += wi.z;
        const bool self = (gj == gi);
        const int lj = gj - atom0;
        const int A = li > lj ? li : lj, B = li > lj ? lj : li;
        bool acc = (q < q_hi) && (d2 < hi32) && !(self && n[0] == 0 && n[1] == 0 && n[2] == 0) &&
                   nlow_nhigh_ok(A, B, nlow, nhigh);
        if (EXACT && acc && !(fast_ok && d2 < lo32)) {
          const bool i_is_a = (li == A);
          const int nab[3] = {i_is_a ? n[0] : -n[0], i_is_a ? n[1] : -n[1], i_is_a ? n[2] : -n[2]};
          const double pj[3] = {pos[3 * (int64_t)gj], pos[3 * (int64_t)gj + 1], pos[3 * (int64_t)gj + 2]};
          double v[3];
          int r[3];
          acc = yaff_row_vector(i_is_a ? pi : pj, i_is_a ? pj : pi, self, nab, S.rvecs, S.gvecs, 3, S.rmax, v, r) &&
                (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < rc2);
        }
        const bool far = (n[0] < -511) | (n[0] > 511) | (n[1] < -511) | (n[1] > 511) | (n[2] < -511) | (n[2] > 511);
        bad |= acc && far;
        acc = acc && !far;
        const uint64_t mask = __ballot(acc);
        if (acc) {
          const int o = cnt + lanes_below(mask);
          if (o < nl.kcap) {
            rowj[o] = gj;
            rowi[o] = pack_img(n[0], n[1], n[2]);
          }
        }
        cnt += __popcll(mask);
      }
    }
    if (active && lane == 0) {
      nl.cnt[gi] = cnt;
      if (cnt > nl.kcap) atomicOr(nl.status, 1);
    }
  }
  if (bad) atomicOr(nl.status, 2);
}


// One block per system: the largest squared displacement since pos_ref
// (order-independent maximum), the rebuild flag, and pos_ref := pos when set.
__global__ __launch_bounds__(256) void k_skin_check(const ynb_sys *sys, const double *pos, double *pos_ref,
                                                    const double *skin, int32_t *rebuild) {
  const int s = blockIdx.x;
  const ynb_sys &S = sys[s];
  const int t = threadIdx.x;
  __shared__ double red[256];
  __shared__ int flag;
  double m = 0.0;
  for (int a = t; a < S.natom; a += 256) {
    const int64_t g = S.atom0 + a;
    const double dx = pos[3 * g] - pos_ref[3 * g], dy = pos[3 * g + 1] - pos_ref[3 * g + 1],
                 dz = pos[3 * g + 2] - pos_ref[3 * g + 2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    m = (d2 > m || d2 != d2) ? (d2 != d2 ? INFINITY : d2) : m;
  }
  red[t] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) red[t] = fmax(red[t], red[t + off]);
    __syncthreads();
  }
  if (t == 0) {
    const double h = 0.5 * skin[s];
    flag = (skin[s] <= 0.0 || !(red[0] < h * h)) ? 1 : 0;
    rebuild[s] = flag;
  }
  __syncthreads();
  if (flag)
    for (int a = t; a < S.natom; a += 256) {
      const int64_t g = S.atom0 + a;
      pos_ref[3 * g] = pos[3 * g];
      pos_ref[3 * g + 1] = pos[3 * g + 1];
      pos_ref[3 * g + 2] = pos[3 * g + 2];
    }
}

// All-pairs build for systems without bins (nvec < 3): every partner and
// every image of the full box [-rmax, rmax]^nvec, tested exactly.
__global__ void k_build_allpairs(const ynb_sys *sys, const int32_t *wave_sys, int64_t npad,
                                 const double *pos, ynb_nlist nl) {
  const int64_t gi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gi >= npad) return;
  const int sid = wave_sys[gi >> 6];
  if (sid < 0 || (nl.rebuild && !nl.rebuild[sid])) return;
  const ynb_sys &S = sys[sid];
  const int li = (int)(gi - S.atom0);
  if (li >= S.natom || S.nbin != 0) return;
  const double rc2 = S.rc * S.rc;
  const int nvec = S.nvec;
  const int R0 = nvec > 0 ? S.rmax[0] : 0, R1 = nvec > 1 ? S.rmax[1] : 0, R2 = nvec > 2 ? S.rmax[2] : 0;
  const double pi[3] = {pos[3 * gi], pos[3 * gi + 1], pos[3 * gi + 2]};
  int cnt = 0;
  for (int lj = 0; lj < S.natom; ++lj) {
    const int64_t gj = S.atom0 + lj;
    const int A = li > lj ? li : lj, B = li > lj ? lj : li;
    if (!nlow_nhigh_ok(A, B, S.nlow, S.nhigh)) continue;
    const bool self = (lj == li);
    const bool i_is_a = (li == A);
    const double pj[3] = {pos[3 * gj], pos[3 * gj + 1], pos[3 * gj + 2]};
    const double *pA = i_is_a ? pi : pj;
    const double *pB = i_is_a ? pj : pi;
    double d0[3] = {0.0, 0.0, 0.0};
    int m[3] = {0, 0, 0};
    if (!self) {
      d0[0] = pB[0] - pA[0];
      d0[1] = pB[1] - pA[1];
      d0[2] = pB[2] - pA[2];
      mic_exact(d0, S.rvecs, S.gvecs, nvec, m);
    }
    for (int r2 = -R2; r2 <= R2; ++r2)
      for (int r1 = -R1; r1 <= R1; ++r1)
        for (int r0 = -R0; r0 <= R0; ++r0) {
          const int r[3] = {r0, r1, r2};
          if (self && r0 == 0 && r1 == 0 && r2 == 0) continue;
          double v[3] = {d0[0], d0[1], d0[2]};
          add_vec_exact(v, S.rvecs, nvec, r);
          const double e2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
          if (!(e2 < rc2)) continue;
          int n[3] = {r0 - m[0], r1 - m[1], r2 - m[2]};
          if (!i_is_a) {
            n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
          }
          store_entry(nl, npad, gi, cnt, (int32_t)gj, pack_img(n[0], n[1], n[2]));
          ++cnt;
        }
  }
  nl.cnt[gi] = cnt;
  if (cnt > nl.kcap) atomicOr(nl.status, 1);
}

// ---- Yaff rows ------------------------------------------------------------

__device__ inline bool canonical_entry(int64_t gi, int32_t gj, int32_t img) {
  if (gj < gi) return true;
  if (gj > gi) return false;
  int n[3];
  unpack_img(img, n);
  return in_half_space(n);
}

// Entries of the list are Yaff rows when they are canonical and pass the
// exact test (the build may keep entries within 2e-4 of rc^2 tentatively).
__device__ inline bool exact_row(const ynb_sys &S, int li, int lj, const double *pi, const double *pj,
                                 const int nn[3], double v[3], int r[3]) {
  const bool in_box = yaff_row_vector(pi, pj, li == lj, nn, S.rvecs, S.gvecs, S.nvec, S.rmax, v, r);
  return in_box && (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < S.rc * S.rc);
}

__global__ void k_rows_count(const ynb_sys *sys, int s, int64_t npad, const double *pos, ynb_nlist nl,
                             int32_t *rowcnt) {
  const ynb_sys &S = sys[s];
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= S.natom) return;
  const int64_t gi = S.atom0 + li;
  const int n = min(nl.cnt[gi], nl.kcap);
  const double pi[3] = {pos[3 * gi], pos[3 * gi + 1], pos[3 * gi + 2]};
  int c = 0;
  for (int k = 0; k < n; ++k) {
    const int32_t gj = nl.j[gi * nl.kcap + k];
    const int32_t img = nl.img[gi * nl.kcap + k];
    if (!canonical_entry(gi, gj, img)) continue;
    int nn[3];
    unpack_img(img, nn);
    const double pj[3] = {pos[3 * (int64_t)gj], pos[3 * (int64_t)gj + 1], pos[3 * (int64_t)gj + 2]};
    double v[3];
    int r[3];
    c += exact_row(S, li, (int)(gj - S.atom0), pi, pj, nn, v, r);
  }
  rowcnt[li] = c;
}

__global__ void k_rows_emit(const ynb_sys *sys, int s, int64_t npad, const double *pos, ynb_nlist nl,
                            const int32_t *rowoff, ynb_neigh_row *tmp, uint64_t *key, int32_t *idx) {
  const ynb_sys &S = sys[s];
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= S.natom) return;
  const int64_t gi = S.atom0 + li;
  const int n = min(nl.cnt[gi], nl.kcap);
  int64_t o = rowoff[li];
  const int nv = S.nvec;
  const int R0 = nv > 0 ? S.rmax[0] : 0, R1 = nv > 1 ? S.rmax[1] : 0, R2 = nv > 2 ? S.rmax[2] : 0;
  const uint64_t nimg = (uint64_t)(2 * R0 + 1) * (2 * R1 + 1) * (2 * R2 + 1);
  const double pi[3] = {pos[3 * gi], pos[3 * gi + 1], pos[3 * gi + 2]};
  for (int k = 0; k < n; ++k) {
    const int32_t gj = nl.j[gi * nl.kcap + k];
    const int32_t img = nl.img[gi * nl.kcap + k];
    if (!canonical_entry(gi, gj, img)) continue;
    int nn[3];
    unpack_img(img, nn);
    const int lj = (int)(gj - S.atom0);
    const bool self = (lj == li);
    // Canonical entries have i == A (the larger index).
    const double pj[3] = {pos[3 * (int64_t)gj], pos[3 * (int64_t)gj + 1], pos[3 * (int64_t)gj + 2]};
    double v[3];
    int r[3];
    if (!exact_row(S, li, lj, pi, pj, nn, v, r)) continue;
    const double d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    ynb_neigh_row row;
    const bool plus = in_half_space(r);
    int rho[3];
    if (plus) {
      row.a = li; row.b = lj;
      row.dx = v[0]; row.dy = v[1]; row.dz = v[2];
      rho[0] = r[0]; rho[1] = r[1]; rho[2] = r[2];
    } else {
      // Yaff's sign -1 row: (-delta0) + rho.R with rho = -r (nlist.c:120-123),
      // evaluated as such so that signed zeros match as well.
      rho[0] = -r[0]; rho[1] = -r[1]; rho[2] = -r[2];
      double d0[3] = {0.0, 0.0, 0.0};
      int m[3];
      if (!self) {
        d0[0] = pj[0] - pi[0];
        d0[1] = pj[1] - pi[1];
        d0[2] = pj[2] - pi[2];
        mic_exact(d0, S.rvecs, S.gvecs, nv, m);
      }
      double w[3] = {-d0[0], -d0[1], -d0[2]};
      add_vec_exact(w, S.rvecs, nv, rho);
      row.a = lj; row.b = li;
      row.dx = w[0]; row.dy = w[1]; row.dz = w[2];
    }
    row.d = sqrt(d2);
    row.r0 = rho[0]; row.r1 = rho[1]; row.r2 = rho[2];
    tmp[o] = row;
    // Yaff's loop order (nlist.c:51-120): A, then B, then the image in
    // odometer order, then sign + before sign -.
    const uint64_t pidx = (uint64_t)li * (uint64_t)S.natom + (uint64_t)lj;
    const uint64_t iidx = ((uint64_t)(rho[2] + R2) * (2 * R1 + 1) + (uint64_t)(rho[1] + R1)) * (2 * R0 + 1) +
                          (uint64_t)(rho[0] + R0);
    key[o] = (pidx * nimg + iidx) * 2 + (plus ? 0 : 1);
    idx[o] = (int32_t)o;
    ++o;
  }
}

__global__ void k_rows_gather(int64_t n, const ynb_neigh_row *tmp, const int32_t *idx, ynb_neigh_row *rows) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < n) rows[q] = tmp[idx[q]];
}

// nlist.c:183-232, one thread per row.
__global__ void k_recompute(ynb_sys S, const double *pos, const double *pos_old, ynb_neigh_row *rows,
                            int64_t n) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  ynb_neigh_row w = rows[q];
  const int64_t a = S.atom0 + w.a, b = S.atom0 + w.b;
  double d0[3] = {pos_old[3 * b] - pos_old[3 * a], pos_old[3 * b + 1] - pos_old[3 * a + 1],
                  pos_old[3 * b + 2] - pos_old[3 * a + 2]};
  int center[3] = {0, 0, 0};
  for (int i = 0; i < S.nvec; ++i)
    center[i] = (int)(-ceil(S.gvecs[3 * i] * d0[0] + S.gvecs[3 * i + 1] * d0[1] + S.gvecs[3 * i + 2] * d0[2] - 0.5));
  d0[0] = pos[3 * b] - pos[3 * a];
  d0[1] = pos[3 * b + 1] - pos[3 * a + 1];
  d0[2] = pos[3 * b + 2] - pos[3 * a + 2];
  add_vec_exact(d0, S.rvecs, S.nvec, center);
  const int r[3] = {(int)w.r0, (int)w.r1, (int)w.r2};
  add_vec_exact(d0, S.rvecs, S.nvec, r);
  w.d = sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
  w.dx = d0[0];
  w.dy = d0[1];
  w.dz = d0[2];
  rows[q] = w;
}

size_t rows_work_layout(int64_t natom, int64_t nrows, char *base, int32_t **rowcnt, int32_t **rowoff,
                        ynb_neigh_row **tmp, uint64_t **key, uint64_t **key2, int32_t **idx, int32_t **idx2,
                        void **cub, size_t *cub_bytes) {
  size_t scan_bytes = 0, sort_bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t *)nullptr, (int32_t *)nullptr, (int)natom);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                     (int32_t *)nullptr, (int32_t *)nullptr, (int)(nrows > 0 ? nrows : 1));
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  *rowcnt = (int32_t *)take(4 * (natom + 1));
  *rowoff = (int32_t *)take(4 * (natom + 1));
  *tmp = (ynb_neigh_row *)take(sizeof(ynb_neigh_row) * nrows);
  *key = (uint64_t *)take(8 * nrows);
  *key2 = (uint64_t *)take(8 * nrows);
  *idx = (int32_t *)take(4 * nrows);
  *idx2 = (int32_t *)take(4 * nrows);
  *cub_bytes = scan_bytes > sort_bytes ? scan_bytes : sort_bytes;
  *cub = take(*cub_bytes);
  return off;
}

}  // namespace

}  // namespace ynb

using namespace ynb;

extern "C" {

int ynb_sys_plan(ynb_sys *s) {
  if (!s) {
    set_error("ynb_sys_plan: null system");
    return YNB_EINVAL;
  }
  if (s->natom < 0 || s->rc <= 0.0 || s->nvec < 0 || s->nvec > 3) {
    set_error("ynb_sys_plan: invalid natom/rc/nvec (%d, %g, %d)", s->natom, s->rc, s->nvec);
    return YNB_EINVAL;
  }
  s->nbin = 0;
  for (int k = 0; k < 3; ++k) {
    s->grid[k] = 1;
    s->range[k] = 0;
  }
  if (s->nvec < 3 || s->natom == 0) return YNB_OK;
  const double rcp = s->rc * (1.0 + 1e-8) + 1e-6;
  // Bins at least rc/2 wide (2-3 bins searched per side: about 4x fewer
  // candidates per atom than rc-wide bins); rc-wide bins when the search box
  // would exceed MAXNB partner bins (very skewed or small cells).
  int64_t nbin = 1;
  for (int div = 2; div >= 1; --div) {
    nbin = 1;
    for (int k = 0; k < 3; ++k) {
      const double *g = s->gvecs + 3 * k;
      const double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);  // 1/plane spacing
      const double spacing = 1.0 / gn;
      int grid = (int)floor(spacing * div / rcp);
      if (grid < 1) grid = 1;
      if (grid > 1024) grid = 1024;
      // Bins searched on each side: the fractional extent of a sphere of
      // radius rc along axis k is rc*|g_k|.
      int range = (int)ceil(rcp * gn * grid);
      if (range < 1) range = 1;
      s->grid[k] = grid;
      s->range[k] = range;
      nbin *= grid;
    }
    // Keep at least a few atoms per bin on average.
    while (nbin > 1 && nbin * 4 > (int64_t)s->natom * 8) {
      int kmax = 0;
      for (int k = 1; k < 3; ++k)
        if (s->grid[k] > s->grid[kmax]) kmax = k;
      if (s->grid[kmax] == 1) break;
      s->grid[kmax] = (s->grid[kmax] + 1) / 2;
      const double *g = s->gvecs + 3 * kmax;
      const double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      s->range[kmax] = (int)ceil(rcp * gn * s->grid[kmax]);
      if (s->range[kmax] < 1) s->range[kmax] = 1;
      nbin = (int64_t)s->grid[0] * s->grid[1] * s->grid[2];
    }
    if ((int64_t)(2 * s->range[0] + 1) * (2 * s->range[1] + 1) * (2 * s->range[2] + 1) <= MAXNB) break;
  }
  if ((int64_t)(2 * s->range[0] + 1) * (2 * s->range[1] + 1) * (2 * s->range[2] + 1) > MAXNB) {
    // Cutoff spans too many periodic copies for the binned build's LDS
    // tables: use the all-pairs path.
    s->nbin = 0;
    for (int k = 0; k < 3; ++k) {
      s->grid[k] = 1;
      s->range[k] = 0;
    }
    return YNB_OK;
  }
  s->nbin = (int32_t)nbin;
  // The image box |r_k| <= rmax_k never rejects a pair within rc when
  // rc*|g_k| < rmax_k + 1/2 on every axis: r_k = g_k.v - g_k.delta0 with
  // |g_k.v| < rc*|g_k| and |g_k.delta0| <= 1/2 after Yaff's wrap.  Then the
  // cheap distance alone decides away from the rc^2 boundary.
  s->box_safe = 1;
  for (int k = 0; k < 3; ++k) {
    const double *g = s->gvecs + 3 * k;
    const double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (!(s->rc * gn < s->rmax[k] + 0.5 - 1e-6)) s->box_safe = 0;
  }
  return YNB_OK;
}

int ynb_build_search_bins(const ynb_sys *s) {
  if (!s || s->nbin <= 0) return 0;
  return build_search_bins(s->range);
}

size_t ynb_nlist_workspace_bytes(int64_t npad, int32_t nbin, int32_t bin_cap) {
  return carve(nullptr, npad, nbin, bin_cap).total;
}

int ynb_nlist_build(const ynb_batch *b, const double *pos, void *work, size_t work_bytes,
                    const ynb_nlist *nl, ynb_stream stream) {
  if (!b || !nl || !pos || b->npad % 64 != 0 || nl->kcap < 1 ||
      (b->nbin > 0 && (b->max_bin_atoms < 1 || b->max_bin_atoms > 1024)) ||
      ((nl->flags & YNB_NLIST_PREBINNED) && nl->rebuild)) {
    set_error("ynb_nlist_build: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  const int32_t cb = b->max_bin_atoms;
  Work w = carve(work, b->npad, b->nbin, cb);
  if (work_bytes < w.total || !work) {
    set_error("ynb_nlist_build: workspace of %zu bytes, need %zu", work_bytes, w.total);
    return YNB_EWORKSPACE;
  }
  const int64_t npad = b->npad;
  const int tb = 256;
  const int nb = (int)ceil_div(npad, tb);
  if (b->nbin > 0) {
    // bin counts are zero here: the caller zero-initialises the workspace
    // and k_bin_order resets the counts after each build
    if (!(nl->flags & YNB_NLIST_PREBINNED))
      YNB_LAUNCH("nlist_bin", k_bin, dim3(nb), dim3(tb), 0, st, b->sys, b->wave_sys, npad, pos, cb, w.bin_count,
                 w.slot_idx, w.xw, w.wr, nl->status, nl->rebuild);
    YNB_LAUNCH("nlist_bin_order", k_bin_order, dim3(b->nbin), dim3(64), 0, st, cb, w.bin_count, w.slot_idx, w.xw,
               w.wr, w.sxw, w.sw, w.bin_end, b->bin_sys, nl->rebuild);
#define YNB_BUILD(EX, NBB)                                                                                    \
  YNB_LAUNCH("nlist_build_binned", (k_build_binned<EX, NBB>), dim3(b->nbin, ychunks), dim3(BW * 64), 0, st,        \
             b->sys, b->bin_sys, npad, pos, w.bin_end, cb, w.sxw, w.sw, *nl)
    // Bins hold a few atoms on average; the block's wavefronts loop over the
    // atoms beyond BW.  Large batches take one block per bin (a block per BW
    // atoms of the largest bin would mostly launch empty blocks); small ones
    // add blocks over the bin's atoms while there are fewer than 8192 blocks.
    const int yfull = (int)ceil_div(cb, BW);
    const int ywant = (int)(8192 / (b->nbin > 0 ? b->nbin : 1));
    const int ychunks = max(1, min(min(yfull, ywant), 64));
    const bool wide = (b->flags & YNB_BATCH_WIDE_SEARCH) != 0;
    if (b->flags & YNB_BATCH_EXACT_BUILD) {
      if (wide) YNB_BUILD(true, MAXNB); else YNB_BUILD(true, NB_SMALL);
    } else {
      if (wide) YNB_BUILD(false, MAXNB); else YNB_BUILD(false, NB_SMALL);
    }
#undef YNB_BUILD
    YNB_LAUNCHED();
  }
  if (b->flags & YNB_BATCH_ALLPAIRS) {
    YNB_LAUNCH("nlist_build_allpairs", k_build_allpairs, dim3(nb), dim3(tb), 0, st, b->sys, b->wave_sys, npad, pos,
               *nl);
    YNB_LAUNCHED();
  }
  return YNB_OK;
}


int ynb_nlist_skin_check(const ynb_batch *b, const double *pos, double *pos_ref, const double *skin,
                         int32_t *rebuild, ynb_stream stream) {
  if (!b || !pos || !pos_ref || !skin || !rebuild) {
    set_error("ynb_nlist_skin_check: invalid arguments");
    return YNB_EINVAL;
  }
  if (b->nsys == 0) return YNB_OK;
  YNB_LAUNCH("nlist_skin_check", k_skin_check, dim3((unsigned)b->nsys), dim3(256), 0, (hipStream_t)stream, b->sys,
             pos, pos_ref, skin, rebuild);
  YNB_LAUNCHED();
  return YNB_OK;
}

size_t ynb_rows_workspace_bytes(int64_t natom, int64_t nrows) {
  int32_t *a, *c, *e, *f;
  ynb_neigh_row *t;
  uint64_t *k, *k2;
  void *cub;
  size_t cb;
  return rows_work_layout(natom, nrows, nullptr, &a, &c, &t, &k, &k2, &e, &f, &cub, &cb);
}

int ynb_nlist_rows(const ynb_batch *b, int32_t s, const double *pos, const ynb_nlist *nl, void *work,
                   size_t work_bytes, ynb_neigh_row *rows, int64_t cap, int64_t *count, ynb_stream stream) {
  if (!b || !nl || !count || s < 0 || s >= b->nsys) {
    set_error("ynb_nlist_rows: invalid arguments");
    return YNB_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  ynb_sys S;
  YNB_HIP(hipMemcpyAsync(&S, b->sys + s, sizeof(ynb_sys), hipMemcpyDeviceToHost, st));
  YNB_HIP(hipStreamSynchronize(st));
  const int64_t natom = S.natom;
  const int64_t nrows_cap = cap > 0 ? cap : 0;
  int32_t *rowcnt, *rowoff, *idx, *idx2;
  ynb_neigh_row *tmp;
  uint64_t *key, *key2;
  void *cub;
  size_t cub_bytes;
  const size_t need = rows_work_layout(natom, nrows_cap, (char *)work, &rowcnt, &rowoff, &tmp, &key, &key2, &idx,
                                       &idx2, &cub, &cub_bytes);
  if (!work || work_bytes < need) {
    set_error("ynb_nlist_rows: workspace of %zu bytes, need %zu", work_bytes, need);
    return YNB_EWORKSPACE;
  }
  if (natom == 0) {
    *count = 0;
    return YNB_OK;
  }
  const int tb = 256;
  const int nb = (int)ceil_div(natom, tb);
  hipLaunchKernelGGL(k_rows_count, dim3(nb), dim3(tb), 0, st, b->sys, s, b->npad, pos, *nl, rowcnt);
  YNB_LAUNCHED();
  size_t cb = cub_bytes;
  YNB_HIP(hipcub::DeviceScan::ExclusiveSum(cub, cb, rowcnt, rowoff, (int)natom, st));
  int32_t last[2];
  YNB_HIP(hipMemcpyAsync(&last[0], rowoff + natom - 1, 4, hipMemcpyDeviceToHost, st));
  YNB_HIP(hipMemcpyAsync(&last[1], rowcnt + natom - 1, 4, hipMemcpyDeviceToHost, st));
  YNB_HIP(hipStreamSynchronize(st));
  const int64_t total = (int64_t)last[0] + last[1];
  *count = total;
  if (total > nrows_cap || !rows) return YNB_OK;
  hipLaunchKernelGGL(k_rows_emit, dim3(nb), dim3(tb), 0, st, b->sys, s, b->npad, pos, *nl, rowoff, tmp, key, idx);
  YNB_LAUNCHED();
  if (total > 0) {
    cb = cub_bytes;
    YNB_HIP(hipcub::DeviceRadixSort::SortPairs(cub, cb, key, key2, idx, idx2, (int)total, 0, 64, st));
    hipLaunchKernelGGL(k_rows_gather, dim3((int)ceil_div(total, tb)), dim3(tb), 0, st, total, tmp, idx2, rows);
    YNB_LAUNCHED();
  }
  YNB_HIP(hipStreamSynchronize(st));
  return YNB_OK;
}

int ynb_nlist_recompute(const ynb_sys *h_sys, const double *pos, const double *pos_old, ynb_neigh_row *rows,
                        int64_t nrow, ynb_stream stream) {
  if (!h_sys || nrow < 0) {
    set_error("ynb_nlist_recompute: invalid arguments");
    return YNB_EINVAL;
  }
  if (nrow == 0) return YNB_OK;
  hipLaunchKernelGGL(k_recompute, dim3((int)ceil_div(nrow, 256)), dim3(256), 0, (hipStream_t)stream, *h_sys, pos,
                     pos_old, rows, nrow);
  YNB_LAUNCHED();
  return YNB_OK;
}

}  // extern "C"
