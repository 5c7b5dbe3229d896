// Device code shared by the single-frame (dts.hip) and batched (batch.hip)
// solvers: colour conversion, differentials, the fixed-point ct scan, the
// window bounds, the horizontal segment and vertical tile passes.  Kernel
// entry points live in the .hip files; everything here is inline device code.
#pragma once

#include "common.h"

namespace dtsamd {

// src/domain_transform_common_kernels.cuh:9-10
constexpr float kINV255_F = 0.00392156863f;
constexpr float kINV255SQ_F = 0.0000153787f;
constexpr double kFX = 4294967296.0;  // 2^32
constexpr int kMaxDim = 32767;        // bounds are packed as 2 x 16 bit

// --------------------------------------------------------------- colour
// src/color_space.cuh:11-40: double arithmetic, truncation, alpha kept.
// Contraction is off so the truncated values equal the oracle's.
__device__ __forceinline__ uchar4 convert_px(uchar4 p, int cs) {
#pragma clang fp contract(off)
  const double r = p.x, g = p.y, b = p.z;
  const double yv = r * 0.299 + 0.587 * g + 0.114 * b;
  if (cs == DTS_COLOR_YCBCR) {
    const double cb = 128 - r * 0.168736 - 0.331264 * g + 0.5 * b;
    const double cr = 128 + r * 0.5 - 0.418688 * g - 0.081312 * b;
    return make_uchar4((unsigned char)yv, (unsigned char)cb,
                       (unsigned char)cr, p.w);
  }
  const unsigned char y8 = (unsigned char)yv;
  return make_uchar4(y8, y8, y8, p.w);
}

// -------------------------------------------------------- differentials
// src/domain_transform_common_kernels.cuh:71-116, in double with the fp32
// constants; neighbours outside the image are (0,0,0,0).
// The step depends on the guide only through the integer four-channel SSD.
__device__ __forceinline__ double dt_diff_ssd(int ssd, double q) {
#pragma clang fp contract(off)
  const double s = (double)ssd;
  double t = s * (double)kINV255SQ_F;
  t = t * q;
  t = t * q;
  return sqrt(t + 1.0);
}

__device__ __forceinline__ double dt_diff(uchar4 a, uchar4 b, double q) {
  const int dx = (int)a.x - (int)b.x, dy = (int)a.y - (int)b.y;
  const int dz = (int)a.z - (int)b.z, dw = (int)a.w - (int)b.w;
  return dt_diff_ssd(dx * dx + dy * dy + dz * dz + dw * dw, q);
}

__device__ __forceinline__ int64_t to_fx(double d) { return __double2ll_rn(d * kFX); }

// ------------------------------------------------------------ integrate
// ct_H: inclusive scan of the 32.32 fixed-point dH along each row.  One
// 256-thread workgroup per row, 4 items per thread per 1024-wide chunk, a
// running carry across chunks.  Integer sums make the result independent of
// the scan order.
__device__ __forceinline__ void ct_line(const uchar4* __restrict__ row, int W, double q,
                                       int64_t* __restrict__ out) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  __shared__ int64_t wsum[4];
  int64_t carry = 0;
  for (int base = 0; base < W; base += 1024) {
    int64_t v[4];
    const int x0 = base + tid * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int x = x0 + k;
      v[k] = 0;
      if (x < W) {
        const uchar4 prv = x > 0 ? row[x - 1] : make_uchar4(0, 0, 0, 0);
        v[k] = to_fx(dt_diff(row[x], prv, q));
      }
    }
    v[1] += v[0];
    v[2] += v[1];
    v[3] += v[2];
    const int64_t tot = v[3];
    const int64_t inc = wave_incl_scan_i64_dpp(tot);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int64_t off = carry + inc - tot, all = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const int64_t s = wsum[w];
      if (w < wid) off += s;
      all += s;
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (x0 + k < W) out[x0 + k] = v[k] + off;
    carry += all;
    __syncthreads();
  }
}

// One wavefront per line of up to 64 NC pixels: chunk c holds pixels
// 64 c + lane, all NC chunks loaded (coalesced) before any arithmetic; the
// previous pixel comes from the neighbouring lane (lane 0: the last lane of
// the previous chunk); each chunk is one DPP wavefront scan plus the running
// carry, stored coalesced.
template <int NC>
__device__ __forceinline__ void ct_line_wave(const uchar4* __restrict__ row, int W, double q,
                                            int64_t* __restrict__ out, int lane) {
  uint32_t px[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int i = 64 * c + lane;
    uint32_t v = 0;
    if (i < W) {
      const uchar4 p = row[i];
      v = (uint32_t)p.x | ((uint32_t)p.y << 8) | ((uint32_t)p.z << 16) | ((uint32_t)p.w << 24);
    }
    px[c] = v;
  }
  auto unpack = [](uint32_t v) {
    return make_uchar4(v & 0xff, (v >> 8) & 0xff, (v >> 16) & 0xff, v >> 24);
  };
  int64_t carry = 0;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int i = 64 * c + lane;
    if (64 * c >= W) break;
    uint32_t prv = (uint32_t)__shfl_up((int)px[c], 1, 64);
    if (lane == 0)
      prv = c > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)px[c > 0 ? c - 1 : 0], 63) : 0u;
    const int64_t d = i < W ? to_fx(dt_diff(unpack(px[c]), unpack(prv), q)) : 0;
    const int64_t inc = wave_incl_scan_i64_dpp(d) + carry;
    if (i < W) out[i] = inc;
    carry = ((int64_t)__builtin_amdgcn_readlane((int)(inc >> 32), 63) << 32) |
            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)inc, 63);
  }
}

// 32.32 fixed point of d >= 0 for d * 2^32 < 2^52: adding 2^52 rounds the
// (exact) product to an integer, ties to even, as __double2ll_rn does, and
// that integer is the low mantissa bits of the sum.
__device__ __forceinline__ int64_t to_fx_small(double d) {
  constexpr double k52 = 4503599627370496.0;  // 2^52
  return __double_as_longlong(d * kFX + k52) - __double_as_longlong(k52);
}

// Largest q = sigma_s / sigma_r for which every step of dt_diff stays below
// 2^20, so to_fx_small applies: d^2 = 1 + s * kINV255SQ_F * q^2 with
// s <= 4 * 255^2, i.e. d^2 <= 1 + 4.0001 q^2 < 2^40.
constexpr double kFxSmallMaxQ = 524000.0;

// One line of up to 2048 pixels, length a multiple of 4 and 16-byte aligned
// rows, by the four wavefronts of a workgroup.  Lane k of a wavefront holds
// 4 consecutive pixels of each 256-pixel chunk (one 16-byte load), so the
// int64 wavefront scan runs once per 4 pixels: the lane sums its 4 steps
// serially, the scan of the lane totals gives the offsets.  Wavefront w
// takes ncw = ceil(chunks / 4) chunks from chunk w * ncw on, scanned from a
// zero carry; the wavefront totals meet in LDS (one barrier) and each
// wavefront adds the totals before it as it stores (two 16-byte stores per
// lane and chunk).  SMALL: every step below 2^20 (to_fx_small).  All 256
// threads must call it.
template <int NCWMAX, bool SMALL>
__device__ __forceinline__ void ct_line_quad(const uchar4* __restrict__ row, int L, double q,
                                            int64_t* __restrict__ out, int64_t* wtot) {
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ncw = (((L + 255) >> 8) + 3) >> 2;
  const int p0 = 256 * w * ncw;  // first pixel of the wavefront
  const uint32_t* row32 = reinterpret_cast<const uint32_t*>(row);
  uint4 px[NCWMAX];
#pragma unroll
  for (int c = 0; c < NCWMAX; c++) {
    const int i = p0 + 256 * c + 4 * lane;
    px[c] = c < ncw && i < L ? *reinterpret_cast<const uint4*>(row32 + i) : make_uint4(0, 0, 0, 0);
  }
  // The pixel before the wavefront's first one (black before the line).
  const uint32_t before = p0 > 0 && p0 < L ? row32[p0 - 1] : 0u;
  auto unpack = [](uint32_t v) {
    return make_uchar4(v & 0xff, (v >> 8) & 0xff, (v >> 16) & 0xff, v >> 24);
  };
  auto fx = [q, unpack](uint32_t a, uint32_t b) -> int64_t {
    const double d = dt_diff(unpack(a), unpack(b), q);
    return SMALL ? to_fx_small(d) : to_fx(d);
  };
  int64_t r[NCWMAX][4];
  int64_t carry = 0;
#pragma unroll
  for (int c = 0; c < NCWMAX; c++) {
#pragma unroll
    for (int j = 0; j < 4; j++) r[c][j] = 0;
    if (c >= ncw || p0 + 256 * c >= L) continue;
    const int i = p0 + 256 * c + 4 * lane;
    uint32_t prv = (uint32_t)__shfl_up((int)px[c].w, 1, 64);
    if (lane == 0)
      prv = c > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)px[c > 0 ? c - 1 : 0].w, 63) : before;
    int64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (i < L) {
      s0 = fx(px[c].x, prv);
      s1 = s0 + fx(px[c].y, px[c].x);
      s2 = s1 + fx(px[c].z, px[c].y);
      s3 = s2 + fx(px[c].w, px[c].z);
    }
    const int64_t ex = wave_incl_scan_i64_dpp(s3) - s3 + carry;
    r[c][0] = s0 + ex;
    r[c][1] = s1 + ex;
    r[c][2] = s2 + ex;
    r[c][3] = s3 + ex;
    carry = ((int64_t)__builtin_amdgcn_readlane((int)(r[c][3] >> 32), 63) << 32) |
            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)r[c][3], 63);
  }
  if (lane == 0) wtot[w] = carry;
  __syncthreads();
  int64_t off = 0;
  for (int k = 0; k < w; k++) off += wtot[k];
#pragma unroll
  for (int c = 0; c < NCWMAX; c++) {
    const int i = p0 + 256 * c + 4 * lane;
    if (c < ncw && i < L) {
      longlong2* o = reinterpret_cast<longlong2*>(out + i);
      o[0] = make_longlong2(r[c][0] + off, r[c][1] + off);
      o[1] = make_longlong2(r[c][2] + off, r[c][3] + off);
    }
  }
}

// Scan of one line by one wavefront (any W; lines longer than 2048 pixels
// fall back to the workgroup loop, which needs all 256 threads).
__device__ __forceinline__ void ct_line_w(const uchar4* __restrict__ row, int W, double q,
                                         int64_t* __restrict__ out, int lane) {
  if (W <= 256)
    ct_line_wave<4>(row, W, q, out, lane);
  else if (W <= 512)
    ct_line_wave<8>(row, W, q, out, lane);
  else if (W <= 1024)
    ct_line_wave<16>(row, W, q, out, lane);
  else
    ct_line_wave<32>(row, W, q, out, lane);
}

// --------------------------------------------------------------- bounds
// A bounds plane holds one packed window per position, (p - l) | (u - p) << S,
// with p - l <= E + 1 and u - p <= E.  Narrow planes (NB) have 2-byte
// elements and S = 8, for extents E <= kNarrowBoundsMaxE; wide planes have
// 4-byte elements and S = 16.  The passes read the plane every iteration, so
// the narrow form saves 2 bytes per pixel and pass.
constexpr int kNarrowBoundsMaxE = 254;

template <bool NB>
__device__ __forceinline__ uint32_t pack_bounds(int dl, int du) {
  return (uint32_t)dl | ((uint32_t)du << (NB ? 8 : 16));
}
template <bool NB>
__device__ __forceinline__ int bounds_dl(uint32_t b) {  // p - l
  return (int)(b & (NB ? 0xffu : 0xffffu));
}
template <bool NB>
__device__ __forceinline__ int bounds_du(uint32_t b) {  // u - p
  return (int)(b >> (NB ? 8 : 16));
}
// Element idx of a bounds plane (32-bit byte offsets, as ld32).
template <bool NB>
__device__ __forceinline__ uint32_t ld_bounds(const void* base, unsigned idx) {
  const char* p = reinterpret_cast<const char*>(base);
  if constexpr (NB)
    return *reinterpret_cast<const uint16_t*>(p + (idx << 1));
  else
    return *reinterpret_cast<const uint32_t*>(p + (idx << 2));
}
template <bool NB>
__device__ __forceinline__ void st_bounds(void* base, size_t idx, uint32_t v) {
  if constexpr (NB)
    reinterpret_cast<uint16_t*>(base)[idx] = (uint16_t)v;
  else
    reinterpret_cast<uint32_t*>(base)[idx] = v;
}

// src/domain_transform_filter.cuh:7-28.  The reference scans linearly from p;
// ct is strictly increasing (d >= 1), so the first hit of the linear search
// is found by bisection inside the radius E = floor(sigma) + 1 that d >= 1
// guarantees.  No hit keeps the reference's collapse (l = p-1, u = p).
// Packed result in the wide form.
__device__ __forceinline__ uint32_t find_bounds(const int64_t* __restrict__ ct, int L, int p,
                                       int64_t sfx, int E) {
  const int64_t cp = ct[p];
  int l = p - 1;
  const int lo_i = max(0, p - E - 1);
  if (lo_i <= p - 1 && cp - ct[lo_i] > sfx) {
    int lo = lo_i, hi = p;  // predicate true at lo, false at hi
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (cp - ct[mid] > sfx)
        lo = mid;
      else
        hi = mid;
    }
    l = lo;
  }
  int u = p;
  const int hi_i = min(L - 1, p + E + 1);
  if (hi_i >= p + 1 && ct[hi_i] - cp > sfx) {
    int lo = p, hi = hi_i;  // predicate false at lo, true at hi
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ct[mid] - cp > sfx)
        hi = mid;
      else
        lo = mid;
    }
    u = hi - 1;
  }
  return (uint32_t)(p - l) | ((uint32_t)(u - p) << 16);
}

// Compiler and LDS ordering point for one wavefront: its LDS writes are
// visible to all of its lanes afterwards.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bounds of a tile of 64 positions x 32 lines with the lines' ct staged in
// LDS: [p0 - E - 1, p0 + 64 + E + 1) per line, then bisection in LDS.
// kCols: lines are image columns (ct_V transposed) and the output is written
// row-major through an LDS tile; otherwise lines are rows.
constexpr int kBPos = 64, kBLines = 32;
constexpr int kYCols = 64;  // columns per vertical-pass workgroup
constexpr int kYRows = 32;  // output rows per vertical-pass workgroup

// Window search and output of one staged tile.  st[li * stage_len + i - s0]
// holds ct(i) of line l0 + li for i in [s0, s1), up to a per-line constant
// (only differences are used).  Each thread owns one position of the tile's
// 8 lines of its wavefront; the 16 searches (lower and upper bound of every
// line) advance in lockstep, so their LDS reads overlap and the wavefront
// does not diverge.  Each search keeps one index, the last position known to
// satisfy its predicate, and tries steps 2^(n-1), ..., 2, 1 ahead of it
// (a branch-free select per step); n = ceil(log2(E + 1)) steps cover the
// E + 1 candidates, 4 steps when two probes 16 positions from the pixel show
// that every window of the wavefront ends within them.  Probes past the
// interval are clamped to its end, where the predicate is false.
// Window extents for the vertical pass are kept per (32-row band, 32-column
// group): the vertical bounds tile of rows [p0, p0 + 64) and columns
// [l0, l0 + 32) is the only writer of its two bands of group l0 / 32, so no
// atomics or clearing are needed.  ext_groups = ceil(W / 32).
// kCols: after the search the staged ct is dead, and its LDS holds the
// [kBPos][kBLines + 1] transpose tile of the results (the launch gives at
// least kBoundsResBytes of dynamic LDS).
constexpr size_t kBoundsResBytes = sizeof(uint32_t) * kBPos * (kBLines + 1);

// Near windows of the search: probes kBNear positions from the pixel, and
// kBNearSteps bisection steps between a probe and the pixel.
constexpr int kBNear = 16, kBNearSteps = 4;

// No restaging: the stage holds the whole search range.
struct NoRestage {
  __device__ bool operator()() const { return false; }
};

// restage: called (by whole wavefronts) when the near test fails for a
// wavefront whose stage holds only the near range; it stages the
// wavefront's lines over the whole range and returns true (from another
// base, so the searches reload ct(p)).  Only with one group of lines per
// wavefront (JH = kBLines / NWB).
template <bool kCols, bool NB, int JH = kBLines / 4, int NWB = 4, class Restage = NoRestage>
__device__ __forceinline__ void bounds_search(int64_t* st, int stage_len, int s0, int L,
                                              int nlines, int nl, int64_t sfx, int E,
                                              void* __restrict__ out,
                                              int* __restrict__ yext, int ext_groups, int p0,
                                              int l0, Restage restage = Restage()) {
  __shared__ int ext_red[NWB][2][2];
  constexpr int J = kBLines / NWB;  // lines per wavefront
  static_assert(J % JH == 0, "whole groups of lines");
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int p = p0 + lane;
  const int pc = min(p, L - 1);
  int nsteps = 0;
  while ((1 << nsteps) < E + 1) nsteps++;
  const int lo_i = max(0, pc - E - 1);
  const int hi_i = min(L - 1, pc + E + 1);
  // Lower bound: the last i in [lo_i, pc) with ct(p) - ct(i) > sfx, i.e.
  // ct(i) < tl = ct(p) - sfx (no overflow: ct >= 0); the predicate is false
  // at pc.  Upper bound: the last i in [pc, hi_i) with ct(i) - ct(p) <= sfx,
  // i.e. ct(i) <= tu = ct(p) + sfx (saturated at INT64_MAX); it is true at
  // pc and, when the window ends inside the line (hasu), false at hi_i.
  // The searches keep LDS byte addresses (lb, ub, and their limits pcx,
  // hix), so a probe is one add and one min.
  using lds_i64 = const int64_t __attribute__((address_space(3)));
  const uint32_t sb = (uint32_t)(uintptr_t)(lds_i64*)st;
  auto ld = [](uint32_t a) { return *(lds_i64*)(uintptr_t)a; };
  int mlo = 0, mhi = 0;
  uint32_t bv[J];
  // JH lines at a time (fewer registers for the searches in flight).
#pragma unroll
  for (int jb = 0; jb < J; jb += JH) {
    uint32_t lb[JH], ub[JH], pcx[JH], hix[JH];
    bool hasl[JH], hasu[JH];
    int64_t tl[JH], tu[JH];
#pragma unroll
    for (int j = 0; j < JH; j++) {
      const int li = min(q * J + jb + j, nl - 1);
      // LDS address of ct(i) is b0 + 8 i for i in [s0, s1) (b0 itself may wrap).
      const uint32_t b0 = sb + 8u * (uint32_t)(li * stage_len - s0);
      pcx[j] = b0 + 8u * (uint32_t)pc;
      const int64_t cp = ld(pcx[j]);
      tl[j] = cp - sfx;
      tu[j] = sfx > INT64_MAX - cp ? INT64_MAX : cp + sfx;
    }
    // Near windows: probes kBNear positions from pc on each side.  When, for
    // every line and position of the wavefront, the lower window ends at or
    // after the lower probe (or the interval starts there) and the upper one
    // before the upper probe (or the interval ends there), kBNearSteps steps
    // from the probes find the same bounds as nsteps from the interval ends:
    // ct increases along a line, so each predicate is monotone.
    bool near = false;
    if (nsteps > kBNearSteps) {
      const int lo2 = max(lo_i, pc - kBNear), hi2 = min(hi_i, pc + kBNear);
      bool ok = true;
#pragma unroll
      for (int j = 0; j < JH; j++) {
        lb[j] = pcx[j] - 8u * (uint32_t)(pc - lo2);
        hix[j] = pcx[j] + 8u * (uint32_t)(hi2 - pc);
        hasl[j] = lo2 <= pc - 1 && ld(lb[j]) < tl[j];
        hasu[j] = hi2 >= pc + 1 && ld(hix[j]) > tu[j];
        const bool live = q * J + jb + j < nl && p < L;
        ok = ok && (!live || ((hasl[j] || lo2 == lo_i) && (hasu[j] || hi2 == hi_i)));
      }
      near = __all(ok);
    }
    if (!near) {
      if (nsteps > kBNearSteps && restage()) {
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < JH; j++) {
          const int64_t cp = ld(pcx[j]);
          tl[j] = cp - sfx;
          tu[j] = sfx > INT64_MAX - cp ? INT64_MAX : cp + sfx;
        }
      }
#pragma unroll
      for (int j = 0; j < JH; j++) {
        lb[j] = pcx[j] - 8u * (uint32_t)(pc - lo_i);
        hix[j] = pcx[j] + 8u * (uint32_t)(hi_i - pc);
        hasl[j] = lo_i <= pc - 1 && ld(lb[j]) < tl[j];
        hasu[j] = hi_i >= pc + 1 && ld(hix[j]) > tu[j];
      }
    }
#pragma unroll
    for (int j = 0; j < JH; j++) ub[j] = pcx[j];
    // A candidate past the interval is clamped to its end, where the
    // predicate is false (lower) or the result is unused (upper, !hasu).
    const int steps = near ? kBNearSteps : nsteps;
    for (uint32_t s8 = steps > 0 ? 8u << (steps - 1) : 0u; s8 >= 8u; s8 >>= 1) {
      // All 2 JH LDS reads of the step are issued before the first compare.
      uint32_t al[JH], au[JH];
      int64_t cl[JH], cu[JH];
#pragma unroll
      for (int j = 0; j < JH; j++) {
        al[j] = min(lb[j] + s8, pcx[j]);
        au[j] = min(ub[j] + s8, hix[j]);
        cl[j] = ld(al[j]);
        cu[j] = ld(au[j]);
      }
#pragma unroll
      for (int j = 0; j < JH; j++) {
        lb[j] = cl[j] < tl[j] ? al[j] : lb[j];
        ub[j] = cu[j] <= tu[j] ? au[j] : ub[j];
      }
    }
#pragma unroll
    for (int j = 0; j < JH; j++) {
      const int li = q * J + jb + j;
      uint32_t bval = 0;
      if (li < nl && p < L) {
        const int l = hasl[j] ? pc - (int)((pcx[j] - lb[j]) >> 3) : p - 1;
        const int u = hasu[j] ? pc + (int)((ub[j] - pcx[j]) >> 3) : p;
        bval = pack_bounds<NB>(p - l, u - p);
        mlo = max(mlo, p - l);
        mhi = max(mhi, u - p);
      }
      bv[jb + j] = bval;
      if (!kCols && li < nl && p < L) st_bounds<NB>(out, (size_t)(l0 + li) * L + p, bval);
    }
  }
  if (kCols) {
    // Extents of the vertical-pass tiles: lanes 0-31 and 32-63 cover rows
    // of two consecutive 32-row tiles.
#pragma unroll
    for (int d = 16; d >= 1; d >>= 1) {
      mlo = max(mlo, __shfl_xor(mlo, d, 64));
      mhi = max(mhi, __shfl_xor(mhi, d, 64));
    }
    if ((lane & 31) == 0) {
      ext_red[q][lane >> 5][0] = mlo;
      ext_red[q][lane >> 5][1] = mhi;
    }
    __syncthreads();  // every wavefront is done reading the staged ct
    uint32_t* res = reinterpret_cast<uint32_t*>(st);  // [kBPos][kBLines + 1]
#pragma unroll
    for (int j = 0; j < J; j++) res[lane * (kBLines + 1) + q * J + j] = bv[j];
    __syncthreads();
    if (threadIdx.x < 2 && p0 + 32 * (int)threadIdx.x < L) {
      const int band = p0 / 32 + (int)threadIdx.x;
      int elo = 0, ehi = 0;
#pragma unroll
      for (int w = 0; w < NWB; w++) {
        elo = max(elo, ext_red[w][threadIdx.x][0]);
        ehi = max(ehi, ext_red[w][threadIdx.x][1]);
      }
      const int e = band * ext_groups + l0 / 32;
      yext[2 * e] = elo;
      yext[2 * e + 1] = ehi;
    }
    // out is row-major over (position = y, line = x): out[p * nlines + line].
    for (int k = threadIdx.x; k < kBPos * kBLines; k += 64 * NWB) {
      const int pi = k / kBLines, li = k % kBLines;
      if (p0 + pi < L && li < nl)
        st_bounds<NB>(out, (size_t)(p0 + pi) * nlines + l0 + li, res[pi * (kBLines + 1) + li]);
    }
  }
}

// Bounds of a tile of 64 positions x 32 lines with the lines' ct staged in
// LDS: [p0 - E - 1, p0 + 64 + E + 1) per line, then bisection in LDS.
// kCols: lines are image columns (ct_V transposed) and the output is written
// row-major through an LDS tile; otherwise lines are rows.
template <bool kCols, bool NB>
__device__ __forceinline__ void bounds_tile(const int64_t* __restrict__ ct, int L, int nlines,
                                            int64_t sfx, int E, int stage_len,
                                            void* __restrict__ out, int* __restrict__ yext,
                                            int ext_groups, int p0, int l0, int64_t* st) {
  // st: dynamic LDS [kBLines][stage_len]
  const int s0 = max(0, p0 - E - 1);
  const int s1 = min(L, p0 + kBPos + E + 1);
  const int sn = s1 - s0;
  const int nl = min(kBLines, nlines - l0);
  // Stage without index division: the two halves of the workgroup take the
  // even and odd lines, 128 positions per round, 8 lines per thread in
  // flight before their LDS stores (a line at a time would leave one load
  // per thread outstanding and the tile latency-bound).
  {
    constexpr int kU = 8;
    const int64_t* src = ct + (size_t)l0 * L + s0;
    const int kk = threadIdx.x & 127, lg = threadIdx.x >> 7;
    for (int k = kk; k < sn; k += 128) {
      for (int li0 = lg; li0 < nl; li0 += 2 * kU) {
        int64_t v[kU];
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (li0 + 2 * u < nl) v[u] = src[(size_t)(li0 + 2 * u) * L + k];
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (li0 + 2 * u < nl) st[(li0 + 2 * u) * stage_len + k] = v[u];
      }
    }
  }
  __syncthreads();
  bounds_search<kCols, NB>(st, stage_len, s0, L, nlines, nl, sfx, E, out, yext, ext_groups, p0,
                           l0);
}

// ------------------------------------------------ bounds from the guide
// The window search compares only differences of ct along a line, and the
// step between pixels i - 1 and i depends on the guide only through their
// four-channel SSD s in [0, kSsdMax].  A table lut[s] = to_fx(dt_diff_ssd(s,
// q)) (kStepLutLen int64 per q) holds the exact 32.32 step of every s, so a
// bounds tile can form its lines' ct differences from the guide itself: no
// ct plane is written or read, and the double-precision square roots are
// paid once per table instead of once per pixel.
constexpr int kSsdMax = 4 * 255 * 255;
constexpr int kStepLutLen = kSsdMax + 1;

// Four-channel SSD of two packed pixels as |a|^2 + |b|^2 - 2 a.b with the
// 8-bit dot-product instruction (exact in 32-bit integers).
__device__ __forceinline__ int px_ssd(uint32_t a, uint32_t b) {
  const uint32_t aa = __builtin_amdgcn_udot4(a, a, 0u, false);
  const uint32_t bb = __builtin_amdgcn_udot4(b, b, 0u, false);
  return (int)(aa + bb - 2u * __builtin_amdgcn_udot4(a, b, 0u, false));
}

// Row stride (int64 elements) of a guide-staged bounds tile: the stage
// starts at s0 rounded down to a multiple of 4 (up to 3 more positions), and
// a multiple-of-4 stride keeps every lane's 4 values 32-byte aligned.
__host__ __device__ constexpr int guide_stage_stride(int stage_len) {
  return (stage_len + 3 + 3) & ~3;
}

// Stages rel(i) = ct(i) - ct(sa) + const for lines [l0, l0 + nl) over the
// positions [sa, s1) (sa a multiple of 4) at st[li * SL + i - sa], from the
// guide g (line l of length L at g + l * L): rel(i) = rel(i - 1) +
// lut[ssd(g(i), g(i - 1))].  Wavefront w stages lines 8 w .. 8 w + 7
// (clamped to nl - 1), GL lines at a time so their loads and table reads are
// in flight together.  Lane k of chunk c holds positions sa + 256 c + 4 k +
// j (j < 4), sums their steps serially, and one int64 wavefront scan per
// line and chunk adds the offsets of the lanes before it.  vec: lines start
// 16-byte aligned (one 16-byte load per lane and line).
template <int GL, int NWB = 4>
__device__ __forceinline__ void stage_rel_guide(const uchar4* __restrict__ g, int L, int l0, int nl,
                                                int sa, int s1, const int64_t* __restrict__ lut,
                                                bool vec, int64_t* st, int SL) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const uint32_t* g32 = reinterpret_cast<const uint32_t*>(g);
  const int nc = (s1 - sa + 255) >> 8;
  constexpr int LPW = kBLines / NWB;  // lines per wavefront
  static_assert(LPW % GL == 0, "whole groups of lines");
  for (int j0 = 0; j0 < LPW; j0 += GL) {
    int64_t carry[GL];
    uint32_t last[GL];
#pragma unroll
    for (int j = 0; j < GL; j++) {
      carry[j] = 0;
      last[j] = 0;
    }
    for (int c = 0; c < nc; c++) {
      const int i0 = sa + 256 * c + 4 * lane;
      uint4 px[GL];
#pragma unroll
      for (int j = 0; j < GL; j++) {
        const int li = min(LPW * w + j0 + j, nl - 1);
        const uint32_t* row = g32 + (size_t)(l0 + li) * L;
        if (vec) {
          px[j] = i0 < s1 ? *reinterpret_cast<const uint4*>(row + i0) : make_uint4(0, 0, 0, 0);
        } else {
          px[j].x = i0 < s1 ? row[i0] : 0u;
          px[j].y = i0 + 1 < s1 ? row[i0 + 1] : 0u;
          px[j].z = i0 + 2 < s1 ? row[i0 + 2] : 0u;
          px[j].w = i0 + 3 < s1 ? row[i0 + 3] : 0u;
        }
      }
      // The step of position sa is arbitrary (only differences of rel are
      // used); lane 0 of the first chunk takes its own pixel as the previous.
      int64_t d[GL][4];
#pragma unroll
      for (int j = 0; j < GL; j++) {
        uint32_t prv = (uint32_t)__shfl_up((int)px[j].w, 1, 64);
        if (lane == 0) prv = c > 0 ? last[j] : px[j].x;
        d[j][0] = lut[px_ssd(px[j].x, prv)];
        d[j][1] = lut[px_ssd(px[j].y, px[j].x)];
        d[j][2] = lut[px_ssd(px[j].z, px[j].y)];
        d[j][3] = lut[px_ssd(px[j].w, px[j].z)];
      }
#pragma unroll
      for (int j = 0; j < GL; j++) {
        const int li = min(LPW * w + j0 + j, nl - 1);
        const int64_t r0 = d[j][0], r1 = r0 + d[j][1], r2 = r1 + d[j][2], r3 = r2 + d[j][3];
        const int64_t ex = wave_incl_scan_i64_dpp(r3) - r3 + (c > 0 ? carry[j] : 0);
        int64_t* o = st + li * SL + (i0 - sa);
        if (i0 + 3 < s1) {
          reinterpret_cast<longlong2*>(o)[0] = make_longlong2(r0 + ex, r1 + ex);
          reinterpret_cast<longlong2*>(o)[1] = make_longlong2(r2 + ex, r3 + ex);
        } else {
          if (i0 < s1) o[0] = r0 + ex;
          if (i0 + 1 < s1) o[1] = r1 + ex;
          if (i0 + 2 < s1) o[2] = r2 + ex;
        }
        const int64_t tot = r3 + ex;
        carry[j] = ((int64_t)__builtin_amdgcn_readlane((int)(tot >> 32), 63) << 32) |
                   (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)tot, 63);
        last[j] = (uint32_t)__builtin_amdgcn_readlane((int)px[j].w, 63);
      }
    }
  }
}

// stage_rel_guide for a range of at most 128 positions (s1 - sa <= 128, sa
// a multiple of 4), two lines per pass: lanes 0-31 take line 2m, lanes
// 32-63 line 2m + 1 of the calling wavefront's lines, lane k' = lane & 31
// holds positions sa + 4 k' + j (j < 4), and one int64 scan per half
// wavefront adds the offsets, so a pair of lines costs what one line of a
// 256-position chunk does.  Position i of line li goes to
// st[li * SL + i - sbase] (sbase <= sa, a multiple of 4).
template <int NWB>
__device__ __forceinline__ void stage_rel_guide_pairs(const uchar4* __restrict__ g, int L, int l0,
                                                      int nl, int sa, int s1, int sbase,
                                                      const int64_t* __restrict__ lut, bool vec,
                                                      int64_t* st, int SL) {
  constexpr int LPW = kBLines / NWB;  // lines per wavefront
  static_assert(LPW % 2 == 0, "pairs of lines");
  constexpr int NP = LPW / 2;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int half = lane >> 5, k = lane & 31;
  const uint32_t* g32 = reinterpret_cast<const uint32_t*>(g);
  const int i0 = sa + 4 * k;
  uint4 px[NP];
  int li[NP];
#pragma unroll
  for (int m = 0; m < NP; m++) {
    li[m] = min(LPW * w + 2 * m + half, nl - 1);
    const uint32_t* row = g32 + (size_t)(l0 + li[m]) * L;
    if (vec) {
      px[m] = i0 < s1 ? *reinterpret_cast<const uint4*>(row + i0) : make_uint4(0, 0, 0, 0);
    } else {
      px[m].x = i0 < s1 ? row[i0] : 0u;
      px[m].y = i0 + 1 < s1 ? row[i0 + 1] : 0u;
      px[m].z = i0 + 2 < s1 ? row[i0 + 2] : 0u;
      px[m].w = i0 + 3 < s1 ? row[i0 + 3] : 0u;
    }
  }
#pragma unroll
  for (int m = 0; m < NP; m++) {
    // The step of position sa is arbitrary (only differences are used):
    // lane k' = 0 of each half takes its own pixel as the previous one.
    uint32_t prv = (uint32_t)__shfl_up((int)px[m].w, 1, 64);
    if (k == 0) prv = px[m].x;
    const int64_t r0 = lut[px_ssd(px[m].x, prv)];
    const int64_t r1 = r0 + lut[px_ssd(px[m].y, px[m].x)];
    const int64_t r2 = r1 + lut[px_ssd(px[m].z, px[m].y)];
    const int64_t r3 = r2 + lut[px_ssd(px[m].w, px[m].z)];
    const int64_t ex = incl_scan_i64_dpp<true>(r3) - r3;
    int64_t* o = st + li[m] * SL + (i0 - sbase);
    if (i0 + 3 < s1) {
      reinterpret_cast<longlong2*>(o)[0] = make_longlong2(r0 + ex, r1 + ex);
      reinterpret_cast<longlong2*>(o)[1] = make_longlong2(r2 + ex, r3 + ex);
    } else {
      if (i0 < s1) o[0] = r0 + ex;
      if (i0 + 1 < s1) o[1] = r1 + ex;
      if (i0 + 2 < s1) o[2] = r2 + ex;
    }
  }
}

// Bounds of a tile of 64 positions x 32 lines as bounds_tile, with the
// lines' ct differences staged from the guide and the step table (SL =
// guide_stage_stride(stage_len) int64 per staged line); GL lines of a
// wavefront staged at a time, JH lines searched at a time, NWB wavefronts.
// Each wavefront stages and searches only its own lines.  With an even
// number of lines per wavefront, lines are staged in pairs: the whole range
// when the search radius is below kBNear (it spans at most 99 positions),
// else the near range [max(0, p0 - kBNear), min(L, p0 + kBPos + kBNear)),
// which holds every position the near test and a near search read (the
// probes lie within kBNear of a pixel of the tile); a wavefront whose
// windows are not all near then restages its lines over the whole range
// (stage_rel_guide) from inside the search.
template <bool kCols, bool NB, int GL = 8, int JH = kBLines / 4, int NWB = 4>
__device__ __forceinline__ void bounds_tile_guide(const uchar4* __restrict__ g, int L, int nlines,
                                                  const int64_t* __restrict__ lut, bool vec,
                                                  int64_t sfx, int E, int SL,
                                                  void* __restrict__ out, int* __restrict__ yext,
                                                  int ext_groups, int p0, int l0, int64_t* st) {
  const int s0 = max(0, p0 - E - 1);
  const int sa = s0 & ~3;
  const int s1 = min(L, p0 + kBPos + E + 1);
  const int nl = min(kBLines, nlines - l0);
  if constexpr ((kBLines / NWB) % 2 == 0 && JH == kBLines / NWB) {
    const bool part = E >= kBNear;
    if (part) {
      const int na = max(0, p0 - kBNear) & ~3, n1 = min(L, p0 + kBPos + kBNear);
      stage_rel_guide_pairs<NWB>(g, L, l0, nl, na, n1, sa, lut, vec, st, SL);
    } else {
      stage_rel_guide_pairs<NWB>(g, L, l0, nl, sa, s1, sa, lut, vec, st, SL);
    }
    __syncthreads();
    auto restage = [&]() -> bool {
      if (!part) return false;
      stage_rel_guide<GL, NWB>(g, L, l0, nl, sa, s1, lut, vec, st, SL);
      return true;
    };
    bounds_search<kCols, NB, JH, NWB>(st, SL, sa, L, nlines, nl, sfx, E, out, yext, ext_groups,
                                      p0, l0, restage);
  } else {
    stage_rel_guide<GL, NWB>(g, L, l0, nl, sa, s1, lut, vec, st, SL);
    __syncthreads();
    bounds_search<kCols, NB, JH, NWB>(st, SL, sa, L, nlines, nl, sfx, E, out, yext, ext_groups,
                                      p0, l0);
  }
}

// ---------------------------------------------------------- line passes
enum PassMode { PM_FILTER = 0, PM_X_L2 = 1, PM_X_CH = 2, PM_Y = 3 };

struct PassArgs {
  const float* vin;       // [H][W]
  const void* bnd;        // [H][W] packed bounds of this direction
  int bnarrow;            // bnd has 2-byte elements (host-side dispatch only)
  const float* t;         // [H][W], horizontal data passes
  const float* c;         // [H][W], horizontal data passes
  float* out;             // [H][W]
  int W, H, ntx, nty;
  int elo_cap, ehi_cap;   // analytic extents the LDS segment is sized for
  int seg_stride;         // floats of LDS per segment array
  const int* yext;        // vertical pass: (lower, upper) extent per (band, group)
  int ext_groups;         // 32-column groups per band
  int tw_div, tw_mod;     // horizontal pass: segment stride in lines + tiles
  float lambda, step, theta;
  const uchar4* left;     // stereo: left guide, right guide, right dH
  const uchar4* right;
  const float* rdH;
  float doff, photo_lambda, grad_alpha;
};

constexpr int kXSeg = 256;  // positions per wavefront in the horizontal pass

// Horizontal pass (box filter, Optimize X half, stereo term).  One wavefront owns a 256-pixel
// segment of one row; lane k owns pixels x0 + 4k .. x0 + 4k + 3.
//  1. loads, all issued together: the lane's packed bounds (and target,
//     confidence) as float4, and v over the segment [sa, sb) =
//     [x0 - Elo + 1, x0 + 256 + Ehi) of the analytic window extents in
//     64-wide coalesced chunks (X passes apply the overwrite to it) -> V;
//  2. prefix sums Q[k] = sum_{sa <= i <= sa+k} (v(i) - vref), vref = v(sa):
//     each lane sums an odd-length run of consecutive elements serially
//     (odd strides keep the LDS reads conflict-free), one DPP wavefront scan
//     of the run totals gives the offsets; subtracting vref keeps the fp32
//     sums small for offset-dominated data;
//  3. B = (Q(u) - Q(l)) / (u - l) + vref, the update, float4 store.
// Divisions use the hardware reciprocal (1 ulp).  No workgroup barrier: each
// wavefront works in its own LDS slice.
// Registers of one horizontal segment: the lane's own bounds, target and
// confidence (pixels x0 + lane + 64 i, i < 4) and NCH 64-wide chunks of the
// segment's v (with confidence and target for the overwrite).
template <int NCH>
struct XBuf {
  uint32_t b[4];
  float t[4], c[4];
  float v[NCH], sc[NCH], st[NCH];
};

__device__ __forceinline__ void xseg_range(const PassArgs& a, int x0, int* sa, int* segn) {
  *sa = max(0, x0 - a.elo_cap + 1);
  *segn = min(a.W, x0 + kXSeg + a.ehi_cap) - *sa;
}

// Issues every global load of one segment (coalesced dwords, 32-bit element
// offsets; chunk indices past the segment are clamped, their values unused).
// NQ <= NCH: chunks actually loaded (the segment needs 64 * NQ >= its length).
template <int MODE, bool OW, int NCH, bool NB, int NQ = NCH>
__device__ __forceinline__ void xseg_load(const PassArgs& a, int line, int x0, int lane, XBuf<NCH>& B,
                                          bool ow = true) {
  static_assert(NQ <= NCH, "chunks loaded <= chunks held");
  constexpr bool kData = (MODE == PM_X_L2 || MODE == PM_X_CH);
  const int W = a.W;
  const unsigned rowoff = (unsigned)line * (unsigned)W;
  int sa, segn;
  xseg_range(a, x0, &sa, &segn);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const unsigned gi = rowoff + (unsigned)min(x0 + lane + 64 * i, W - 1);
    B.b[i] = ld_bounds<NB>(a.bnd, gi);
    if (kData) {
      B.t[i] = ld32(a.t, gi);
      B.c[i] = ld32(a.c, gi);
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const unsigned gi = rowoff + (unsigned)(sa + min(64 * q + lane, segn - 1));
    B.v[q] = ld32(a.vin, gi);
    if (OW && kData) {
      if (ow) {
        B.sc[q] = ld32(a.c, gi);
        B.st[q] = ld32(a.t, gi);
      } else {
        B.sc[q] = 0.f;
        B.st[q] = 0.f;
      }
    }
  }
}

// The horizontal update of pixel p of a row (Optimize X half, stereo term): Bm is the
// window mean of the pass input, rn = 1 / n, v the pass input at p (after the
// overwrite), t and c the target and confidence at p.
template <int MODE, bool OW, bool STEREO>
__device__ __forceinline__ float x_update(const PassArgs& a, unsigned rowoff, int p, float v,
                                          float Bm, float rn, float t, float c, bool ow) {
  constexpr bool kData = (MODE == PM_X_L2 || MODE == PM_X_CH);
  if (MODE == PM_FILTER) return Bm;
  float g = 0.f;
  if (kData) {
    const float cn = c * rn;
    const float d = v - t;
    if (MODE == PM_X_L2)
      g = 2.f * cn * d;
    else
      g = cn * d * __builtin_amdgcn_rsqf(d * d + 1e-6f);
  }
  if (STEREO) {
    const float xi = a.doff + (float)p - v;
    if (xi >= 0.f && xi < (float)a.W) {
      const int xr = (int)xi;
      const uchar4 lv = a.left[rowoff + p];
      const uchar4 rv = a.right[rowoff + xr];
      const float sad = (fabsf((float)((int)lv.x - (int)rv.x)) +
                         fabsf((float)((int)lv.y - (int)rv.y)) +
                         fabsf((float)((int)lv.z - (int)rv.z)) +
                         fabsf((float)((int)lv.w - (int)rv.w))) *
                        kINV255_F;
      const float gr = a.rdH[rowoff + xr];
      g += a.photo_lambda * 2.f * sad * ((gr * gr + (-1.f)) * a.grad_alpha);
    }
  }
  float res = v + a.step * (-(a.lambda * (v - Bm) + 0.5f * g));
  if (OW && kData && ow && c > a.theta) res = t;
  return res;
}

// Scans and updates one segment from its registers; V, Q are the
// wavefront's LDS slices of 64 * (NCH | 1) floats each.  NQ: chunks loaded
// (as in xseg_load).  ow: runtime switch of the overwrite when OW is set.
template <int MODE, bool OW, bool STEREO, int NCH, bool NB, int NQ = NCH>
__device__ __forceinline__ void xseg_process(const PassArgs& a, int line, int x0, int lane,
                                    const XBuf<NCH>& B, float* V, float* Q, bool ow = true) {
  constexpr bool kData = (MODE == PM_X_L2 || MODE == PM_X_CH);
  constexpr int EPL = NQ | 1;  // odd run length, EPL * 64 >= 64 * NQ >= segn
  static_assert(NQ <= NCH, "chunks loaded <= chunks held");
  const int W = a.W;
  const unsigned rowoff = (unsigned)line * (unsigned)W;
  int sa, segn;
  xseg_range(a, x0, &sa, &segn);
  float vals[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    vals[q] = B.v[q];
    if (OW && kData && ow) {
      if (B.sc[q] > a.theta) vals[q] = B.st[q];
    }
  }
  const float vref = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vals[0])));
#pragma unroll
  for (int q = 0; q < NQ; q++) V[64 * q + lane] = vals[q];
  wave_lds_sync();
  // Lane-serial runs of EPL consecutive elements.  Entries at or past segn
  // (unwritten or clamped loads) only reach lanes and entries at or past
  // segn, which are never read.
  const int e0 = lane * EPL;
  float loc[EPL];
  float run = 0.f;
#pragma unroll
  for (int i = 0; i < EPL; i++) {
    run += V[e0 + i] - vref;
    loc[i] = run;
  }
  const float off = wave_incl_scan_dpp(run) - run;
#pragma unroll
  for (int i = 0; i < EPL; i++) Q[e0 + i] = loc[i] + off;
  wave_lds_sync();

  // Every LDS read of the segment is issued before the first use; indices
  // are clamped into the segment (results of positions past W are not
  // stored).
  float qlv[4], quv[4], vv[4];
  int nw[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int p = x0 + lane + 64 * i;
    const int l = p - bounds_dl<NB>(B.b[i]);
    const int u = p + bounds_du<NB>(B.b[i]);
    const int il = l - sa, iu = max(0, min(u - sa, segn - 1));
    const float rl = Q[max(0, min(il, segn - 1))];
    quv[i] = Q[iu];
    vv[i] = V[min(p - sa, segn - 1)];
    qlv[i] = il >= 0 ? rl : 0.f;
    nw[i] = u - l;
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int p = x0 + lane + 64 * i;
    if (p >= W) continue;
    const float rn = __builtin_amdgcn_rcpf((float)nw[i]);
    const float Bm = (quv[i] - qlv[i]) * rn + vref;
    st32(a.out, rowoff + (unsigned)p,
         x_update<MODE, OW, STEREO>(a, rowoff, p, vv[i], Bm, rn, kData ? B.t[i] : 0.f,
                                    kData ? B.c[i] : 0.f, ow));
  }
  wave_lds_sync();
}

// Horizontal pass (box filter, Optimize X half, stereo term).  One wavefront owns 256-pixel row
// segments; lane k owns pixels x0 + k + 64 i (i < 4).  Per segment:
//  1. loads, all issued together: the lane's packed bounds (and target,
//     confidence) and v over [sa, sb) = [x0 - Elo + 1, x0 + 256 + Ehi) of
//     the analytic window extents in NCH 64-wide chunks (X passes apply the
//     overwrite to it); every access is a coalesced row segment;
//  2. prefix sums Q[k] = sum_{sa <= i <= sa+k} (v(i) - vref), vref = v(sa):
//     each lane sums an odd-length run of consecutive elements serially
//     (odd strides keep the LDS reads conflict-free), one DPP wavefront scan
//     of the run totals gives the offsets; subtracting vref keeps the fp32
//     sums small for offset-dominated data;
//  3. B = (Q(u) - Q(l)) / (u - l) + vref (hardware reciprocal), the update.
// The wavefronts are persistent: each walks every TW-th segment and issues
// the next segment's loads before processing the current one (registers
// double-buffered), so memory traffic overlaps the scan and update work.
constexpr int kYStride = 65; // padded LDS row stride

// Window extents of the vertical tile with columns [64 tile_x, 64 tile_x +
// 64) and rows [y0, y0 + rows): the maximum over its 32-row bands and its
// two 32-column groups, clamped to the analytic caps (at least one row
// below, as the staged segment starts at y0 - elo + 1).
__device__ __forceinline__ void tile_extents(const PassArgs& a, int tile_x, int y0, int rows,
                                             int* elo, int* ehi) {
  int lo = 1, hi = 0;
  const int nb = (a.H + 31) / 32;
  for (int b = y0 / 32; b < min(nb, (y0 + rows + 31) / 32); b++) {
#pragma unroll
    for (int g = 0; g < 2; g++) {
      const int gi = 2 * tile_x + g;
      if (gi < a.ext_groups) {
        const int e = b * a.ext_groups + gi;
        lo = max(lo, a.yext[2 * e]);
        hi = max(hi, a.yext[2 * e + 1]);
      }
    }
  }
  *elo = min(lo, a.elo_cap);
  *ehi = min(hi, a.ehi_cap);
}

// Output rows y0 + q + 4k of a vertical tile from the column prefix sums Q
// (rows [sa, sa + segn) of the segment, stride kYStride): every LDS read of
// the tile is issued before the first use.  Indices are clamped into the
// segment; results of masked rows are not stored.
template <int MODE, int PR, bool NB>
__device__ __forceinline__ void ytile_out(const PassArgs& a, const float* Q, int sa, int segn,
                                          int y0, int q, int lane, int x, bool xok, float vref,
                                          const uint32_t (&b)[PR], const float (&v)[PR]) {
  float ql[PR], qu[PR];
  int nw[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + 4 * k;
    const int l = y - bounds_dl<NB>(b[k]);
    const int u = y + bounds_du<NB>(b[k]);
    const int il = min(l - sa, segn - 1), iu = min(u - sa, segn - 1);
    const float rl = Q[max(il, 0) * kYStride + lane];
    const float ru = Q[max(iu, 0) * kYStride + lane];
    ql[k] = il >= 0 ? rl : 0.f;
    qu[k] = iu >= 0 ? ru : 0.f;
    nw[k] = u - l;
  }
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + 4 * k;
    if (!xok || y >= a.H) continue;
    const float B = (qu[k] - ql[k]) * __builtin_amdgcn_rcpf((float)nw[k]) + vref;
    float res;
    if (MODE == PM_FILTER)
      res = B;
    else
      res = v[k] + a.step * (-(a.lambda * (v[k] - B)));
    st32(a.out, (unsigned)y * (unsigned)a.W + (unsigned)x, res);
  }
}

// Vertical pass (box filter, Optimize Y half).  A workgroup owns 64 columns x 32 rows;
// lanes walk adjacent columns, so every global access is a contiguous row
// segment.  The tile's window extents come from the bounds pass (per tile),
// so only the rows [y0 - Elo + 1, y0 + 32 + Ehi) are loaded into LDS.
// Column prefix sums of v - vref (vref = the column's first segment value)
// are formed in place, one thread per column and quarter of the rows.
template <int MODE, bool NB, int YR = kYRows>
__device__ __forceinline__ void ytile(const PassArgs& a, int tile_x, int tile_y, float* lds) {
  static_assert(YR % kYRows == 0, "whole extent tiles");
  float* Q = lds;  // [seg rows][65]
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int x0 = tile_x * kYCols, y0 = tile_y * YR;
  const int W = a.W, H = a.H;
  const int x = x0 + lane;
  const bool xok = x < W;
  constexpr int PR = YR / 4;
  int elo, ehi;
  tile_extents(a, tile_x, y0, YR, &elo, &ehi);
  const int sa = max(0, y0 - elo + 1);
  const int sb = min(H, y0 + YR + ehi);
  const int segn = sb - sa;

  uint32_t b[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + 4 * k;
    b[k] = ld_bounds<NB>(a.bnd, (unsigned)min(y, H - 1) * (unsigned)W + (unsigned)min(x, W - 1));
  }
  for (int rr = q; rr < segn; rr += 4 * 12) {
    float vals[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const int row = rr + 4 * k;
      vals[k] = ld32(a.vin, (unsigned)(sa + min(row, segn - 1)) * (unsigned)W +
                                (unsigned)min(x, W - 1));
    }
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const int row = rr + 4 * k;
      if (row < segn) Q[row * kYStride + lane] = vals[k];
    }
  }
  __syncthreads();
  const float vref = Q[lane];
  float v[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + 4 * k;
    v[k] = y < H ? Q[(y - sa) * kYStride + lane] : 0.f;
  }
  __syncthreads();
  // Column prefix sums: wavefront q scans rows [q R, (q+1) R) of every
  // column serially (lanes = columns), then adds the totals of the quarters
  // above it.
  float* tot = Q + a.seg_stride * kYStride;  // [4][64]
  const int R = (segn + 3) / 4;
  const int r0 = q * R, r1 = min(segn, r0 + R);
  float run = 0.f;
  for (int row = r0; row < r1; row++) {
    run += Q[row * kYStride + lane] - vref;
    Q[row * kYStride + lane] = run;
  }
  tot[q * kYCols + lane] = run;
  __syncthreads();
  float off = 0.f;
  for (int k = 0; k < q; k++) off += tot[k * kYCols + lane];
  if (q > 0)
    for (int row = r0; row < r1; row++) Q[row * kYStride + lane] += off;
  __syncthreads();

  ytile_out<MODE, PR, NB>(a, Q, sa, segn, y0, q, lane, x, xok, vref, b, v);
}

// Vertical pass with the staged rows in registers.  Wavefront q owns the
// contiguous rows [q R, q R + R) of the tile's segment (R = ceil(segn / 4) <=
// RMAX): each of its loads is one 64-column row, all issued before the first
// use.  The column prefix sums of v - vref run serially in registers, the
// four quarter totals meet in LDS (one barrier), and each wavefront writes
// its prefix sums with the offset of the quarters above (second barrier);
// the window sums are then read from LDS as in ytile.
// YR: output rows per tile (32 or 64); the window extents come from the
// 32-row tiles of the bounds pass that it covers.  The loads and the rest
// are split (ytile_reg_load, ytile_reg_compute) so that a persistent
// workgroup can issue the next tile's loads before processing the current.

// Registers of one vertical register tile between its loads and their use.
template <int RMAX, int PR>
struct YTileRegs {
  float vals[RMAX];  // rows [sa + r0, sa + r0 + RMAX) of the wavefront's quarter
  uint32_t b[PR];    // packed bounds of the output rows y0 + q + 4k
  float v[PR];       // v' at the output rows
  float vref;        // v' at row sa
  int sa, segn;      // staged segment rows [sa, sa + segn)
};

// Issues every global load of the register tile (tile_x, tile_y); nothing
// waits for them here.
template <int RMAX, int YR, bool NB>
__device__ __forceinline__ void ytile_reg_load(const PassArgs& a, int tile_x, int tile_y,
                                               YTileRegs<RMAX, YR / 4>& T) {
  static_assert(YR % kYRows == 0, "whole extent tiles");
  constexpr int PR = YR / 4;
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x0 = tile_x * kYCols, y0 = tile_y * YR;
  const int W = a.W, H = a.H;
  const unsigned xc = (unsigned)min(x0 + lane, W - 1);
  int elo, ehi;
  tile_extents(a, tile_x, y0, YR, &elo, &ehi);
  const int sa = max(0, y0 - elo + 1);
  const int sb = min(H, y0 + YR + ehi);
  const int segn = sb - sa;
  const int R = (segn + 3) >> 2;
  const int r0 = q * R;
  // Only the wavefront's own rows are loaded (rlim is wave-uniform): a tile
  // whose window extents, from the bounds pass, are below its register
  // class loads fewer rows (64-frame depth batch: 4.99 -> 4.81 ms).
  const int rlim = min(R, segn - r0);
#pragma unroll
  for (int k = 0; k < RMAX; k++)
    if (k < rlim) T.vals[k] = ld32(a.vin, (unsigned)(sa + r0 + k) * (unsigned)W + xc);
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const unsigned gi = (unsigned)min(y0 + q + 4 * k, H - 1) * (unsigned)W + xc;
    T.b[k] = ld_bounds<NB>(a.bnd, gi);
    T.v[k] = ld32(a.vin, gi);
  }
  T.vref = ld32(a.vin, (unsigned)sa * (unsigned)W + xc);
  T.sa = sa;
  T.segn = segn;
}

// Scan, LDS exchange and output of a loaded register tile.  Two barriers;
// a following tile of the same workgroup may reuse the LDS at once (its
// first barrier orders its writes after these reads).
template <int MODE, int RMAX, int YR, bool NB>
__device__ __forceinline__ void ytile_reg_compute(const PassArgs& a, int tile_x, int tile_y,
                                                  YTileRegs<RMAX, YR / 4>& T, float* lds) {
  constexpr int PR = YR / 4;
  float* Q = lds;                              // [segn][kYStride]
  float* tot = lds + a.seg_stride * kYStride;  // [4][64]
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x0 = tile_x * kYCols, y0 = tile_y * YR;
  const int x = x0 + lane;
  const bool xok = x < a.W;
  const int sa = T.sa, segn = T.segn;
  const int R = (segn + 3) >> 2;
  const int r0 = q * R;
  float run = 0.f;
  const int rlim = min(R, segn - r0);  // rows of this wavefront inside the segment
#pragma unroll
  for (int k = 0; k < RMAX; k++) {
    run += k < rlim ? T.vals[k] - T.vref : 0.f;
    T.vals[k] = run;
  }
  tot[q * kYCols + lane] = run;
  __syncthreads();
  float off = 0.f;
  for (int k = 0; k < q; k++) off += tot[k * kYCols + lane];
#pragma unroll
  for (int k = 0; k < RMAX; k++) {
    if (k < rlim) Q[(r0 + k) * kYStride + lane] = T.vals[k] + off;
  }
  __syncthreads();
  ytile_out<MODE, PR, NB>(a, Q, sa, segn, y0, q, lane, x, xok, T.vref, T.b, T.v);
}

template <int MODE, int RMAX, bool NB, int YR = kYRows>
__device__ __forceinline__ void ytile_reg(const PassArgs& a, int tile_x, int tile_y, float* lds) {
  YTileRegs<RMAX, YR / 4> T;
  ytile_reg_load<RMAX, YR, NB>(a, tile_x, tile_y, T);
  ytile_reg_compute<MODE, RMAX, YR, NB>(a, tile_x, tile_y, T, lds);
}

// Register budget of ytile_reg: rows per wavefront for a segment of up to
// seg_rows rows; 0 when the LDS-loop tile must be used.
__host__ __device__ inline int ytile_rmax(int seg_rows) {
  const int r = (seg_rows + 3) / 4;
  return r <= 16 ? 16 : r <= 24 ? 24 : r <= 32 ? 32 : r <= 48 ? 48 : 0;
}

}  // namespace dtsamd
