// Shared host/device helpers for the gfx950 DTS / HFBS library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "dts_capi.h"

namespace dtsamd {

// Thread-local last-error text behind dts_last_error().
void set_error(const char* fmt, ...);
const char* last_error();

#define DTS_HIP_CHECK(call)                                                  \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::dtsamd::set_error("%s failed: %s (%s:%d)", #call,                    \
                          hipGetErrorString(e_), __FILE__, __LINE__);        \
      return DTS_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define DTS_LAUNCH_CHECK() DTS_HIP_CHECK(hipGetLastError())

// A named range on the calling host thread around one C ABI call (roctx):
// rocprofv3 --marker-trace shows it beside the kernels the call enqueues.
// Without a tool attached the calls return at once.
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

// Blocks -> XCD-aware order.  The dispatcher hands workgroup g to XCD g % 8;
// remapping g to start_of(xcd) + g / 8 gives each XCD a contiguous run of
// tiles, so tiles that share halo lines and output cache lines meet in the
// same L2.  The map is a bijection of [0, n).
__device__ inline int xcd_remap(int g, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, r = n % kXcd;
  const int x = g % kXcd, k = g / kXcd;
  return x * q + (x < r ? x : r) + k;
}

constexpr int kWave = 64;

// Inclusive prefix sum over one 64-lane wavefront with DPP moves (gfx9
// family): Hillis-Steele inside each 16-lane row (row_shr 1, 2, 4, 8), then
// row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.  A lane
// whose DPP source is out of range or masked off adds the `old` operand, 0.
__device__ inline float wave_incl_scan_dpp(float v) {
  float s = v;
#define DTS_DPP_ADD(ctrl, rmask) \
  s += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s), ctrl, rmask, 0xf, false))
  DTS_DPP_ADD(0x111, 0xf);  // row_shr:1
  DTS_DPP_ADD(0x112, 0xf);  // row_shr:2
  DTS_DPP_ADD(0x114, 0xf);  // row_shr:4
  DTS_DPP_ADD(0x118, 0xf);  // row_shr:8
  DTS_DPP_ADD(0x142, 0xa);  // row_bcast:15
  DTS_DPP_ADD(0x143, 0xc);  // row_bcast:31
#undef DTS_DPP_ADD
  return s;
}

// Inclusive 64-bit integer prefix sum with DPP moves of the two 32-bit
// halves (same lane pattern as wave_incl_scan_dpp): over the whole
// wavefront, or (HALVES) over lanes 0-31 and 32-63 separately, whose last
// step is then the row_bcast:15 into rows 1 and 3 only.
template <bool HALVES>
__device__ inline int64_t incl_scan_i64_dpp(int64_t v) {
#define DTS_DPP_ADD64(ctrl, rmask)                                                          \
  do {                                                                                     \
    const int lo_ = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, ctrl, rmask, 0xf, false); \
    const int hi_ = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), ctrl, rmask, 0xf, false);   \
    v += (int64_t)(((uint64_t)(uint32_t)hi_ << 32) | (uint32_t)lo_);                         \
  } while (0)
  DTS_DPP_ADD64(0x111, 0xf);
  DTS_DPP_ADD64(0x112, 0xf);
  DTS_DPP_ADD64(0x114, 0xf);
  DTS_DPP_ADD64(0x118, 0xf);
  DTS_DPP_ADD64(0x142, 0xa);
  if (!HALVES) DTS_DPP_ADD64(0x143, 0xc);
#undef DTS_DPP_ADD64
  return v;
}

__device__ inline int64_t wave_incl_scan_i64_dpp(int64_t v) { return incl_scan_i64_dpp<false>(v); }

// Inclusive scan of first-order linear recurrences over one wavefront: lane
// k holds the map r -> a r + b; on return it holds the composition of the
// maps of lanes 0..k applied in lane order, so r_k = a * r_(-1) + b.  Same
// DPP lane pattern as wave_incl_scan_dpp; a masked or out-of-range source
// contributes the identity map (a = 1, b = 0).
__device__ inline void wave_linrec_scan_dpp(float& a, float& b) {
#define DTS_DPP_LINREC(ctrl, rmask)                                                          \
  do {                                                                                      \
    const float pa_ = __int_as_float(                                                       \
        __builtin_amdgcn_update_dpp(__float_as_int(1.0f), __float_as_int(a), ctrl, rmask, 0xf, \
                                    false));                                                \
    const float pb_ = __int_as_float(                                                       \
        __builtin_amdgcn_update_dpp(0, __float_as_int(b), ctrl, rmask, 0xf, false));        \
    b = a * pb_ + b;                                                                        \
    a = a * pa_;                                                                            \
  } while (0)
  DTS_DPP_LINREC(0x111, 0xf);
  DTS_DPP_LINREC(0x112, 0xf);
  DTS_DPP_LINREC(0x114, 0xf);
  DTS_DPP_LINREC(0x118, 0xf);
  DTS_DPP_LINREC(0x142, 0xa);
  DTS_DPP_LINREC(0x143, 0xc);
#undef DTS_DPP_LINREC
}

// Two recurrences with the same multipliers (r -> a r + b1, s -> a s + b2)
// scanned together: the prefix products of a are formed once.
__device__ inline void wave_linrec_scan2_dpp(float& a, float& b1, float& b2) {
#define DTS_DPP_LINREC2(ctrl, rmask)                                                         \
  do {                                                                                      \
    const float pa_ = __int_as_float(                                                       \
        __builtin_amdgcn_update_dpp(__float_as_int(1.0f), __float_as_int(a), ctrl, rmask, 0xf, \
                                    false));                                                \
    const float p1_ = __int_as_float(                                                       \
        __builtin_amdgcn_update_dpp(0, __float_as_int(b1), ctrl, rmask, 0xf, false));       \
    const float p2_ = __int_as_float(                                                       \
        __builtin_amdgcn_update_dpp(0, __float_as_int(b2), ctrl, rmask, 0xf, false));       \
    b1 = a * p1_ + b1;                                                                      \
    b2 = a * p2_ + b2;                                                                      \
    a = a * pa_;                                                                            \
  } while (0)
  DTS_DPP_LINREC2(0x111, 0xf);
  DTS_DPP_LINREC2(0x112, 0xf);
  DTS_DPP_LINREC2(0x114, 0xf);
  DTS_DPP_LINREC2(0x118, 0xf);
  DTS_DPP_LINREC2(0x142, 0xa);
  DTS_DPP_LINREC2(0x143, 0xc);
#undef DTS_DPP_LINREC2
}

inline int div_up(int a, int b) { return (a + b - 1) / b; }

// LDS per workgroup on gfx950 (hipDeviceProp_t::sharedMemPerBlock = 160 KiB).
constexpr size_t kLdsMax = 160 * 1024;
// Largest bounds tile stage kept in LDS (more would cut occupancy to one
// workgroup per CU); wider windows search ct in global memory.
constexpr size_t kBoundsLdsMax = 96 * 1024;

// 4-byte element access through a 32-bit byte offset from a base pointer,
// which the backend maps onto the SGPR-base + 32-bit VGPR-offset form of the
// global memory instructions.  Element indices stay below 2^30 (the largest
// image is 32767 x 32767).
template <typename T>
__device__ inline T ld32(const T* base, unsigned idx) {
  static_assert(sizeof(T) == 4, "4-byte elements");
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (idx << 2));
}

template <typename T>
__device__ inline void st32(T* base, unsigned idx, T v) {
  static_assert(sizeof(T) == 4, "4-byte elements");
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (idx << 2)) = v;
}

}  // namespace dtsamd
