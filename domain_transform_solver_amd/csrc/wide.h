// Inclusive line prefix sums in double precision for windows too wide for
// the LDS segments of the pass kernels (used by the forward passes in dts.hip
// and the adjoint passes in adjoint.hip).  Elem is a device functor giving
// the value of element i (row-major index) of a W x H plane; S receives the
// prefixes along rows or along columns, tot holds block totals.  Prefixes are
// built from 64-element blocks: block totals first, then every block adds
// the totals before it, so no launch walks a whole line serially.
#pragma once

#include <algorithm>

#include "common.h"

namespace dtsamd {

// Block totals along rows: grid (ceil(W/64), H) of one wavefront each;
// tot[y * gridDim.x + bx].
template <class Elem>
__global__ __launch_bounds__(64) void k_line_tot_rows(Elem e, int W, double* __restrict__ tot) {
  const int nbx = (int)gridDim.x, bx = (int)blockIdx.x, y = (int)blockIdx.y;
  const int lane = threadIdx.x, x = bx * 64 + lane;
  double v = x < W ? (double)e((unsigned)y * (unsigned)W + (unsigned)x) : 0.0;
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (lane == 0) tot[(size_t)y * nbx + bx] = v;
}

template <class Elem>
__global__ __launch_bounds__(64) void k_line_prefix_rows(Elem e, int W,
                                                         const double* __restrict__ tot,
                                                         double* __restrict__ S) {
  const int nbx = (int)gridDim.x, bx = (int)blockIdx.x, y = (int)blockIdx.y;
  const int lane = threadIdx.x, x = bx * 64 + lane;
  double off = 0.0;
  for (int k = lane; k < bx; k += 64) off += tot[(size_t)y * nbx + k];
  for (int d = 32; d >= 1; d >>= 1) off += __shfl_xor(off, d, 64);
  double v = x < W ? (double)e((unsigned)y * (unsigned)W + (unsigned)x) : 0.0;
  for (int d = 1; d < 64; d <<= 1) {
    const double t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  if (x < W) S[(size_t)y * W + x] = off + v;
}

// Block totals along columns: grid (ceil(W/256), ceil(H/64)), a thread per
// column; tot[by * W + x].
template <class Elem>
__global__ __launch_bounds__(256) void k_line_tot_cols(Elem e, int W, int H,
                                                       double* __restrict__ tot) {
  const int x = (int)blockIdx.x * 256 + threadIdx.x, by = (int)blockIdx.y;
  if (x >= W) return;
  const int y0 = by * 64, y1 = min(H, y0 + 64);
  double s = 0.0;
  for (int y = y0; y < y1; y++) s += (double)e((unsigned)y * (unsigned)W + (unsigned)x);
  tot[(size_t)by * W + x] = s;
}

template <class Elem>
__global__ __launch_bounds__(256) void k_line_prefix_cols(Elem e, int W, int H,
                                                          const double* __restrict__ tot,
                                                          double* __restrict__ S) {
  const int x = (int)blockIdx.x * 256 + threadIdx.x, by = (int)blockIdx.y;
  if (x >= W) return;
  double s = 0.0;
  for (int k = 0; k < by; k++) s += tot[(size_t)k * W + x];
  const int y0 = by * 64, y1 = min(H, y0 + 64);
  for (int y = y0; y < y1; y++) {
    s += (double)e((unsigned)y * (unsigned)W + (unsigned)x);
    S[(size_t)y * W + x] = s;
  }
}

// Prefixes of a W x H plane along rows (rows = true) or columns into S;
// returns the launch status.
template <class Elem>
hipError_t line_prefix(hipStream_t st, Elem e, int W, int H, bool rows, double* tot, double* S) {
  if (rows) {
    const dim3 g(div_up(W, 64), H);
    k_line_tot_rows<Elem><<<g, 64, 0, st>>>(e, W, tot);
    k_line_prefix_rows<Elem><<<g, 64, 0, st>>>(e, W, tot, S);
  } else {
    const dim3 g(div_up(W, 256), div_up(H, 64));
    k_line_tot_cols<Elem><<<g, 256, 0, st>>>(e, W, H, tot);
    k_line_prefix_cols<Elem><<<g, 256, 0, st>>>(e, W, H, tot, S);
  }
  return hipGetLastError();
}

// Block-total doubles line_prefix needs for a W x H plane (either direction).
inline size_t line_prefix_tot(int W, int H) {
  return std::max((size_t)H * div_up(W, 64), (size_t)div_up(H, 64) * W);
}

}  // namespace dtsamd
