// One Optimize iteration without the target overwrite as one launch: the
// horizontal half-step of the rows a vertical tile needs, recomputed per tile
// into LDS, then the vertical half-step of the tile.  Included by dts.hip.
// src/domain_transform_solver.cu:100-202 (the X then Y launches of one
// iteration).
#pragma once

#include "dts_kernels.h"

namespace dtsamd {

// Registers of one row of the horizontal half inside a fused tile: NCX
// 64-wide chunks of v over [xa, xa + segx), and the lane's own bounds,
// target and confidence at column x0 + lane.  OW (the target overwrite):
// also the confidence and target of the chunks.
template <int NCX, bool OW = false>
struct FRow {
  float v[NCX];
  float sc[OW ? NCX : 1], st[OW ? NCX : 1];
  uint32_t b;
  float t, c;
};

template <int NCX, bool NBH, bool OW = false>
__device__ __forceinline__ void frow_load(const PassArgs& ax, int row, int x0, int xa, int segx,
                                          int lane, FRow<NCX, OW>& R) {
  const unsigned rowoff = (unsigned)row * (unsigned)ax.W;
#pragma unroll
  for (int k = 0; k < NCX; k++) {
    const unsigned gi = rowoff + (unsigned)(xa + min(64 * k + lane, segx - 1));
    R.v[k] = ld32(ax.vin, gi);
    if (OW) {
      R.sc[k] = ld32(ax.c, gi);
      R.st[k] = ld32(ax.t, gi);
    }
  }
  const unsigned gi = rowoff + (unsigned)min(x0 + lane, ax.W - 1);
  R.b = ld_bounds<NBH>(ax.bnd, gi);
  R.t = ld32(ax.t, gi);
  R.c = ld32(ax.c, gi);
}

// Horizontal update of pixel (row, x0 + lane) from a loaded row: prefix sums
// of v - vref over the row's window span in the wavefront's LDS slices V, Q
// (64 * (NCX | 1) floats each), as in xseg_process.  OW: pixels whose
// confidence exceeds theta enter the sums with their target, and their
// output is the target.
template <int MODE, int NCX, bool NBH, bool STEREO, bool OW = false>
__device__ __forceinline__ float frow_process(const PassArgs& ax, int row, int x0, int xa,
                                              int segx, int lane, const FRow<NCX, OW>& R,
                                              float* V, float* Q) {
  constexpr int EPL = NCX | 1;
  float vals[NCX];
#pragma unroll
  for (int k = 0; k < NCX; k++) vals[k] = OW && R.sc[k] > ax.theta ? R.st[k] : R.v[k];
  const float vref = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vals[0])));
#pragma unroll
  for (int k = 0; k < NCX; k++) V[64 * k + lane] = vals[k];
  wave_lds_sync();
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
  const int p = x0 + lane;
  const int l = p - bounds_dl<NBH>(R.b), u = p + bounds_du<NBH>(R.b);
  const int il = l - xa, iu = max(0, min(u - xa, segx - 1));
  const float rl = Q[max(0, min(il, segx - 1))];
  const float qu = Q[iu];
  const float vv = V[max(0, min(p - xa, segx - 1))];
  const float ql = il >= 0 ? rl : 0.f;
  const float rn = __builtin_amdgcn_rcpf((float)(u - l));
  const float Bm = (qu - ql) * rn + vref;
  // Lanes past the row end compute from clamped loads and are not stored;
  // the clamped position keeps the stereo term's gathers inside the row.
  const float res = x_update<MODE, OW, STEREO>(ax, (unsigned)row * (unsigned)ax.W,
                                              min(p, ax.W - 1), vv, Bm, rn, R.t, R.c, true);
  wave_lds_sync();
  return res;
}

// LDS bytes of a fused tile: the staged rows (stride kYStride), the NW
// wavefronts' column totals and their two horizontal scan slices.
__host__ __device__ constexpr size_t fused_tile_lds(int rows, int ncx, int nw) {
  return sizeof(float) * ((size_t)rows * kYStride + (size_t)nw * kYCols +
                          (size_t)nw * 2 * 64 * (ncx | 1));
}

// One fused tile: 64 columns x YR (16, 32 or 64) output rows, NW wavefronts.
// Its vertical windows need the horizontal output of rows [sa, sa + segn)
// (per-tile extents from the bounds pass); wavefront q computes rows sa + q,
// sa + q + NW, ... into the LDS tile T, the next row's loads in flight while
// the current one is processed.  The vertical half then runs on T as in
// ytile.  NBH, NBV: 2-byte horizontal, vertical bounds elements.  STEREO: the
// photometric term in the horizontal half (x_update gathers the right view).
// OW: the target overwrite in the horizontal half.
template <int MODE, int NCX, bool NBH, bool NBV, int YR = kYRows, bool STEREO = false,
          int NW = 4, bool OW = false>
__device__ __forceinline__ void xytile(const PassArgs& ax, const PassArgs& ay, int tile_x,
                                       int tile_y, float* lds) {
  static_assert(YR == 16 || YR == 32 || YR == 64, "16, 32 or 64 output rows");
  static_assert(NW == 4 || NW == 8 || NW == 16, "4, 8 or 16 wavefronts per tile");
  static_assert(YR % NW == 0, "whole output rows per wavefront");
  constexpr int PR = YR / NW;
  float* T = lds;                                      // [ay.seg_stride][kYStride]
  float* tot = T + ay.seg_stride * kYStride;           // [NW][64]
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* V = tot + NW * kYCols + q * 2 * 64 * (NCX | 1);
  float* Q = V + 64 * (NCX | 1);
  const int W = ay.W, H = ay.H;
  const int x0 = tile_x * kYCols, y0 = tile_y * YR;
  const int x = x0 + lane;
  const bool xok = x < W;
  int elo, ehi;
  tile_extents(ay, tile_x, y0, YR, &elo, &ehi);
  const int sa = max(0, y0 - elo + 1);
  const int segn = min(H, y0 + YR + ehi) - sa;
  const int xa = max(0, x0 - ax.elo_cap + 1);
  const int segx = min(W, x0 + kYCols + ax.ehi_cap) - xa;
  // Vertical bounds of the output rows, loaded before the horizontal work.
  uint32_t b[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const unsigned gi = (unsigned)min(y0 + q + NW * k, H - 1) * (unsigned)W + (unsigned)min(x, W - 1);
    b[k] = ld_bounds<NBV>(ay.bnd, gi);
  }
  // Horizontal half for the staged rows.
  {
    int r = q;
    FRow<NCX, OW> A, B;
    if (r < segn) frow_load<NCX, NBH, OW>(ax, sa + r, x0, xa, segx, lane, A);
    while (r < segn) {
      const int r1 = r + NW;
      if (r1 < segn) frow_load<NCX, NBH, OW>(ax, sa + r1, x0, xa, segx, lane, B);
      T[r * kYStride + lane] =
          frow_process<MODE, NCX, NBH, STEREO, OW>(ax, sa + r, x0, xa, segx, lane, A, V, Q);
      if (r1 >= segn) break;
      const int r2 = r1 + NW;
      if (r2 < segn) frow_load<NCX, NBH, OW>(ax, sa + r2, x0, xa, segx, lane, A);
      T[r1 * kYStride + lane] =
          frow_process<MODE, NCX, NBH, STEREO, OW>(ax, sa + r1, x0, xa, segx, lane, B, V, Q);
      r = r2;
    }
  }
  __syncthreads();
  // Vertical half on T (as ytile).
  const float vref = T[lane];
  float v[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + NW * k;
    v[k] = y < H ? T[(y - sa) * kYStride + lane] : 0.f;
  }
  __syncthreads();
  const int R = (segn + NW - 1) / NW;
  const int r0 = q * R, r1 = min(segn, r0 + R);
  float run = 0.f;
  for (int row = r0; row < r1; row++) {
    run += T[row * kYStride + lane] - vref;
    T[row * kYStride + lane] = run;
  }
  tot[q * kYCols + lane] = run;
  __syncthreads();
  float off = 0.f;
  for (int k = 0; k < q; k++) off += tot[k * kYCols + lane];
  if (q > 0)
    for (int row = r0; row < r1; row++) T[row * kYStride + lane] += off;
  __syncthreads();
  float ql[PR], qu[PR];
  int nw[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + NW * k;
    const int l = y - bounds_dl<NBV>(b[k]);
    const int u = y + bounds_du<NBV>(b[k]);
    const int il = min(l - sa, segn - 1), iu = min(u - sa, segn - 1);
    const float rl = T[max(il, 0) * kYStride + lane];
    const float ru = T[max(iu, 0) * kYStride + lane];
    ql[k] = il >= 0 ? rl : 0.f;
    qu[k] = iu >= 0 ? ru : 0.f;
    nw[k] = u - l;
  }
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + NW * k;
    if (!xok || y >= H) continue;
    const float Bv = (qu[k] - ql[k]) * __builtin_amdgcn_rcpf((float)nw[k]) + vref;
    const float res = v[k] + ay.step * (-(ay.lambda * (v[k] - Bv)));
    st32(ay.out, (unsigned)y * (unsigned)W + (unsigned)x, res);
  }
}

}  // namespace dtsamd
