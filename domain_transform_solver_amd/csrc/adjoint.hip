// Adjoint passes of the DTS solve (no reference counterpart: the reference
// has no gradients).  The forward passes (src/domain_transform_solver.cuh:
// 37-116) are affine in v except the data term, so one iteration's adjoint is
//   Y:  vbar' = (1 - s lam) g + s lam B_V^T(g)
//   X:  with xbar = vbar' where c <= theta (else 0, and tbar += vbar'),
//       wbar  = (1 - s lam) xbar + s lam B_H^T(xbar) - s/2 dg/dw xbar,
//       tbar += -s/2 dg/dt xbar,  cbar += -s/2 dg/dc xbar,
//       then the pre-pass overwrite routes wbar to tbar where c > theta.
// B is the box mean over (l(p), u(p)] of n(p) = u - l samples, so
// B^T(y)(j) = sum_p [l(p) < j <= u(p)] y(p) / n(p).  Away from the collapse
// the window relation is symmetric (j in W(p) iff |ct(j) - ct(p)| <= sigma
// iff p in W(j)), which turns B^T into a windowed sum of z = y / n over the
// symmetric window of j, read from the same prefix sums the forward uses.
// Lines whose lower bound collapses (positions below a) keep only [p, u(p)],
// those whose upper bound collapses (above b) only (l(p), p]; the two sums
// over the removed parts are subtracted.  With S the prefix sum of z:
//   l* = -1 if j < a else l(j),  u* = L-1 if j > b else u(j),
//   B^T(j) = S(u*) - S(l*) - [S(max(j, min(u*, a-1))) - S(j)]
//                          - [S(j-1) - S(min(j-1, max(l*, b)))].
// Guide derivatives are zero almost everywhere (the guide only moves the
// integer bounds) and are not formed.
#include "adjoint.h"
#include "dts_kernels.h"
#include "wide.h"

namespace dtsamd {

// Wavefronts per vertical adjoint tile: 8 for the 16-row tiles of small
// frames (each wavefront's chain of staged rows halves), else 4.
constexpr int adj_y_waves(int yr) { return yr == 16 ? 8 : 4; }

__device__ __forceinline__ int2 collapse_limits(const int64_t* __restrict__ c, int L,
                                                int64_t sfx, int E) {
  const int64_t c0 = c[0], cl = c[L - 1];
  // a: first x with c[x] - c0 > sfx (false at 0; true at E + 1 < L when
  // sigma is finite, since every step adds at least 1).
  int a = L;
  const int hi = min(L - 1, E + 1);
  if (c[hi] - c0 > sfx) {
    int lo = 0, h2 = hi;
    while (h2 - lo > 1) {
      const int mid = (lo + h2) >> 1;
      if (c[mid] - c0 > sfx) h2 = mid; else lo = mid;
    }
    a = h2;
  }
  // b: last x with cl - c[x] > sfx (false at L - 1).
  int b = -1;
  const int lo0 = max(0, L - 2 - E);
  if (cl - c[lo0] > sfx) {
    int lo = lo0, h2 = L - 1;
    while (h2 - lo > 1) {
      const int mid = (lo + h2) >> 1;
      if (cl - c[mid] > sfx) lo = mid; else h2 = mid;
    }
    b = lo;
  }
  return make_int2(a, b);
}

__global__ __launch_bounds__(256) void k_adj_collapse(const int64_t* __restrict__ ct, int L,
                                                      int nlines, int64_t sfx, int E,
                                                      int2* __restrict__ ab) {
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  ab[line] = collapse_limits(ct + (size_t)line * L, L, sfx, E);
}

// Threads [0, H) take the rows, [H, H + W) the columns.
__global__ __launch_bounds__(256) void k_adj_collapse2(const int64_t* __restrict__ ctH,
                                                       const int64_t* __restrict__ ctV, int W,
                                                       int H, int64_t sfx, int Eh, int Ev,
                                                       int2* __restrict__ abH,
                                                       int2* __restrict__ abV) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < H)
    abH[i] = collapse_limits(ctH + (size_t)i * W, W, sfx, Eh);
  else if (i < H + W)
    abV[i - H] = collapse_limits(ctV + (size_t)(i - H) * H, H, sfx, Ev);
}

// B^T of one position j from the prefix sums S over the staged range
// [sa, sb) (S(sa - 1) = 0), split into the six prefix-sum positions it reads
// (every one in [sa - 1, sb - 1]) and their combination, so that a caller
// can issue the LDS reads of several positions before the first use.
// Indices outside the staged range only occur for a collapsed window of a
// non-finite sigma, where the clamped terms cancel to z(j).
__device__ __forceinline__ void box_adjoint_idx(int j, int l, int u, int2 ab, int L, int sa,
                                                int sb, int (&k)[6]) {
  int ls = j < ab.x ? -1 : l;
  int us = j > ab.y ? L - 1 : u;
  ls = max(ls, sa - 1);
  us = min(us, sb - 1);
  k[0] = us;
  k[1] = ls;
  k[2] = max(j, min(us, ab.x - 1));
  k[3] = j;
  k[4] = j - 1;
  k[5] = max(min(j - 1, max(ls, ab.y)), sa - 1);
}

__device__ __forceinline__ float box_adjoint_eval(const float (&v)[6]) {
  return (v[0] - v[1]) - (v[2] - v[3]) - (v[4] - v[5]);
}

// The per-pixel end of the horizontal adjoint at element gi: from the window
// size n, B^T of the cotangent (BT), and the pixel's cotangent, confidence,
// target, pass state and accumulated d/dt, d/dc, the cotangent of the pass
// input and the new d/dt, d/dc.
struct AdjXOut {
  float vo, tb, cb;  // cotangent of the pass input, new d/dt, d/dc
};

__device__ __forceinline__ AdjXOut adj_x_values(const AdjArgs& a, int n, float BT, float go,
                                                float co, float to, float so, float tbo,
                                                float cbo) {
  const float sl = a.step * a.lambda;
  const bool over = a.ow && co > a.theta;
  const float xb = over ? 0.f : go;
  const float w = over ? to : so;
  const float r = w - to;
  const float rn = __builtin_amdgcn_rcpf((float)n);
  const float cn = co * rn;
  float dv, dc;
  if (a.loss == DTS_LOSS_L2) {
    dv = 2.f * cn;
    dc = 2.f * r * rn;
  } else {
    const float is = __builtin_amdgcn_rsqf(r * r + 1e-6f);
    dv = cn * 1e-6f * is * is * is;
    dc = r * rn * is;
  }
  const float wbar = (1.f - sl) * xb + sl * BT - 0.5f * a.step * dv * xb;
  AdjXOut o;
  o.tb = tbo + 0.5f * a.step * dv * xb + (over ? go : 0.f);
  o.cb = cbo - 0.5f * a.step * dc * xb;
  o.vo = wbar;
  if (over) {
    o.tb += wbar;
    o.vo = 0.f;
  }
  return o;
}

__device__ __forceinline__ void adj_x_finish(const AdjArgs& a, unsigned gi, int n, float BT,
                                             float go, float co, float to, float so, float tbo,
                                             float cbo) {
  const AdjXOut o = adj_x_values(a, n, BT, go, co, to, so, tbo, cbo);
  if (a.acc_last) {
    st32(a.tbar, gi, o.tb + o.vo);
  } else {
    st32(a.vout, gi, o.vo);
    st32(a.tbar, gi, o.tb);
  }
  st32(a.cbar, gi, o.cb);
}

template <int NCH, bool NB>
__global__ __launch_bounds__(256) void k_adj_x(AdjArgs a, int ntx) {
  extern __shared__ float lds[];
  constexpr int EPL = NCH | 1;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* Z = lds + wid * 2 * 64 * EPL;
  float* Q = Z + 64 * EPL;
  const int s = (int)blockIdx.x * 4 + wid;
  if (s >= ntx * a.H) return;
  const int line = s / ntx, x0 = (s - line * ntx) * kXSeg;
  const int W = a.W;
  const unsigned row = (unsigned)line * (unsigned)W;
  const int sa = max(0, x0 - a.elo + 1);
  const int sb = min(W, x0 + kXSeg + a.ehi);
  const int segn = sb - sa;

  // Own pixels x0 + lane + 64 i.
  uint32_t bo[4];
  float go[4], co[4], to[4], so[4], tbo[4], cbo[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const unsigned gi = row + (unsigned)min(x0 + lane + 64 * i, W - 1);
    bo[i] = ld_bounds<NB>(a.bnd, gi);
    go[i] = ld32(a.gin, gi);
    co[i] = ld32(a.c, gi);
    to[i] = ld32(a.t, gi);
    so[i] = ld32(a.state, gi);
    tbo[i] = a.acc_first ? 0.f : ld32(a.tbar, gi);
    cbo[i] = a.acc_first ? 0.f : ld32(a.cbar, gi);
  }
  // z = xbar / n over the staged segment.
  float zc[NCH];
#pragma unroll
  for (int q = 0; q < NCH; q++) {
    const unsigned gi = row + (unsigned)(sa + min(64 * q + lane, segn - 1));
    const float g = ld32(a.gin, gi);
    const uint32_t b = ld_bounds<NB>(a.bnd, gi);
    const bool over = a.ow && ld32(a.c, gi) > a.theta;
    zc[q] = (over ? 0.f : g) * __builtin_amdgcn_rcpf((float)(bounds_dl<NB>(b) + bounds_du<NB>(b)));
  }
#pragma unroll
  for (int q = 0; q < NCH; q++) Z[64 * q + lane] = zc[q];
  wave_lds_sync();
  // Lane-serial odd runs + one wavefront scan (entries at or past segn only
  // reach entries at or past segn, which are never read).
  const int e0 = lane * EPL;
  float loc[EPL];
  float run = 0.f;
#pragma unroll
  for (int i = 0; i < EPL; i++) {
    run += Z[e0 + i];
    loc[i] = run;
  }
  const float off = wave_incl_scan_dpp(run) - run;
#pragma unroll
  for (int i = 0; i < EPL; i++) Q[e0 + i] = loc[i] + off;
  wave_lds_sync();

  const int2 ab = a.ab[line];
  float BT[4];
  if (x0 >= ab.x && x0 + kXSeg - 1 <= ab.y) {
    // No position of the segment collapses: B^T = S(u) - S(l), the other
    // four terms of box_adjoint_idx cancel exactly.
    float su[4], sl[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int p = x0 + lane + 64 * i;
      const int il = p - bounds_dl<NB>(bo[i]) - sa;
      su[i] = Q[p + bounds_du<NB>(bo[i]) - sa];
      const float r = Q[max(il, 0)];
      sl[i] = il >= 0 ? r : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) BT[i] = su[i] - sl[i];
  } else {
    // Every LDS read of the four positions before the first use (positions
    // past W are clamped to W - 1, as their bounds were, and not stored).
    float sv[4][6];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int pc = min(x0 + lane + 64 * i, W - 1);
      int kk[6];
      box_adjoint_idx(pc, pc - bounds_dl<NB>(bo[i]), pc + bounds_du<NB>(bo[i]), ab, W, sa, sb,
                      kk);
#pragma unroll
      for (int m = 0; m < 6; m++) {
        const float r = Q[max(kk[m] - sa, 0)];
        sv[i][m] = kk[m] >= sa ? r : 0.f;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) BT[i] = box_adjoint_eval(sv[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int p = x0 + lane + 64 * i;
    if (p >= W) continue;
    adj_x_finish(a, row + (unsigned)p, bounds_dl<NB>(bo[i]) + bounds_du<NB>(bo[i]), BT[i],
                 go[i], co[i], to[i], so[i], tbo[i], cbo[i]);
  }
}

// Staged rows [*sa, *sb) of the vertical adjoint tile at (tile_x, y0).  B^T
// of a row j whose window does not collapse reads the prefix sums at u(j),
// l(j), j and j - 1: rows inside j's forward window, which the extents of
// its band from the bounds pass (yext) cover.  A row whose lower window
// collapses (j < a <= E + 1) also sums every row from 0, one whose upper
// window collapses (j > b >= H - 2 - E) every row to H - 1, so tiles that
// can hold such rows stage to the line end.  Without yext: the caps.
template <int YR>
__device__ __forceinline__ void adj_tile_rows(const AdjArgs& a, int tile_x, int y0, int* sa,
                                              int* sb) {
  const int H = a.H;
  int lo = a.elo, hi = a.ehi;
  if (a.yext) {
    lo = 1;
    hi = 0;
    const int band = y0 / 32;
#pragma unroll
    for (int g = 0; g < 2; g++) {
      const int gi = 2 * tile_x + g;
      if (gi < a.ext_groups) {
        const int e = band * a.ext_groups + gi;
        lo = max(lo, a.yext[2 * e]);
        hi = max(hi, a.yext[2 * e + 1]);
      }
    }
    lo = min(lo, a.elo);
    hi = min(hi, a.ehi);
  }
  *sa = y0 <= a.ehi ? 0 : max(0, y0 - lo + 1);
  *sb = y0 + YR - 1 >= H - 1 - a.ehi ? H : min(H, y0 + YR + hi);
}

// Vertical adjoint over 64-column x 32-row tiles, staged rows from
// adj_tile_rows (at most cap), NW wavefronts.  RMAX > 0: wavefront q loads
// its contiguous share of the staged rows straight into registers (all loads
// in flight together, only its own rows) and scans it there, as the
// forward's register tile does; RMAX = 0: the rows go through LDS in rounds
// of 8.
template <int RMAX, bool NB, int YR, int NW = 4>
__global__ __launch_bounds__(64 * NW) void k_adj_y(AdjArgs a, int ntx, int nty, int cap) {
  extern __shared__ float lds[];
  float* Q = lds;                      // [cap][kYStride]
  float* tot = lds + cap * kYStride;   // [NW][64]
  const int r = xcd_remap((int)blockIdx.x, ntx * nty);
  const int tile_x = r % ntx, tile_y = r / ntx;
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = a.W, H = a.H;
  const int x0 = tile_x * kYCols, y0 = tile_y * YR;
  const int x = x0 + lane;
  const bool xok = x < W;
  const unsigned xc = (unsigned)min(x, W - 1);
  int sa, sb;
  adj_tile_rows<YR>(a, tile_x, y0, &sa, &sb);
  const int segn = sb - sa;
  constexpr int PR = YR / NW;
  uint32_t bo[PR];
  float go[PR];
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const unsigned gi = (unsigned)min(y0 + q + NW * k, H - 1) * (unsigned)W + xc;
    bo[k] = ld_bounds<NB>(a.bnd, gi);
    go[k] = ld32(a.gin, gi);
  }
  const int R = (segn + NW - 1) / NW;
  const int r0 = q * R, r1 = min(segn, r0 + R);
  if (RMAX > 0) {
    constexpr int RM = RMAX > 0 ? RMAX : 1;
    const int rlim = r1 - r0;  // wave-uniform
    float z[RM];
#pragma unroll
    for (int k = 0; k < RM; k++) {
      z[k] = 0.f;
      if (k < rlim) {
        const unsigned gi = (unsigned)(sa + r0 + k) * (unsigned)W + xc;
        const float g = ld32(a.gin, gi);
        const uint32_t b = ld_bounds<NB>(a.bnd, gi);
        z[k] = g * __builtin_amdgcn_rcpf((float)(bounds_dl<NB>(b) + bounds_du<NB>(b)));
      }
    }
    float run = 0.f;
#pragma unroll
    for (int k = 0; k < RM; k++) {
      run += k < rlim ? z[k] : 0.f;
      z[k] = run;
    }
    tot[q * kYCols + lane] = run;
    __syncthreads();
    float off = 0.f;
    for (int k = 0; k < q; k++) off += tot[k * kYCols + lane];
#pragma unroll
    for (int k = 0; k < RM; k++)
      if (k < rlim) Q[(r0 + k) * kYStride + lane] = z[k] + off;
    __syncthreads();
  } else {
    for (int rr = q; rr < segn; rr += NW * 8) {
      float z[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const unsigned gi = (unsigned)(sa + min(rr + NW * k, segn - 1)) * (unsigned)W + xc;
        const float g = ld32(a.gin, gi);
        const uint32_t b = ld_bounds<NB>(a.bnd, gi);
        z[k] = g * __builtin_amdgcn_rcpf((float)(bounds_dl<NB>(b) + bounds_du<NB>(b)));
      }
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (rr + NW * k < segn) Q[(rr + NW * k) * kYStride + lane] = z[k];
    }
    __syncthreads();
    float run = 0.f;
    for (int row = r0; row < r1; row++) {
      run += Q[row * kYStride + lane];
      Q[row * kYStride + lane] = run;
    }
    tot[q * kYCols + lane] = run;
    __syncthreads();
    float off = 0.f;
    for (int k = 0; k < q; k++) off += tot[k * kYCols + lane];
    if (q > 0)
      for (int row = r0; row < r1; row++) Q[row * kYStride + lane] += off;
    __syncthreads();
  }
  if (!xok) return;
  const int2 ab = a.ab[x];
  const float sl = a.step * a.lambda;
  float BT[PR];
  if (y0 >= ab.x && y0 + YR - 1 <= ab.y) {
    // No row of the column's tile collapses: B^T = S(u) - S(l), the other
    // four terms of box_adjoint_idx cancel exactly.
    float su[PR], sl0[PR];
#pragma unroll
    for (int k = 0; k < PR; k++) {
      const int y = y0 + q + NW * k;
      const int il = y - bounds_dl<NB>(bo[k]) - sa;
      su[k] = Q[(y + bounds_du<NB>(bo[k]) - sa) * kYStride + lane];
      const float r = Q[max(il, 0) * kYStride + lane];
      sl0[k] = il >= 0 ? r : 0.f;
    }
#pragma unroll
    for (int k = 0; k < PR; k++) BT[k] = su[k] - sl0[k];
  } else {
    // Every LDS read of the tile's rows before the first use (rows past H
    // are clamped to H - 1, as their bounds were, and not stored).
    float sv[PR][6];
#pragma unroll
    for (int k = 0; k < PR; k++) {
      const int yc = min(y0 + q + NW * k, H - 1);
      int kk[6];
      box_adjoint_idx(yc, yc - bounds_dl<NB>(bo[k]), yc + bounds_du<NB>(bo[k]), ab, H, sa, sb,
                      kk);
#pragma unroll
      for (int m = 0; m < 6; m++) {
        const float r = Q[max(kk[m] - sa, 0) * kYStride + lane];
        sv[k][m] = kk[m] >= sa ? r : 0.f;
      }
    }
#pragma unroll
    for (int k = 0; k < PR; k++) BT[k] = box_adjoint_eval(sv[k]);
  }
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + NW * k;
    if (y >= H) continue;
    st32(a.vout, (unsigned)y * (unsigned)W + (unsigned)x, (1.f - sl) * go[k] + sl * BT[k]);
  }
}

// -------------------------------- one launch per pair of adjoint passes
// The horizontal adjoint of iteration k and the vertical adjoint of
// iteration k - 1 as one launch over 64-column x YR-row tiles of NW
// wavefronts.  The tile's vertical windows need the horizontal pass's output
// on rows [sa, sb) (adj_tile_rows), which the tile recomputes into LDS
// (wavefront q takes rows sa + q, sa + q + NW, ..., the next row's loads in
// flight while the current one is processed) from the horizontal cotangent
// ax.gin, storing d/dt and d/dc on its own rows only; the vertical adjoint
// then runs on those rows as k_adj_y does.  The pass between them never
// reaches global memory.  ay.vout must differ from ax.gin (tiles read rows
// that other tiles write).

// One staged row of the horizontal adjoint: the masked cotangent over n of
// the NCX chunks of the row's window span [xa, xa + segx), and the lane's
// own pixel (x0 + lane): horizontal bounds, cotangent, confidence, target,
// pass state, vertical bounds and (own rows) the accumulated d/dt, d/dc.
template <int NCX>
struct AdjRow {
  float z[NCX];
  uint32_t bo, bv;
  float go, co, to, so, tbo, cbo;
};

template <int NCX, bool NBH>
__device__ __forceinline__ void adj_row_load(const AdjArgs& ax, const AdjArgs& ay, int row,
                                             int x0, int xa, int segx, int lane, bool own,
                                             AdjRow<NCX>& R) {
  const unsigned rowoff = (unsigned)row * (unsigned)ax.W;
#pragma unroll
  for (int q = 0; q < NCX; q++) {
    const unsigned gi = rowoff + (unsigned)(xa + min(64 * q + lane, segx - 1));
    const float g = ld32(ax.gin, gi);
    const uint32_t b = ld_bounds<NBH>(ax.bnd, gi);
    const bool over = ax.ow && ld32(ax.c, gi) > ax.theta;
    R.z[q] = (over ? 0.f : g) *
             __builtin_amdgcn_rcpf((float)(bounds_dl<NBH>(b) + bounds_du<NBH>(b)));
  }
  const unsigned gi = rowoff + (unsigned)min(x0 + lane, ax.W - 1);
  R.bo = ld_bounds<NBH>(ax.bnd, gi);
  R.go = ld32(ax.gin, gi);
  R.co = ld32(ax.c, gi);
  R.to = ld32(ax.t, gi);
  R.so = ld32(ax.state, gi);
  R.bv = ld_bounds<false>(ay.bnd, gi);
  R.tbo = own && !ax.acc_first ? ld32(ax.tbar, gi) : 0.f;
  R.cbo = own && !ax.acc_first ? ld32(ax.cbar, gi) : 0.f;
}

// Horizontal adjoint of pixel (row, x0 + lane) from a loaded row: prefix
// sums of z (as in k_adj_x) formed in place in the wavefront's LDS slice Q
// (z is not read again); stores d/dt, d/dc on own rows and returns the
// cotangent of the pass input.  Consecutive rows of a wavefront alternate
// two slices, so a row's writes need no wait for the previous row's reads.
template <int NCX, bool NBH>
__device__ __forceinline__ float adj_row_process(const AdjArgs& ax, int row, int x0, int xa,
                                                 int segx, int lane, bool own,
                                                 const AdjRow<NCX>& R, float* Q) {
  constexpr int EPL = NCX | 1;
  const int W = ax.W;
#pragma unroll
  for (int q = 0; q < NCX; q++) Q[64 * q + lane] = R.z[q];
  wave_lds_sync();
  // Entries at or past segx only reach entries at or past segx, never read.
  const int e0 = lane * EPL;
  float loc[EPL];
  float run = 0.f;
#pragma unroll
  for (int i = 0; i < EPL; i++) {
    run += Q[e0 + i];
    loc[i] = run;
  }
  // Each lane rewrites only its own run.
  const float off = wave_incl_scan_dpp(run) - run;
#pragma unroll
  for (int i = 0; i < EPL; i++) Q[e0 + i] = loc[i] + off;
  wave_lds_sync();
  const int2 ab = ax.ab[row];
  const int pc = min(x0 + lane, W - 1);
  const int dl = bounds_dl<NBH>(R.bo), du = bounds_du<NBH>(R.bo);
  float BT;
  if (x0 >= ab.x && x0 + kYCols - 1 <= ab.y) {
    // No position of the tile's columns collapses: B^T = S(u) - S(l).
    const int il = pc - dl - xa;
    const float su = Q[pc + du - xa];
    const float r = Q[max(il, 0)];
    BT = su - (il >= 0 ? r : 0.f);
  } else {
    int kk[6];
    box_adjoint_idx(pc, pc - dl, pc + du, ab, W, xa, xa + segx, kk);
    float sv[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
      const float r = Q[max(kk[m] - xa, 0)];
      sv[m] = kk[m] >= xa ? r : 0.f;
    }
    BT = box_adjoint_eval(sv);
  }
  const AdjXOut o = adj_x_values(ax, dl + du, BT, R.go, R.co, R.to, R.so, R.tbo, R.cbo);
  if (own && x0 + lane < W) {
    const unsigned gi = (unsigned)row * (unsigned)W + (unsigned)(x0 + lane);
    st32(ax.tbar, gi, o.tb);
    st32(ax.cbar, gi, o.cb);
  }
  return o.vo;
}

// LDS floats of a pair tile: the staged rows (z of the vertical pass, then
// its column prefix), the own rows' horizontal output, the column totals and
// the wavefronts' horizontal scan slices.
__host__ __device__ constexpr size_t adj_xy_lds(int cap, int yr, int nw, int ncx) {
  return sizeof(float) * ((size_t)cap * kYStride + (size_t)yr * kYCols + (size_t)nw * kYCols +
                          (size_t)nw * 2 * 64 * (ncx | 1));
}

template <int NCX, bool NBH, int YR>
__global__ __launch_bounds__(64 * adj_y_waves(YR)) void k_adj_xy(AdjArgs ax, AdjArgs ay,
                                                                  int ntx, int nty, int cap) {
  constexpr int NW = adj_y_waves(YR);
  constexpr int PR = YR / NW;
  extern __shared__ float lds[];
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* T = lds;                   // [cap][kYStride]
  float* G = T + cap * kYStride;    // [YR][kYCols]
  float* tot = G + YR * kYCols;     // [NW][kYCols]
  float* Q0 = tot + NW * kYCols + q * 2 * 64 * (NCX | 1);  // two slices per wavefront
  float* Q1 = Q0 + 64 * (NCX | 1);
  const int r = xcd_remap((int)blockIdx.x, ntx * nty);
  const int tile_x = r % ntx, tile_y = r / ntx;
  const int W = ay.W, H = ay.H;
  const int x0 = tile_x * kYCols, y0 = tile_y * YR;
  const int x = x0 + lane;
  const bool xok = x < W;
  const unsigned xc = (unsigned)min(x, W - 1);
  int sa, sb;
  adj_tile_rows<YR>(ay, tile_x, y0, &sa, &sb);
  const int segn = sb - sa;
  const int xa = max(0, x0 - ax.elo + 1);
  const int segx = min(W, x0 + kYCols + ax.ehi) - xa;
  // Vertical bounds of the output rows, loaded before the horizontal work.
  uint32_t bo[PR];
#pragma unroll
  for (int k = 0; k < PR; k++)
    bo[k] = ld_bounds<false>(ay.bnd, (unsigned)min(y0 + q + NW * k, H - 1) * (unsigned)W + xc);
  // Horizontal adjoint of the staged rows: z = its output over the vertical
  // window size into T, the output itself into G on own rows.
  {
    auto own = [&](int rr) { return sa + rr >= y0 && sa + rr < y0 + YR; };
    auto put = [&](int rr, const AdjRow<NCX>& R) {
      const bool ow = own(rr);
      // Wavefront q's rows rr = q, q + NW, ... alternate the slices.
      float* Q = ((rr / NW) & 1) ? Q1 : Q0;
      const float vo = adj_row_process<NCX, NBH>(ax, sa + rr, x0, xa, segx, lane, ow, R, Q);
      T[rr * kYStride + lane] =
          vo * __builtin_amdgcn_rcpf((float)(bounds_dl<false>(R.bv) + bounds_du<false>(R.bv)));
      if (ow) G[(sa + rr - y0) * kYCols + lane] = vo;
    };
    int rr = q;
    AdjRow<NCX> A, B;
    if (rr < segn) adj_row_load<NCX, NBH>(ax, ay, sa + rr, x0, xa, segx, lane, own(rr), A);
    while (rr < segn) {
      const int r1 = rr + NW;
      if (r1 < segn) adj_row_load<NCX, NBH>(ax, ay, sa + r1, x0, xa, segx, lane, own(r1), B);
      put(rr, A);
      if (r1 >= segn) break;
      const int r2 = r1 + NW;
      if (r2 < segn) adj_row_load<NCX, NBH>(ax, ay, sa + r2, x0, xa, segx, lane, own(r2), A);
      put(r1, B);
      rr = r2;
    }
  }
  __syncthreads();
  // Column prefix sums of z: each wavefront scans a contiguous share of the
  // staged rows, the shares' totals meet in LDS.
  const int R = (segn + NW - 1) / NW;
  const int r0 = q * R, r1 = min(segn, r0 + R);
  float run = 0.f;
  for (int row = r0; row < r1; row++) {
    run += T[row * kYStride + lane];
    T[row * kYStride + lane] = run;
  }
  tot[q * kYCols + lane] = run;
  __syncthreads();
  float off = 0.f;
  for (int k = 0; k < q; k++) off += tot[k * kYCols + lane];
  if (q > 0)
    for (int row = r0; row < r1; row++) T[row * kYStride + lane] += off;
  __syncthreads();
  if (!xok) return;
  const int2 ab = ay.ab[x];
  const float sl = ay.step * ay.lambda;
  float BT[PR];
  if (y0 >= ab.x && y0 + YR - 1 <= ab.y) {
    float su[PR], sl0[PR];
#pragma unroll
    for (int k = 0; k < PR; k++) {
      const int y = y0 + q + NW * k;
      const int il = y - bounds_dl<false>(bo[k]) - sa;
      su[k] = T[(y + bounds_du<false>(bo[k]) - sa) * kYStride + lane];
      const float rl = T[max(il, 0) * kYStride + lane];
      sl0[k] = il >= 0 ? rl : 0.f;
    }
#pragma unroll
    for (int k = 0; k < PR; k++) BT[k] = su[k] - sl0[k];
  } else {
    float sv[PR][6];
#pragma unroll
    for (int k = 0; k < PR; k++) {
      const int yc = min(y0 + q + NW * k, H - 1);
      int kk[6];
      box_adjoint_idx(yc, yc - bounds_dl<false>(bo[k]), yc + bounds_du<false>(bo[k]), ab, H, sa,
                      sb, kk);
#pragma unroll
      for (int m = 0; m < 6; m++) {
        const float rv = T[max(kk[m] - sa, 0) * kYStride + lane];
        sv[k][m] = kk[m] >= sa ? rv : 0.f;
      }
    }
#pragma unroll
    for (int k = 0; k < PR; k++) BT[k] = box_adjoint_eval(sv[k]);
  }
#pragma unroll
  for (int k = 0; k < PR; k++) {
    const int y = y0 + q + NW * k;
    if (y >= H) continue;
    st32(ay.vout, (unsigned)y * (unsigned)W + (unsigned)x,
         (1.f - sl) * G[(q + NW * k) * kYCols + lane] + sl * BT[k]);
  }
}

// ------------------------------------------- wide windows (any extent)
// Windows beyond the LDS segments: prefixes of z = xbar / n along the lines
// in double precision (wide.h), then B^T of each pixel from its six prefix
// positions (box_adjoint_idx over the whole line).
template <bool NB>
struct AdjZRow {  // horizontal: the overwritten pixels pass no cotangent
  AdjArgs a;
  __device__ float operator()(unsigned i) const {
    const uint32_t b = ld_bounds<NB>(a.bnd, i);
    const bool over = a.ow && ld32(a.c, i) > a.theta;
    return (over ? 0.f : ld32(a.gin, i)) *
           __builtin_amdgcn_rcpf((float)(bounds_dl<NB>(b) + bounds_du<NB>(b)));
  }
};
struct AdjZCol {  // vertical planes of a single frame are wide
  AdjArgs a;
  __device__ float operator()(unsigned i) const {
    const uint32_t b = ld_bounds<false>(a.bnd, i);
    return ld32(a.gin, i) *
           __builtin_amdgcn_rcpf((float)(bounds_dl<false>(b) + bounds_du<false>(b)));
  }
};

__device__ __forceinline__ float box_adjoint_eval_d(const double* S, size_t stride,
                                                    const int (&k)[6]) {
  double v[6];
#pragma unroll
  for (int m = 0; m < 6; m++) v[m] = k[m] >= 0 ? S[(size_t)k[m] * stride] : 0.0;
  return (float)((v[0] - v[1]) - (v[2] - v[3]) - (v[4] - v[5]));
}

template <bool NB>
__global__ __launch_bounds__(256) void k_adj_wide_x(AdjArgs a, const double* __restrict__ S) {
  const int p = (int)blockIdx.x * 256 + threadIdx.x, line = (int)blockIdx.y;
  if (p >= a.W) return;
  const unsigned row = (unsigned)line * (unsigned)a.W, gi = row + (unsigned)p;
  const uint32_t b = ld_bounds<NB>(a.bnd, gi);
  int kk[6];
  box_adjoint_idx(p, p - bounds_dl<NB>(b), p + bounds_du<NB>(b), a.ab[line], a.W, 0, a.W, kk);
  const float BT = box_adjoint_eval_d(S + row, 1, kk);
  adj_x_finish(a, gi, bounds_dl<NB>(b) + bounds_du<NB>(b), BT, ld32(a.gin, gi), ld32(a.c, gi),
               ld32(a.t, gi), ld32(a.state, gi), a.acc_first ? 0.f : ld32(a.tbar, gi),
               a.acc_first ? 0.f : ld32(a.cbar, gi));
}

__global__ __launch_bounds__(256) void k_adj_wide_y(AdjArgs a, const double* __restrict__ S) {
  const int x = (int)blockIdx.x * 256 + threadIdx.x, y = (int)blockIdx.y;
  if (x >= a.W) return;
  const unsigned gi = (unsigned)y * (unsigned)a.W + (unsigned)x;
  const uint32_t b = ld_bounds<false>(a.bnd, gi);
  int kk[6];
  box_adjoint_idx(y, y - bounds_dl<false>(b), y + bounds_du<false>(b), a.ab[x], a.H, 0, a.H, kk);
  const float BT = box_adjoint_eval_d(S + x, (size_t)a.W, kk);
  const float sl = a.step * a.lambda;
  st32(a.vout, gi, (1.f - sl) * ld32(a.gin, gi) + sl * BT);
}

int adj_wide_scratch_check(const AdjArgs& a) {
  if (a.wS && a.wtot) return DTS_OK;
  set_error("window extent %d+%d needs the wide-window scratch", a.elo, a.ehi);
  return DTS_ERR_STATE;
}

__global__ void k_adj_acc(float* __restrict__ d, const float* __restrict__ s, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] += s[i];
}

int adj_collapse(hipStream_t s, const int64_t* ct, int L, int nlines, int64_t sfx, int E,
                 int2* ab) {
  k_adj_collapse<<<div_up(nlines, 256), 256, 0, s>>>(ct, L, nlines, sfx, E, ab);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int adj_collapse2(hipStream_t s, const int64_t* ctH, const int64_t* ctV, int W, int H,
                  int64_t sfx, int Eh, int Ev, int2* abH, int2* abV) {
  k_adj_collapse2<<<div_up(W + H, 256), 256, 0, s>>>(ctH, ctV, W, H, sfx, Eh, Ev, abH, abV);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

// Compute units of the current device (queried once).
int adj_num_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    cus = hipGetDevice(&dev) == hipSuccess &&
                  hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) ==
                      hipSuccess &&
                  n > 0
              ? n
              : 256;
  }
  return cus;
}

template <int NCH>
int adj_xpass_n(hipStream_t s, const AdjArgs& a) {
  const int ntx = div_up(a.W, kXSeg);
  const int nseg = ntx * a.H;
  const size_t lds = sizeof(float) * 4 * 2 * 64 * (NCH | 1);
  if (a.bnarrow)
    k_adj_x<NCH, true><<<div_up(nseg, 4), 256, lds, s>>>(a, ntx);
  else
    k_adj_x<NCH, false><<<div_up(nseg, 4), 256, lds, s>>>(a, ntx);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int adj_xpass(hipStream_t s, const AdjArgs& a) {
  const int need = std::min(kXSeg + a.elo + a.ehi - 1, a.W);
  if (need <= 64 * 5) return adj_xpass_n<5>(s, a);
  if (need <= 64 * 6) return adj_xpass_n<6>(s, a);
  if (need <= 64 * 8) return adj_xpass_n<8>(s, a);
  int rc;
  if ((rc = adj_wide_scratch_check(a))) return rc;
  if (a.bnarrow) {
    DTS_HIP_CHECK(line_prefix(s, AdjZRow<true>{a}, a.W, a.H, true, a.wtot, a.wS));
    k_adj_wide_x<true><<<dim3(div_up(a.W, 256), a.H), 256, 0, s>>>(a, a.wS);
  } else {
    DTS_HIP_CHECK(line_prefix(s, AdjZRow<false>{a}, a.W, a.H, true, a.wtot, a.wS));
    k_adj_wide_x<false><<<dim3(div_up(a.W, 256), a.H), 256, 0, s>>>(a, a.wS);
  }
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

// Vertical bounds planes of a single frame are wide (launch_bounds).
template <int RMAX, int YR>
void adj_y_launch(hipStream_t s, const AdjArgs& a, dim3 grid, size_t lds, int ntx, int nty,
                  int cap) {
  constexpr int NW = adj_y_waves(YR);
  k_adj_y<RMAX, false, YR, NW><<<grid, 64 * NW, lds, s>>>(a, ntx, nty, cap);
}

template <int YR>
void adj_y_dispatch(hipStream_t s, const AdjArgs& a, size_t lds, int cap) {
  const int ntx = div_up(a.W, kYCols), nty = div_up(a.H, YR);
  const dim3 grid((unsigned)(ntx * nty));
  // Register class of one wavefront's share of the staged rows: with 8
  // wavefronts, the class of half the rows split four ways.
  switch (ytile_rmax(adj_y_waves(YR) == 8 ? (cap + 1) / 2 : cap)) {
    case 16: adj_y_launch<16, YR>(s, a, grid, lds, ntx, nty, cap); break;
    case 24: adj_y_launch<24, YR>(s, a, grid, lds, ntx, nty, cap); break;
    case 32: adj_y_launch<32, YR>(s, a, grid, lds, ntx, nty, cap); break;
    case 48: adj_y_launch<48, YR>(s, a, grid, lds, ntx, nty, cap); break;
    default: adj_y_launch<0, YR>(s, a, grid, lds, ntx, nty, cap); break;
  }
}


int adj_ypass(hipStream_t s, const AdjArgs& a) {
  if (a.bnarrow) {
    set_error("vertical bounds planes of a single frame are wide");
    return DTS_ERR_STATE;
  }
  // 16-row tiles where 32-row ones would give fewer than two per compute
  // unit: the tile's chain of staged rows, not bandwidth, then sets the time.
  const bool short_tiles =
      div_up(a.W, kYCols) * div_up(a.H, kYRows) < 2 * adj_num_cus();
  const int yr = short_tiles ? 16 : kYRows;
  const int cap = yr + a.elo + a.ehi;
  const size_t lds =
      sizeof(float) * ((size_t)cap * kYStride + (size_t)adj_y_waves(yr) * kYCols);
  if (lds > kLdsMax) {
    int rc;
    if ((rc = adj_wide_scratch_check(a))) return rc;
    DTS_HIP_CHECK(line_prefix(s, AdjZCol{a}, a.W, a.H, false, a.wtot, a.wS));
    k_adj_wide_y<<<dim3(div_up(a.W, 256), a.H), 256, 0, s>>>(a, a.wS);
    DTS_LAUNCH_CHECK();
    return DTS_OK;
  }
  if (short_tiles)
    adj_y_dispatch<16>(s, a, lds, cap);
  else
    adj_y_dispatch<kYRows>(s, a, lds, cap);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

// Launch shape of the pair launch (k_adj_xy); false when it does not apply
// (a horizontal window span beyond 4 chunks, the tile beyond LDS).
bool adj_xy_plan(const AdjArgs& ax, const AdjArgs& ay, int* ncx, int* yr, int* cap,
                 size_t* lds) {
  if (ay.bnarrow) return false;
  *ncx = std::max(2, div_up(63 + ax.elo + ax.ehi, 64));
  if (*ncx > 4) return false;
  const bool short_tiles = div_up(ay.W, kYCols) * div_up(ay.H, kYRows) < 2 * adj_num_cus();
  *yr = short_tiles ? 16 : kYRows;
  *cap = *yr + ay.elo + ay.ehi;
  *lds = adj_xy_lds(*cap, *yr, adj_y_waves(*yr), *ncx);
  return *lds <= kLdsMax;
}

bool adj_xy_fits(const AdjArgs& ax, const AdjArgs& ay) {
  int ncx, yr, cap;
  size_t lds;
  return adj_xy_plan(ax, ay, &ncx, &yr, &cap, &lds);
}

template <int NCX, bool NBH>
void adj_xy_launch(hipStream_t s, const AdjArgs& ax, const AdjArgs& ay, int yr, int cap,
                   size_t lds) {
  const int ntx = div_up(ay.W, kYCols), nty = div_up(ay.H, yr);
  const dim3 grid((unsigned)(ntx * nty));
  if (yr == 16)
    k_adj_xy<NCX, NBH, 16><<<grid, 64 * adj_y_waves(16), lds, s>>>(ax, ay, ntx, nty, cap);
  else
    k_adj_xy<NCX, NBH, kYRows><<<grid, 64 * adj_y_waves(kYRows), lds, s>>>(ax, ay, ntx, nty,
                                                                            cap);
}

int adj_xypass(hipStream_t s, const AdjArgs& ax, const AdjArgs& ay) {
  int ncx, yr, cap;
  size_t lds;
  if (!adj_xy_plan(ax, ay, &ncx, &yr, &cap, &lds)) {
    set_error("windows too wide for the adjoint pair launch");
    return DTS_ERR_UNSUPPORTED;
  }
  const bool nb = ax.bnarrow != 0;
  switch (ncx) {
    case 2: nb ? adj_xy_launch<2, true>(s, ax, ay, yr, cap, lds)
               : adj_xy_launch<2, false>(s, ax, ay, yr, cap, lds); break;
    case 3: nb ? adj_xy_launch<3, true>(s, ax, ay, yr, cap, lds)
               : adj_xy_launch<3, false>(s, ax, ay, yr, cap, lds); break;
    default: nb ? adj_xy_launch<4, true>(s, ax, ay, yr, cap, lds)
                : adj_xy_launch<4, false>(s, ax, ay, yr, cap, lds); break;
  }
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int adj_accumulate(hipStream_t s, float* dst, const float* src, size_t n) {
  if (n == 0) return DTS_OK;
  k_adj_acc<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dst, src, n);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

}  // namespace dtsamd
