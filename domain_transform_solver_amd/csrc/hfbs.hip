// HFBS (bilateral-grid solver) on gfx950: kernels and the hfbs_* C ABI.
//
// src/hfbs.cu:586-743.  Grids are packed [gl][gy][gx] fp32 arrays of the
// current grid size; reads outside [0, grid) are 0.  Pipeline per Optimize:
//   splat        one workgroup per 4x4-cell tile accumulates b = sum c*t,
//                c_g = sum c and S = pixel count for all luma bins in LDS
//                (ds_add_f32), then stores its voxels (each voxel has one
//                owner, no global atomics);
//   bistoch.     K_b Jacobi steps n' = sqrt(n (S + eps) / (2n + sum6 n + eps));
//   inv diag     invA = 1 / max(1e-3, lambda (S - 2 n^2) + c_g);
//   blur         separable symmetric exponential sweeps of c_g and b;
//   init         z = blur(b) / (blur(c_g) + 1e-7);
//   solve        K heavy-ball steps with the Huber-clamped gradient;
//   slice        trilinear, splat-count weighted.
#include <cmath>
#include <cstring>
#include <vector>

#include "dts_kernels.h"

namespace dtsamd {

struct Grid {
  int gx, gy, gl;
};

__device__ inline size_t gidx(const Grid& g, int x, int y, int z) {
  return ((size_t)z * g.gy + y) * g.gx + x;
}

__device__ inline float gget(const float* __restrict__ a, const Grid& g, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= g.gx || y >= g.gy || z >= g.gl) return 0.f;
  return a[gidx(g, x, y, z)];
}

// Grid coordinate as the reference computes it (x * 1.0f / sigma in fp32),
// clamped into the grid.
__device__ inline float gcoord(int v, float sigma, int gmax) {
  const float g = (float)v * 1.0f / sigma;
  return fminf(g, (float)(gmax - 1));
}

__global__ void k_hf_convert(uchar4* __restrict__ img, int n, int cs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) img[i] = convert_px(img[i], cs);
}

__global__ void k_hf_fill(float* __restrict__ p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_hf_copy(const float* __restrict__ s, float* __restrict__ d, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

constexpr int kCellTile = 4;  // cells per splat tile side

// Splat (src/hfbs.cu:71-120).  Workgroup = kCellTile x kCellTile spatial
// cells x all luma bins.  Pixels of a slightly larger rectangle are visited
// and kept when their own cell falls in the tile.
__global__ __launch_bounds__(256) void k_hf_splat(const uchar4* __restrict__ guide,
                                                  const float* __restrict__ t,
                                                  const float* __restrict__ c, int W, int H,
                                                  float sx, float sy, float sl, Grid g,
                                                  float* __restrict__ b,
                                                  float* __restrict__ cg,
                                                  float* __restrict__ S,
                                                  float* __restrict__ n0,
                                                  float* __restrict__ hb) {
  extern __shared__ float acc[];  // [3][gl][kCellTile*kCellTile]
  const int ntx = (g.gx + kCellTile - 1) / kCellTile;
  const int cx0 = (blockIdx.x % ntx) * kCellTile, cy0 = (blockIdx.x / ntx) * kCellTile;
  constexpr int kCells = kCellTile * kCellTile;
  const int nacc = 3 * g.gl * kCells;
  for (int i = threadIdx.x; i < nacc; i += 256) acc[i] = 0.f;
  __syncthreads();
  const int xlo = max(0, (int)floorf(cx0 * sx) - 1);
  const int xhi = min(W, (int)ceilf((cx0 + kCellTile) * sx) + 1);
  const int ylo = max(0, (int)floorf(cy0 * sy) - 1);
  const int yhi = min(H, (int)ceilf((cy0 + kCellTile) * sy) + 1);
  const int rw = xhi - xlo, n = rw * max(0, yhi - ylo);
  const int lane = threadIdx.x & 63;
  // Whole wavefronts per round (the scan below needs every lane); lanes
  // past n carry no pixel.
  for (int i0 = threadIdx.x - lane; i0 < n; i0 += 256) {
    const int i = i0 + lane;
    int cell = -1;
    float vb = 0.f, vc = 0.f, vs = 0.f;
    if (i < n) {
      const int y = ylo + i / rw, x = xlo + i % rw;
      const int ix = (int)floorf(gcoord(x, sx, g.gx)) - cx0;
      const int iy = (int)floorf(gcoord(y, sy, g.gy)) - cy0;
      if (ix >= 0 && iy >= 0 && ix < kCellTile && iy < kCellTile) {
        const size_t pi = (size_t)y * W + x;
        const int iz = (int)floorf(gcoord(guide[pi].x, sl, g.gl));
        const float cv = c[pi];
        cell = (iz * kCells) + iy * kCellTile + ix;
        vb = cv * t[pi];
        vc = cv;
        vs = 1.f;
      }
    }
    // Consecutive lanes are consecutive pixels of a row, so runs of lanes
    // share a voxel: a segmented wavefront scan sums each run, and its last
    // lane adds the sums with one atomic per grid (instead of one per lane
    // on a contended LDS address).
    // Both neighbour shuffles run on every lane (a shuffle inside a
    // short-circuit would leave its source lanes inactive).
    const int cprev = __shfl_up(cell, 1, 64), cnext = __shfl_down(cell, 1, 64);
    int f = (lane == 0) | (cprev != cell);  // run head
    for (int d = 1; d < 64; d <<= 1) {
      const float tb = __shfl_up(vb, d, 64), tc = __shfl_up(vc, d, 64), ts = __shfl_up(vs, d, 64);
      const int tf = __shfl_up(f, d, 64);
      if (lane >= d) {
        if (!f) {
          vb += tb;
          vc += tc;
          vs += ts;
        }
        f |= tf;
      }
    }
    const bool tail = (lane == 63) | (cnext != cell);
    if (tail && cell >= 0) {
      atomicAdd(&acc[cell], vb);
      atomicAdd(&acc[g.gl * kCells + cell], vc);
      atomicAdd(&acc[2 * g.gl * kCells + cell], vs);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < g.gl * kCells; i += 256) {
    const int iz = i / kCells, rem = i % kCells;
    const int x = cx0 + rem % kCellTile, y = cy0 + rem / kCellTile;
    if (x < g.gx && y < g.gy) {
      const size_t v = gidx(g, x, y, iz);
      b[v] = acc[i];
      cg[v] = acc[g.gl * kCells + i];
      S[v] = acc[2 * g.gl * kCells + i];
      // The starting values of the bistochastization (n = 1) and of the
      // heavy-ball momentum (0), stored by the voxel's owner.
      n0[v] = 1.f;
      hb[v] = 0.f;
    }
  }
}

// src/hfbs.cu:125-155
// Stencil kernels run 256 consecutive (x, y) columns (column c = y gx + x)
// per workgroup, each thread two consecutive z (kGridZ): the values at z and
// z + 1 serve as each other's z neighbour.  Numbering the columns through
// the rows leaves no idle lanes at the row ends, so the 1080p colorization
// grid (240 x 135 x 32) is 8,128 wavefronts, all resident at once, where
// 64 x 4 blocks were 8,704, a second partial round every launch.
constexpr int kStencilCols = 256;
constexpr int kGridZ = 2;

// Thread's column c and z block of a stencil launch (a 1-D grid of
// ceil(gx gy / kStencilCols) * ceil(gl / kGridZ) workgroups).  The
// workgroups of one XCD take consecutive block numbers (xcd_remap), and the
// numbering runs z fastest, then the column block: each XCD holds every z
// of a contiguous band of columns, so the z and x neighbours a block reads,
// and most y neighbours, sit in its own L2.
struct StencilBlock {
  int c, bz;
};

// A thread's column (x, y) of a stencil block: 32-bit voxel offsets (grids
// stay below 2^31 voxels, hfbs_create) and which of the four in-plane
// neighbours lie inside the grid; reads outside are 0.
struct StencilCol {
  int base;        // offset of (x, y, 0)
  int plane;       // gx * gy
  int gx, gl;
  bool xl, xr, yl, yr;
  __device__ float at(const float* __restrict__ a, int z) const {  // (x, y, z), any z
    return z >= 0 && z < gl ? a[base + z * plane] : 0.f;
  }
  // Sum of the four in-plane neighbours at voxel v, in the order
  // (x - 1) + (x + 1) + (y - 1) + (y + 1).
  __device__ float sum4(const float* __restrict__ a, int v) const {
    const float a0 = xl ? a[v - 1] : 0.f, a1 = xr ? a[v + 1] : 0.f;
    const float a2 = yl ? a[v - gx] : 0.f, a3 = yr ? a[v + gx] : 0.f;
    return a0 + a1 + a2 + a3;
  }
};

__device__ inline StencilCol stencil_col(const Grid& g, int x, int y) {
  StencilCol c;
  c.base = y * g.gx + x;
  c.plane = g.gx * g.gy;
  c.gx = g.gx;
  c.gl = g.gl;
  c.xl = x > 0;
  c.xr = x + 1 < g.gx;
  c.yl = y > 0;
  c.yr = y + 1 < g.gy;
  return c;
}

__device__ inline StencilBlock stencil_block(const Grid& g) {
  const int nbz = (g.gl + kGridZ - 1) / kGridZ;
  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  StencilBlock b;
  b.bz = L % nbz;
  b.c = (L / nbz) * kStencilCols + (int)threadIdx.x;
  return b;
}

// invA non-null (the last step): also the inverse diagonal of the final n,
// as k_hf_invdiag.
__global__ __launch_bounds__(256) void k_hf_bistoch(const float* __restrict__ S,
                                                    const float* __restrict__ ns,
                                                    float* __restrict__ nd, Grid g,
                                                    const float* __restrict__ cg,
                                                    float* __restrict__ invA, float lambda) {
  const StencilBlock sb = stencil_block(g);
  if (sb.c >= g.gx * g.gy) return;
  const int x = sb.c % g.gx, y = sb.c / g.gx, z0 = sb.bz * kGridZ;
  const StencilCol col = stencil_col(g, x, y);
  float c[kGridZ + 2];  // n at z0 - 1 .. z0 + kGridZ
#pragma unroll
  for (int k = 0; k < kGridZ + 2; k++) c[k] = col.at(ns, z0 - 1 + k);
#pragma unroll
  for (int k = 0; k < kGridZ; k++) {
    const int z = z0 + k;
    if (z >= g.gl) break;
    const int v = col.base + z * col.plane;
    const float d = col.sum4(ns, v) + c[k] + c[k + 2];
    const float n = c[k + 1];
    const float bl = 2.f * n + d;
    const float sv = S[v];
    const float nn = sqrtf(n * (sv + 0.00001f) / (bl + 0.00001f));
    nd[v] = nn;
    if (invA) {
      const float diag = lambda * (sv - 2.f * nn * nn) + cg[v];
      invA[v] = (float)(1.0 / fmax(1e-3, (double)diag));
    }
  }
}

// src/hfbs.cu:158-195
__global__ void k_hf_invdiag(const float* __restrict__ S, const float* __restrict__ n,
                             const float* __restrict__ cg, float* __restrict__ invA,
                             size_t nv, float lambda) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float nn = n[v];
  const float diag = lambda * (S[v] - 2.f * nn * nn) + cg[v];
  // The reference takes max(1e-3, diag) in double (A_VERT_DIAG_MIN is a
  // double literal) and divides 1.0f by it.
  invA[v] = (float)(1.0 / fmax(1e-3, (double)diag));
}

// Symmetric exponential sweep y_i = sum_j beta^|i-j| x_j along one axis
// (src/hfbs.cu:357-427): forward recurrence, then the backward tail added.
// One thread per line; lines along x are strided across lanes, lines along
// y and z have adjacent lanes on adjacent x.
__global__ void k_hf_sweep(const float* __restrict__ src, float* __restrict__ dst, Grid g,
                           int axis, float beta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n;
  size_t base, stride;
  if (axis == 0) {
    if (i >= g.gy * g.gl) return;
    n = g.gx;
    base = (size_t)i * g.gx;
    stride = 1;
  } else if (axis == 1) {
    if (i >= g.gx * g.gl) return;
    const int x = i % g.gx, z = i / g.gx;
    n = g.gy;
    base = (size_t)z * g.gx * g.gy + x;
    stride = g.gx;
  } else {
    if (i >= g.gx * g.gy) return;
    n = g.gl;
    base = i;
    stride = (size_t)g.gx * g.gy;
  }
  float run = src[base];
  dst[base] = run;
  for (int k = 1; k < n; k++) {
    run = src[base + k * stride] + beta * run;
    dst[base + k * stride] = run;
  }
  run = src[base + (size_t)(n - 1) * stride];
  for (int k = n - 2; k >= 0; k--) {
    run = beta * run;
    dst[base + k * stride] += run;
    run = run + src[base + k * stride];
  }
}

// The same sweep as parallel scans.  y_i = sum_j beta^|i-j| x_j splits into
// f_i = sum_{j<=i} beta^(i-j) x_j and g_i = sum_{j>=i} beta^(j-i) x_j with
// y = f + g - x, and both are decayed prefix sums: per-chunk serial sums from
// zero, plus beta^(distance) times the carry of all earlier chunks.

// Lines along x (contiguous): one wavefront per line, EP consecutive
// elements per lane (n <= 64 EP), carries by a shuffle scan whose step d
// scales by beta^(EP d).
template <int EP>
// blockIdx.y = 1 sweeps src2 into dst2 (two grids in one launch).
__global__ __launch_bounds__(256) void k_hf_sweep_x(const float* __restrict__ src,
                                                    float* __restrict__ dst,
                                                    const float* __restrict__ src2,
                                                    float* __restrict__ dst2, int n, int nlines,
                                                    float beta) {
  const int lane = threadIdx.x & 63;
  const int line = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (line >= nlines) return;
  const float* s = (blockIdx.y ? src2 : src) + (size_t)line * n;
  float* d = (blockIdx.y ? dst2 : dst) + (size_t)line * n;
  const int i0 = lane * EP;
  float x[EP], f[EP], g[EP];
#pragma unroll
  for (int k = 0; k < EP; k++) x[k] = i0 + k < n ? s[i0 + k] : 0.f;
  float bE = 1.f;
#pragma unroll
  for (int k = 0; k < EP; k++) bE *= beta;
  float run = 0.f;
#pragma unroll
  for (int k = 0; k < EP; k++) {
    run = x[k] + beta * run;
    f[k] = run;
  }
  float S = run, fac = bE;
#pragma unroll
  for (int dd = 1; dd < 64; dd <<= 1) {
    const float o = __shfl_up(S, dd, 64);
    if (lane >= dd) S += fac * o;
    fac *= fac;
  }
  float C = __shfl_up(S, 1, 64);
  if (lane == 0) C = 0.f;
  run = 0.f;
#pragma unroll
  for (int k = EP - 1; k >= 0; k--) {
    run = x[k] + beta * run;
    g[k] = run;
  }
  float R = run;
  fac = bE;
#pragma unroll
  for (int dd = 1; dd < 64; dd <<= 1) {
    const float o = __shfl_down(R, dd, 64);
    if (lane + dd < 64) R += fac * o;
    fac *= fac;
  }
  float D = __shfl_down(R, 1, 64);
  if (lane == 63) D = 0.f;
  float pf = beta, pg = bE;
#pragma unroll
  for (int k = 0; k < EP; k++) {
    f[k] += pf * C;
    pf *= beta;
  }
#pragma unroll
  for (int k = 0; k < EP; k++) {
    g[k] += pg * D;
    pg *= (1.f / beta);
  }
#pragma unroll
  for (int k = 0; k < EP; k++)
    if (i0 + k < n) d[i0 + k] = f[k] + g[k] - x[k];
}

// Lines along y (axis 1) or z (axis 2): 64 lines at adjacent x per
// workgroup, each wavefront one contiguous quarter of the line (CH <= CMAX
// elements) in registers; the four quarter carries meet in LDS.
struct SweepLine {
  int n;          // line length
  size_t base;    // offset of the line's first element
  size_t stride;  // between its elements
};

__device__ inline SweepLine sweep_line(const Grid& gd, int axis, int xc, int other) {
  SweepLine L;
  if (axis == 1) {  // line (x, z), along y
    L.n = gd.gy;
    L.base = (size_t)other * gd.gx * gd.gy + xc;
    L.stride = gd.gx;
  } else {  // line (x, y), along z
    L.n = gd.gl;
    L.base = (size_t)other * gd.gx + xc;
    L.stride = (size_t)gd.gx * gd.gy;
  }
  return L;
}

// The swept values of wavefront q's quarter [k0, k0 + len) of line L of
// src into out (one barrier; Fs, Gs are free again after the next one).
template <int CMAX>
__device__ __forceinline__ void sweep_quarter(const float* __restrict__ src, const SweepLine& L,
                                              int q, int lane, float beta, float (*Fs)[64],
                                              float (*Gs)[64], float (&out)[CMAX]) {
  const int n = L.n;
  const int CH = (n + 3) / 4;
  const int k0 = q * CH;
  const int len = max(0, min(CH, n - k0));
  float x[CMAX], f[CMAX], g[CMAX];
#pragma unroll
  for (int k = 0; k < CMAX; k++) x[k] = k < len ? src[L.base + (size_t)(k0 + k) * L.stride] : 0.f;
  float run = 0.f;
#pragma unroll
  for (int k = 0; k < CMAX; k++) {
    if (k < len) run = x[k] + beta * run;
    f[k] = run;
  }
  Fs[q][lane] = run;
  run = 0.f;
#pragma unroll
  for (int k = CMAX - 1; k >= 0; k--) {
    if (k < len) run = x[k] + beta * run;
    g[k] = run;
  }
  Gs[q][lane] = run;
  __syncthreads();
  // beta^CH (every quarter before the last non-empty one is full).
  float bC = 1.f;
  for (int k = 0; k < CH; k++) bC *= beta;
  float P = 0.f;  // f at the end of quarter q - 1, all earlier quarters included
  for (int m = 0; m < q; m++) P = Fs[m][lane] + bC * P;
  float Qn = 0.f;  // g at the start of quarter q + 1, all later quarters included
  for (int m = 3; m > q; m--) {
    const int lm = max(0, min(CH, n - m * CH));
    float bl = 1.f;
    for (int k = 0; k < lm; k++) bl *= beta;
    Qn = Gs[m][lane] + bl * Qn;
  }
  float pf = beta;
#pragma unroll
  for (int k = 0; k < CMAX; k++) {
    f[k] += pf * P;
    pf *= beta;
  }
  // g gains beta^(len - k) Qn.
  float pg = 1.f;
  for (int k = 0; k < len; k++) pg *= beta;
#pragma unroll
  for (int k = 0; k < CMAX; k++) {
    g[k] += pg * Qn;
    pg *= (1.f / beta);
  }
#pragma unroll
  for (int k = 0; k < CMAX; k++) out[k] = f[k] + g[k] - x[k];
}

// blockIdx.y = 1 sweeps src2 into dst2 (two grids in one launch).  INITZ
// (axis 2 only): every workgroup sweeps the same lines of src (the XY-swept
// splat confidence) and src2 (the XY-swept splat values), stores the first
// into dst and the initial solution blur(b) / (blur(c_g) + 1e-7) into dst2.
template <int CMAX, bool INITZ = false>
__global__ __launch_bounds__(256) void k_hf_sweep_yz(const float* __restrict__ src,
                                                     float* __restrict__ dst,
                                                     const float* __restrict__ src2,
                                                     float* __restrict__ dst2, Grid gd,
                                                     int axis, float beta) {
  __shared__ float Fs[4][64], Gs[4][64];
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nbx = (gd.gx + 63) / 64;
  const int bx = (int)blockIdx.x % nbx, other = (int)blockIdx.x / nbx;
  const int xl = bx * 64 + lane;
  const bool ok = xl < gd.gx;
  const SweepLine L = sweep_line(gd, axis, min(xl, gd.gx - 1), other);
  const int CH = (L.n + 3) / 4;
  const int k0 = q * CH;
  const int len = max(0, min(CH, L.n - k0));
  float a[CMAX];
  sweep_quarter<CMAX>(INITZ || !blockIdx.y ? src : src2, L, q, lane, beta, Fs, Gs, a);
  if (!INITZ) {
    float* d = blockIdx.y ? dst2 : dst;
    if (!ok) return;
#pragma unroll
    for (int k = 0; k < CMAX; k++)
      if (k < len) d[L.base + (size_t)(k0 + k) * L.stride] = a[k];
  } else {
    float b[CMAX];
    __syncthreads();  // Fs, Gs are read above
    sweep_quarter<CMAX>(src2, L, q, lane, beta, Fs, Gs, b);
    if (!ok) return;
#pragma unroll
    for (int k = 0; k < CMAX; k++)
      if (k < len) {
        const size_t v = L.base + (size_t)(k0 + k) * L.stride;
        dst[v] = a[k];
        dst2[v] = (float)((double)b[k] / ((double)a[k] + 1e-7));
      }
  }
}

// src/hfbs.cu:335-353 (kUseLaplacianBlur)
__global__ void k_hf_initz(const float* __restrict__ bb, const float* __restrict__ cb,
                           float* __restrict__ z, size_t nv) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) z[v] = (float)((double)bb[v] / ((double)cb[v] + 1e-7));
}

// src/hfbs.cu:200-239
__global__ __launch_bounds__(256) void k_hf_solve(const float* __restrict__ n,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ invA,
                                                  const float* __restrict__ ys,
                                                  float* __restrict__ yd, float* __restrict__ hb,
                                                  Grid g, float lambda, float alpha,
                                                  float beta) {
  const StencilBlock sb = stencil_block(g);
  if (sb.c >= g.gx * g.gy) return;
  const int x = sb.c % g.gx, y = sb.c / g.gx, z0 = sb.bz * kGridZ;
  const StencilCol col = stencil_col(g, x, y);
  float cy[kGridZ + 2], cn[kGridZ + 2];  // Y and n at z0 - 1 .. z0 + kGridZ
#pragma unroll
  for (int k = 0; k < kGridZ + 2; k++) {
    cy[k] = col.at(ys, z0 - 1 + k);
    cn[k] = col.at(n, z0 - 1 + k);
  }
#pragma unroll
  for (int k = 0; k < kGridZ; k++) {
    const int z = z0 + k;
    if (z >= g.gl) break;
    const int v = col.base + z * col.plane;
    // In-plane neighbours (x - 1, x + 1, y - 1, y + 1), 0 outside the grid.
    const int o[4] = {-1, 1, -col.gx, col.gx};
    const bool in[4] = {col.xl, col.xr, col.yl, col.yr};
    float yn[4], nn[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      yn[j] = in[j] ? ys[v + o[j]] : 0.f;
      nn[j] = in[j] ? n[v + o[j]] : 0.f;
    }
    const float dif = yn[0] * nn[0] + yn[1] * nn[1] + yn[2] * nn[2] + yn[3] * nn[3] +
                      cy[k] * cn[k] + cy[k + 2] * cn[k + 2];
    const float q = (b[v] + lambda * cn[k + 1] * dif) * invA[v];
    const float yv = cy[k + 1];
    const float grad = yv - q;
    // The reference compares the float |grad| with the double 1.345
    // (src/hfbs.cu:233); 1.345f lies just above 1.345 and the float below it
    // just under, so ">= 1.345f" selects exactly the same floats.
    const float gc = fabsf(grad) >= 1.345f ? copysignf(1.f, grad) : grad;
    const float hn = beta * hb[v] + gc;
    hb[v] = hn;
    yd[v] = yv - alpha * hn;
  }
}

// src/hfbs.cu:244-330 (kSliceTrilinear)
__global__ void k_hf_slice(const uchar4* __restrict__ guide, const float* __restrict__ S,
                           const float* __restrict__ Y, int W, int H, float sx, float sy,
                           float sl, Grid g, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W || y >= H) return;
  const size_t pi = (size_t)y * W + x;
  const float fx = gcoord(x, sx, g.gx), fy = gcoord(y, sy, g.gy);
  const float fz = gcoord(guide[pi].x, sl, g.gl);
  const int x0 = (int)floorf(fx), y0 = (int)floorf(fy), z0 = (int)floorf(fz);
  const float ax = fx - x0, ay = fy - y0, az = fz - z0;
  // The eight corners: 32-bit voxel offsets (grids stay below 2^31 voxels);
  // x0, y0, z0 lie inside the grid (gcoord clamps), a corner one above may
  // not, and reads outside the grid are 0.
  const int ox = 1, oy = g.gx, oz = g.gx * g.gy;
  const int o = (z0 * g.gy + y0) * g.gx + x0;
  const bool bx = x0 + 1 < g.gx, by = y0 + 1 < g.gy, bz = z0 + 1 < g.gl;
  float sv[2][2][2], yv[2][2][2];  // [dz][dy][dx]
#pragma unroll
  for (int dz = 0; dz < 2; dz++)
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
      for (int dx = 0; dx < 2; dx++) {
        const bool in = (!dx || bx) && (!dy || by) && (!dz || bz);
        const int v = o + dx * ox + dy * oy + dz * oz;
        sv[dz][dy][dx] = in ? S[v] : 0.f;
        yv[dz][dy][dx] = in ? Y[v] : 0.f;
      }
#define HSY(DX, DY, DZ) (sv[DZ][DY][DX] * yv[DZ][DY][DX])
#define HSS(DX, DY, DZ) sv[DZ][DY][DX]
  const float c00 = HSY(0, 0, 0) * (1 - ax) + HSY(1, 0, 0) * ax;
  const float c01 = HSY(0, 0, 1) * (1 - ax) + HSY(1, 0, 1) * ax;
  const float c10 = HSY(0, 1, 0) * (1 - ax) + HSY(1, 1, 0) * ax;
  const float c11 = HSY(0, 1, 1) * (1 - ax) + HSY(1, 1, 1) * ax;
  const float c0 = c00 * (1 - ay) + c10 * ay;
  const float c1 = c01 * (1 - ay) + c11 * ay;
  const float m00 = HSS(0, 0, 0) * (1 - ax) + HSS(1, 0, 0) * ax;
  const float m01 = HSS(0, 0, 1) * (1 - ax) + HSS(1, 0, 1) * ax;
  const float m10 = HSS(0, 1, 0) * (1 - ax) + HSS(1, 1, 0) * ax;
  const float m11 = HSS(0, 1, 1) * (1 - ax) + HSS(1, 1, 1) * ax;
  const float m0 = m00 * (1 - ay) + m10 * ay;
  const float m1 = m01 * (1 - ay) + m11 * ay;
#undef HSY
#undef HSS
  out[pi] = (c0 * (1 - az) + c1 * az) / (m0 * (1 - az) + m1 * az);
}

}  // namespace dtsamd

using namespace dtsamd;

struct hfbs_handle {
  int max_w = 0, max_h = 0, max_gx = 0, max_gy = 0, max_gl = 0, w = 0, h = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  uchar4* color = nullptr;
  float *target = nullptr, *conf = nullptr, *out = nullptr;
  float *b = nullptr, *cg = nullptr, *S = nullptr, *n0 = nullptr, *n1 = nullptr;
  float *invA = nullptr, *y0 = nullptr, *y1 = nullptr, *hb = nullptr, *t1 = nullptr,
        *t2 = nullptr;
  float* nfinal = nullptr;  // n after bistochastization (n0 or n1)
  float* yfinal = nullptr;  // z after the solve (y0 or y1)
  hfbs_grid_params p{};
  Grid g{1, 1, 1};
  bool graph_mode = false;
  hipGraphExec_t gexec = nullptr;
  std::vector<char> gkey;
  bool last_fell_back = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

namespace {

size_t hplane(const hfbs_handle* h) { return (size_t)h->w * h->h; }
size_t gvox(const Grid& g) { return (size_t)g.gx * g.gy * g.gl; }

int hf_fill(hfbs_handle* h, float* p, size_t n, float v) {
  if (!n) return DTS_OK;
  k_hf_fill<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(p, n, v);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int hf_copy(hfbs_handle* h, const float* s, float* d, size_t n) {
  if (!n || s == d) return DTS_OK;
  k_hf_copy<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(s, d, n);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

Grid grid_for(const hfbs_handle* h) {
  // src/hfbs.cu:589-591: x/y in fp32, luma via double.
  Grid g;
  g.gx = (int)std::ceil((float)h->w / h->p.sigma_x);
  g.gy = (int)std::ceil((float)h->h / h->p.sigma_y);
  g.gl = (int)std::ceil(255.0 / (double)h->p.sigma_l);
  return g;
}

// Sweeps along x of src into dst and, when src2 is set, of src2 into dst2.
int sweep_x(hfbs_handle* h, const float* src, float* dst, float beta,
            const float* src2 = nullptr, float* dst2 = nullptr) {
  const Grid& g = h->g;
  const int nl = g.gy * g.gl;
  const dim3 grid((unsigned)div_up(nl, 4), src2 ? 2u : 1u);
  if (beta > 0.f && g.gx <= 64 * 4)
    k_hf_sweep_x<4><<<grid, 256, 0, h->stream>>>(src, dst, src2, dst2, g.gx, nl, beta);
  else if (beta > 0.f && g.gx <= 64 * 8)
    k_hf_sweep_x<8><<<grid, 256, 0, h->stream>>>(src, dst, src2, dst2, g.gx, nl, beta);
  else if (beta > 0.f && g.gx <= 64 * 16)
    k_hf_sweep_x<16><<<grid, 256, 0, h->stream>>>(src, dst, src2, dst2, g.gx, nl, beta);
  else {
    k_hf_sweep<<<div_up(nl, 64), 64, 0, h->stream>>>(src, dst, g, 0, beta);
    if (src2) k_hf_sweep<<<div_up(nl, 64), 64, 0, h->stream>>>(src2, dst2, g, 0, beta);
  }
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

template <bool INITZ>
void sweep_yz_launch(hfbs_handle* h, int ch, dim3 grid, const float* src, float* dst,
                     const float* src2, float* dst2, int axis, float beta) {
  const Grid& g = h->g;
  if (ch <= 8)
    k_hf_sweep_yz<8, INITZ><<<grid, 256, 0, h->stream>>>(src, dst, src2, dst2, g, axis, beta);
  else if (ch <= 16)
    k_hf_sweep_yz<16, INITZ><<<grid, 256, 0, h->stream>>>(src, dst, src2, dst2, g, axis, beta);
  else if (ch <= 40)
    k_hf_sweep_yz<40, INITZ><<<grid, 256, 0, h->stream>>>(src, dst, src2, dst2, g, axis, beta);
  else
    k_hf_sweep_yz<64, INITZ><<<grid, 256, 0, h->stream>>>(src, dst, src2, dst2, g, axis, beta);
}

// Sweeps along y (axis 1) or z (axis 2) of src into dst and, when src2 is
// set, of src2 into dst2.
int sweep_yz(hfbs_handle* h, const float* src, float* dst, int axis, float beta,
             const float* src2 = nullptr, float* dst2 = nullptr) {
  const Grid& g = h->g;
  const int n = axis == 1 ? g.gy : g.gl;
  const int others = axis == 1 ? g.gl : g.gy;
  const dim3 grid((unsigned)(div_up(g.gx, 64) * others), src2 ? 2u : 1u);
  const int ch = (n + 3) / 4;
  if (beta > 0.f && ch <= 64) {
    sweep_yz_launch<false>(h, ch, grid, src, dst, src2, dst2, axis, beta);
  } else {
    const int nl = axis == 1 ? g.gx * g.gl : g.gx * g.gy;
    k_hf_sweep<<<div_up(nl, 64), 64, 0, h->stream>>>(src, dst, g, axis, beta);
    if (src2) k_hf_sweep<<<div_up(nl, 64), 64, 0, h->stream>>>(src2, dst2, g, axis, beta);
  }
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

// The z sweeps of the XY-swept confidence c (into blur_c) and values b, and
// the initial solution z0 = blur(b) / (blur(c) + 1e-7) (src/hfbs.cu:335-353);
// tmp: scratch grid for the general path.
int sweep_z_initz(hfbs_handle* h, const float* c, float* blur_c, const float* b, float* z0,
                  float* tmp, float beta) {
  const Grid& g = h->g;
  const int ch = (g.gl + 3) / 4;
  if (beta > 0.f && ch <= 64) {
    sweep_yz_launch<true>(h, ch, dim3((unsigned)(div_up(g.gx, 64) * g.gy)), c, blur_c, b, z0,
                          2, beta);
    DTS_LAUNCH_CHECK();
    return DTS_OK;
  }
  int rc;
  if ((rc = sweep_yz(h, c, blur_c, 2, beta, b, tmp))) return rc;
  const size_t nv = gvox(g);
  k_hf_initz<<<(unsigned)((nv + 255) / 256), 256, 0, h->stream>>>(tmp, blur_c, z0, nv);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int run_hfbs(hfbs_handle* h) {
  const Grid& g = h->g;
  const size_t nv = gvox(g);
  const hfbs_grid_params& p = h->p;
  int rc;
  // Splat.
  const int ntx = div_up(g.gx, kCellTile), nty = div_up(g.gy, kCellTile);
  const size_t lds = sizeof(float) * 3 * (size_t)g.gl * kCellTile * kCellTile;
  k_hf_splat<<<ntx * nty, 256, lds, h->stream>>>(h->color, h->target, h->conf, h->w, h->h,
                                                 p.sigma_x, p.sigma_y, p.sigma_l, g, h->b,
                                                 h->cg, h->S, h->n0, h->hb);
  DTS_LAUNCH_CHECK();
  // Bistochastization from n = 1 (stored by the splat); the last step also
  // forms the inverse diagonal.
  const unsigned gg = (unsigned)(div_up(g.gx * g.gy, kStencilCols) * div_up(g.gl, kGridZ));
  const dim3 gb(kStencilCols);
  float* ns = h->n0;
  float* nd = h->n1;
  for (int it = 0; it < p.bistoch_iter_max; it++) {
    const bool last = it == p.bistoch_iter_max - 1;
    k_hf_bistoch<<<gg, gb, 0, h->stream>>>(h->S, ns, nd, g, h->cg, last ? h->invA : nullptr,
                                          p.lambda);
    DTS_LAUNCH_CHECK();
    std::swap(ns, nd);
  }
  h->nfinal = ns;
  if (p.bistoch_iter_max <= 0) {
    k_hf_invdiag<<<(unsigned)((nv + 255) / 256), 256, 0, h->stream>>>(h->S, ns, h->cg,
                                                                     h->invA, nv, p.lambda);
    DTS_LAUNCH_CHECK();
  }
  // Laplacian initialisation, both grids per launch: x (c_g -> t1, b -> t2),
  // y (t1 -> c_g, t2 -> y1), then z with the initial solution: blur(c_g) ->
  // t1, blur(b) / (blur(c_g) + 1e-7) -> y0 (b kept).
  if ((rc = sweep_x(h, h->cg, h->t1, p.neg_exp_beta, h->b, h->t2))) return rc;
  if ((rc = sweep_yz(h, h->t1, h->cg, 1, p.neg_exp_beta, h->t2, h->y1))) return rc;
  if ((rc = sweep_z_initz(h, h->cg, h->t1, h->y1, h->y0, h->t2, p.neg_exp_beta))) return rc;
  // Heavy-ball solve.
  float* ys = h->y0;
  float* yd = h->y1;
  for (int it = 0; it < p.optim_iter_max; it++) {
    k_hf_solve<<<gg, gb, 0, h->stream>>>(ns, h->b, h->invA, ys, yd, h->hb, g, p.lambda,
                                        p.alpha, p.beta);
    DTS_LAUNCH_CHECK();
    std::swap(ys, yd);
  }
  h->yfinal = ys;
  k_hf_slice<<<dim3(div_up(h->w, 256), h->h), 256, 0, h->stream>>>(
      h->color, h->S, ys, h->w, h->h, p.sigma_x, p.sigma_y, p.sigma_l, g, h->out);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

void hf_drop_graph(hfbs_handle* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  h->gexec = nullptr;
  h->gkey.clear();
}

}  // namespace

extern "C" {

int hfbs_default_params(hfbs_grid_params* p) {
  if (!p) return DTS_ERR_INVALID_ARG;
  // src/hfbs.h:20-52
  *p = hfbs_grid_params{8.0f, 8.0f, 4.0f, 8.0f, 1.0f, 0.9f, 20, 140, 0.9f, 1, 1, 1};
  return DTS_OK;
}

int hfbs_create(int max_w, int max_h, int max_gx, int max_gy, int max_gl, hfbs_handle** out) {
  if (!out) return DTS_ERR_INVALID_ARG;
  *out = nullptr;
  if (max_w <= 0 || max_h <= 0 || max_gx <= 0 || max_gy <= 0 || max_gl <= 0) {
    set_error("invalid HFBS max dims");
    return DTS_ERR_INVALID_DIMS;
  }
  // The stencil kernels index voxels with 32-bit offsets.
  if ((size_t)max_gx * max_gy * max_gl > (size_t)INT32_MAX) {
    set_error("HFBS grid of %d x %d x %d exceeds 2^31 voxels", max_gx, max_gy, max_gl);
    return DTS_ERR_INVALID_DIMS;
  }
  hfbs_handle* h = new hfbs_handle();
  h->max_w = h->w = max_w;
  h->max_h = h->h = max_h;
  h->max_gx = max_gx;
  h->max_gy = max_gy;
  h->max_gl = max_gl;
  hfbs_default_params(&h->p);
  const size_t n = (size_t)max_w * max_h;
  const size_t nv = (size_t)max_gx * max_gy * max_gl;
  hipError_t e;
  auto fail = [&](hipError_t err) {
    set_error("HFBS allocation failed: %s", hipGetErrorString(err));
    hfbs_destroy(h);
    return DTS_ERR_HIP;
  };
  if ((e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking))) return fail(e);
  h->stream = h->own_stream;
#define HF_ALLOC(ptr, bytes) \
  if ((e = hipMalloc((void**)&(ptr), (bytes)))) return fail(e);
  HF_ALLOC(h->color, n * sizeof(uchar4));
  HF_ALLOC(h->target, n * 4);
  HF_ALLOC(h->conf, n * 4);
  HF_ALLOC(h->out, n * 4);
  float** grids[] = {&h->b, &h->cg, &h->S, &h->n0, &h->n1, &h->invA,
                     &h->y0, &h->y1, &h->hb, &h->t1, &h->t2};
  for (float** gp : grids) HF_ALLOC(*gp, nv * 4);
#undef HF_ALLOC
  if ((e = hipEventCreate(&h->e0))) return fail(e);
  if ((e = hipEventCreate(&h->e1))) return fail(e);
  h->nfinal = h->n0;
  h->yfinal = h->y0;
  const int rc = hfbs_clear_all(h);
  if (rc) {
    hfbs_destroy(h);
    return rc;
  }
  *out = h;
  return DTS_OK;
}

int hfbs_destroy(hfbs_handle* h) {
  if (!h) return DTS_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  hf_drop_graph(h);
  void* ptrs[] = {h->color, h->target, h->conf, h->out, h->b, h->cg, h->S, h->n0,
                  h->n1,    h->invA,   h->y0,   h->y1,  h->hb, h->t1, h->t2};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
  return DTS_OK;
}

int hfbs_get_stream(hfbs_handle* h, void** s) {
  if (!h || !s) return DTS_ERR_INVALID_ARG;
  *s = (void*)h->stream;
  return DTS_OK;
}

int hfbs_set_stream(hfbs_handle* h, void* s) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  hf_drop_graph(h);
  h->stream = s ? (hipStream_t)s : h->own_stream;
  return DTS_OK;
}

int hfbs_synchronize(hfbs_handle* h) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int hfbs_set_image_dim(hfbs_handle* h, int w, int hh) {
  if (!h) return DTS_ERR_INVALID_ARG;
  // src/hfbs.cu:527-538 rejects only sizes above the allocation.
  if (w <= 0 || hh <= 0 || w > h->max_w || hh > h->max_h) {
    set_error("Image size is more than the reserved memory: %dx%d", w, hh);
    return DTS_ERR_INVALID_DIMS;
  }
  h->w = w;
  h->h = hh;
  return DTS_OK;
}

int hfbs_set_params(hfbs_handle* h, const hfbs_grid_params* p) {
  if (!h || !p) return DTS_ERR_INVALID_ARG;
  if (!(p->sigma_x > 0.f) || !(p->sigma_y > 0.f) || !(p->sigma_l > 0.f) ||
      p->bistoch_iter_max < 0 || p->optim_iter_max < 0) {
    set_error("invalid HFBS params");
    return DTS_ERR_INVALID_ARG;
  }
  h->p = *p;
  return DTS_OK;
}

static int hfbs_init_impl(hfbs_handle* h, int cs, const void* rgba, const float* t,
                          const float* c, bool device) {
  if (!h || !rgba || !t) return DTS_ERR_INVALID_ARG;
  if (cs < DTS_COLOR_RGB || cs > DTS_COLOR_YYY) {
    set_error("unknown colour space %d", cs);
    return DTS_ERR_INVALID_ARG;
  }
  const size_t n = hplane(h);
  const hipMemcpyKind k = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  DTS_HIP_CHECK(hipMemcpyAsync(h->color, rgba, n * sizeof(uchar4), k, h->stream));
  if (cs != DTS_COLOR_RGB) {
    k_hf_convert<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(h->color, (int)n, cs);
    DTS_LAUNCH_CHECK();
  }
  DTS_HIP_CHECK(hipMemcpyAsync(h->target, t, n * 4, k, h->stream));
  if (c) {
    DTS_HIP_CHECK(hipMemcpyAsync(h->conf, c, n * 4, k, h->stream));
  } else {
    const int rc = hf_fill(h, h->conf, (size_t)h->max_w * h->max_h, 1.0f);
    if (rc) return rc;
  }
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int hfbs_init_frame(hfbs_handle* h, int cs, const void* rgba, const float* t, const float* c) {
  return hfbs_init_impl(h, cs, rgba, t, c, false);
}

int hfbs_init_frame_device(hfbs_handle* h, int cs, const void* rgba, const float* t,
                           const float* c) {
  return hfbs_init_impl(h, cs, rgba, t, c, true);
}

int hfbs_clear_all(hfbs_handle* h) {
  if (!h) return DTS_ERR_INVALID_ARG;
  const size_t n = (size_t)h->max_w * h->max_h;
  const size_t nv = (size_t)h->max_gx * h->max_gy * h->max_gl;
  DTS_HIP_CHECK(hipMemsetAsync(h->color, 0, n * sizeof(uchar4), h->stream));
  float* planes[] = {h->target, h->conf, h->out};
  for (float* p : planes) DTS_HIP_CHECK(hipMemsetAsync(p, 0, n * 4, h->stream));
  float* grids[] = {h->b, h->cg, h->S, h->n0, h->n1, h->invA, h->y0, h->y1, h->hb, h->t1, h->t2};
  for (float* p : grids) DTS_HIP_CHECK(hipMemsetAsync(p, 0, nv * 4, h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int hfbs_optimize(hfbs_handle* h, int* fell_back) {
  const TraceRange trace_("hfbs_optimize");
  if (!h) return DTS_ERR_INVALID_ARG;
  h->g = grid_for(h);
  const Grid& g = h->g;
  const bool too_big = g.gx > h->max_gx || g.gy > h->max_gy || g.gl > h->max_gl ||
                       g.gx <= 0 || g.gy <= 0 || g.gl <= 0;
  h->last_fell_back = too_big;
  if (fell_back) *fell_back = too_big ? 1 : 0;
  if (too_big) {
    // src/hfbs.cu:594-611: the output equals the target.
    fprintf(stderr,
            "Trying to use too large a bilateral space. Increase the blur sigmas or allocate "
            "more memory via constructor.\n");
    return hf_copy(h, h->target, h->out, hplane(h));
  }
  if (3 * (size_t)g.gl * kCellTile * kCellTile * sizeof(float) > 64 * 1024) {
    set_error("grid_l %d too large for the splat tile", g.gl);
    return DTS_ERR_UNSUPPORTED;
  }
  if (!h->graph_mode) return run_hfbs(h);
  std::vector<char> key(sizeof(int) * 2 + sizeof(hfbs_grid_params));
  std::memcpy(key.data(), &h->w, sizeof(int));
  std::memcpy(key.data() + sizeof(int), &h->h, sizeof(int));
  std::memcpy(key.data() + 2 * sizeof(int), &h->p, sizeof(hfbs_grid_params));
  if (!h->gexec || key != h->gkey) {
    hf_drop_graph(h);
    DTS_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = run_hfbs(h);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return DTS_ERR_HIP;
    }
    const hipError_t e2 = hipGraphInstantiate(&h->gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e2 != hipSuccess) {
      h->gexec = nullptr;
      set_error("graph instantiate failed: %s", hipGetErrorString(e2));
      return DTS_ERR_HIP;
    }
    h->gkey = key;
  }
  DTS_HIP_CHECK(hipGraphLaunch(h->gexec, h->stream));
  return DTS_OK;
}

int hfbs_download(hfbs_handle* h, int type, float* out) {
  if (!h || !out) return DTS_ERR_INVALID_ARG;
  const float* src = nullptr;
  switch (type) {
    case DTS_IMAGE_TARGET: src = h->target; break;
    case DTS_IMAGE_CONFIDENCE: src = h->conf; break;
    case DTS_IMAGE_OPTIMIZED_QUANTITY: src = h->out; break;
    default:
      set_error("unsupported image type %d", type);
      return DTS_ERR_INVALID_ARG;
  }
  DTS_HIP_CHECK(hipMemcpyAsync(out, src, hplane(h) * 4, hipMemcpyDeviceToHost, h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int hfbs_result_device(hfbs_handle* h, const float** out) {
  if (!h || !out) return DTS_ERR_INVALID_ARG;
  *out = h->out;
  return DTS_OK;
}

int hfbs_debug_grid_dims(hfbs_handle* h, int* gx, int* gy, int* gl) {
  if (!h || !gx || !gy || !gl) return DTS_ERR_INVALID_ARG;
  const Grid g = grid_for(h);
  *gx = g.gx;
  *gy = g.gy;
  *gl = g.gl;
  return DTS_OK;
}

int hfbs_debug_grid(hfbs_handle* h, int which, float* out) {
  if (!h || !out) return DTS_ERR_INVALID_ARG;
  const float* src = nullptr;
  switch (which) {
    case 0: src = h->b; break;
    case 1: src = h->t1; break;
    case 2: src = h->S; break;
    case 3: src = h->nfinal; break;
    case 4: src = h->invA; break;
    case 5: src = h->yfinal; break;
    default:
      set_error("unknown grid %d", which);
      return DTS_ERR_INVALID_ARG;
  }
  DTS_HIP_CHECK(hipMemcpyAsync(out, src, gvox(h->g) * 4, hipMemcpyDeviceToHost, h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int hfbs_set_graph_mode(hfbs_handle* h, int enable) {
  if (!h) return DTS_ERR_INVALID_ARG;
  h->graph_mode = enable != 0;
  if (!h->graph_mode) hf_drop_graph(h);
  return DTS_OK;
}

int hfbs_timer_start(hfbs_handle* h) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipEventRecord(h->e0, h->stream));
  return DTS_OK;
}

int hfbs_timer_stop(hfbs_handle* h, float* ms) {
  if (!h || !ms) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipEventRecord(h->e1, h->stream));
  DTS_HIP_CHECK(hipEventSynchronize(h->e1));
  DTS_HIP_CHECK(hipEventElapsedTime(ms, h->e0, h->e1));
  return DTS_OK;
}

}  // extern "C"
