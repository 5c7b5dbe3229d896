// Domain Transform Solver on gfx950: kernels and the dts_* C ABI.
//
// Data layout in HBM (all planes tightly packed at the current image size):
//   color, target, conf, vA, vB, boundsH, boundsV, ctH : row-major [y][x]
//   colorT, ctV                                        : transposed [x][y]
// One solver iteration is two kernel launches:
//   X pass  reads vA (+ target, conf, boundsH), writes vB;
//   Y pass  reads vB (+ boundsV), writes vA.
// The X pass gives each wavefront a 256-pixel row segment plus its window
// halo, forms the segment's prefix sum with DPP wavefront scans in LDS and
// evaluates the box mean and update per pixel; the Y pass gives each
// workgroup a 64-column tile plus its vertical halo and forms column prefix
// sums in LDS.  Every global access is a contiguous row segment.  The window
// bounds depend only on the guide and sigma, so they are found once per solve
// (exactly, on 32.32 fixed-point ct) instead of once per pixel per iteration.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include <hip/hip_ext.h>

#include "adjoint.h"
#include "dts_kernels.h"
#include "fused.h"
#include "rf.h"
#include "wide.h"

namespace dtsamd {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

const char* last_error() { return g_err.c_str(); }

__global__ void k_convert_color(uchar4* __restrict__ img, int n, int cs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) img[i] = convert_px(img[i], cs);
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_copy_f32(const float* __restrict__ s, float* __restrict__ d,
                           size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

__global__ void k_square_f32(const float* __restrict__ s, float* __restrict__ d,
                             size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float v = s[i];
    d[i] = v * v;
  }
}

// Horizontal differential as fp32 (Download(DIFFERENTIAL), stereo right dH).
__global__ void k_diff_h(const uchar4* __restrict__ color, int W, int H,
                         double q, float* __restrict__ dHf) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W || y >= H) return;
  const uchar4* row = color + (size_t)y * W;
  const uchar4 prv = x > 0 ? row[x - 1] : make_uchar4(0, 0, 0, 0);
  dHf[(size_t)y * W + x] = (float)dt_diff(row[x], prv, q);
}

// ------------------------------------------------------------ integrate
// ct_H: inclusive scan of the 32.32 fixed-point dH along each row (one
// workgroup per line); the same kernel scans the transposed guide for ct_V.
__global__ __launch_bounds__(256) void k_ct_rows(const uchar4* __restrict__ color, int W,
                                                 double q, int64_t* __restrict__ ct) {
  const size_t y = blockIdx.x;
  ct_line(color + y * W, W, q, ct + y * W);
}

// Both directions in one launch, one workgroup per line (ct_line_quad: W
// and H multiples of 4, at most 2048): lines [0, H) are the rows of the
// guide (into ct_H), lines [H, H + W) the rows of the transposed guide, i.e.
// the columns (into ct_V).
template <bool SMALL>
__global__ __launch_bounds__(256) void k_ct_quad(const uchar4* __restrict__ color,
                                                 const uchar4* __restrict__ colorT, int W, int H,
                                                 double qx, double qy, int64_t* __restrict__ ctH,
                                                 int64_t* __restrict__ ctV) {
  __shared__ int64_t wtot[4];
  const int id = (int)blockIdx.x;
  const bool rows = id < H;
  const int L = rows ? W : H;
  const uchar4* line = rows ? color + (size_t)id * W : colorT + (size_t)(id - H) * H;
  int64_t* out = rows ? ctH + (size_t)id * W : ctV + (size_t)(id - H) * H;
  const double q = rows ? qx : qy;
  if (L <= 1024)
    ct_line_quad<1, SMALL>(line, L, q, out, wtot);
  else
    ct_line_quad<2, SMALL>(line, L, q, out, wtot);
}

// Both directions in one launch, one wavefront per line: lines [0, H) are
// the rows of the guide (into ct_H), lines [H, H + W) the rows of the
// transposed guide, i.e. the columns (into ct_V).  Lines longer than 2048
// use k_ct_rows.
__global__ __launch_bounds__(256) void k_ct_both(const uchar4* __restrict__ color,
                                                 const uchar4* __restrict__ colorT, int W, int H,
                                                 double qx, double qy, int64_t* __restrict__ ctH,
                                                 int64_t* __restrict__ ctV) {
  const int lane = threadIdx.x & 63;
  const int id = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (id >= H + W) return;
  if (id < H)
    ct_line_w(color + (size_t)id * W, W, qx, ctH + (size_t)id * W, lane);
  else
    ct_line_w(colorT + (size_t)(id - H) * H, H, qy, ctV + (size_t)(id - H) * H, lane);
}

// Guide transpose [y][x] -> [x][y] through a 64x64 LDS tile, done once per
// frame at upload so the vertical differential and its scan run along
// contiguous lines with the same kernel as the horizontal ones.
__global__ __launch_bounds__(256) void k_transpose_u8x4(const uchar4* __restrict__ src,
                                                        int W, int H,
                                                        uchar4* __restrict__ dst) {
  __shared__ uint32_t tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 64;
  for (int k = ty; k < 64; k += 4) {
    const int x = x0 + tx, y = y0 + k;
    uint32_t v = 0;
    if (x < W && y < H) {
      const uchar4 c = src[(size_t)y * W + x];
      v = (uint32_t)c.x | ((uint32_t)c.y << 8) | ((uint32_t)c.z << 16) | ((uint32_t)c.w << 24);
    }
    tile[k][tx] = v;
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int x = x0 + k, y = y0 + tx;
    if (x < W && y < H) {
      const uint32_t v = tile[tx][k];
      dst[(size_t)x * H + y] = make_uchar4(v & 0xff, (v >> 8) & 0xff, (v >> 16) & 0xff, v >> 24);
    }
  }
}

// Horizontal bounds, searched in global memory (any extent): lines are the
// rows of ct_H, output row-major.
__global__ __launch_bounds__(256) void k_bounds_rows(const int64_t* __restrict__ ct,
                                                     int L, int64_t sfx, int E,
                                                     uint32_t* __restrict__ out) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const size_t line = blockIdx.y;
  if (p < L) out[line * L + p] = find_bounds(ct + line * L, L, p, sfx, E);
}

// Vertical bounds, searched in global memory (any extent): lines are the
// columns, contiguous in the transposed ct_V ([x][y]); the result is written
// row-major ([y][x]) through a 64x64 LDS tile.
__global__ __launch_bounds__(256) void k_bounds_cols(const int64_t* __restrict__ ctT,
                                                     int W, int H, int64_t sfx, int E,
                                                     uint32_t* __restrict__ out) {
  __shared__ uint32_t tile[64][65];  // [x][y]
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int y0 = blockIdx.x * 64, x0 = blockIdx.y * 64;
  for (int k = q; k < 64; k += 4) {
    const int x = x0 + k, y = y0 + lane;
    uint32_t b = 0;
    if (x < W && y < H) b = find_bounds(ctT + (size_t)x * H, H, y, sfx, E);
    tile[k][lane] = b;
  }
  __syncthreads();
  for (int k = q; k < 64; k += 4) {
    const int y = y0 + k, x = x0 + lane;
    if (x < W && y < H) out[(size_t)y * W + x] = tile[lane][k];
  }
}

template <bool kCols, bool NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_bounds_tile(const int64_t* __restrict__ ct, int L, int nlines, int64_t sfx, int E,
                   int stage_len, void* __restrict__ out, int* __restrict__ yext,
                   int ext_groups) {
  extern __shared__ int64_t st[];
  bounds_tile<kCols, NB>(ct, L, nlines, sfx, E, stage_len, out, yext, ext_groups,
                     blockIdx.x * kBPos, blockIdx.y * kBLines, st);
}

// Horizontal and vertical tiles of one bounds computation in one launch:
// workgroups [0, ntiles_h) are horizontal tiles (ntx_h per tile row), the
// rest vertical tiles (ntx_v per tile row).
struct BoundsPair {
  const int64_t* ctH;
  const int64_t* ctV;
  uint32_t* bH;
  uint32_t* bV;
  int* yext;
  int W, H, ext_groups;
  int64_t sfx_h, sfx_v;
  int E_h, E_v, stage_h, stage_v;
  int ntiles_h, ntx_h, ntx_v;
  int nb_h;  // narrow (2-byte) horizontal bounds plane; the vertical one is wide
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_bounds_both(BoundsPair b) {
  extern __shared__ int64_t st[];
  // Contiguous tile ranges per XCD: neighbouring tiles of a line share their
  // staged halo in one L2.
  const int id = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  if (id < b.ntiles_h) {
    const int p0 = (id % b.ntx_h) * kBPos, l0 = (id / b.ntx_h) * kBLines;
    if (b.nb_h)
      bounds_tile<false, true>(b.ctH, b.W, b.H, b.sfx_h, b.E_h, b.stage_h, b.bH, nullptr, 0, p0,
                               l0, st);
    else
      bounds_tile<false, false>(b.ctH, b.W, b.H, b.sfx_h, b.E_h, b.stage_h, b.bH, nullptr, 0, p0,
                                l0, st);
  } else {
    const int t = id - b.ntiles_h;
    const int p0 = (t % b.ntx_v) * kBPos, l0 = (t / b.ntx_v) * kBLines;
    bounds_tile<true, false>(b.ctV, b.H, b.W, b.sfx_v, b.E_v, b.stage_v, b.bV, b.yext,
                             b.ext_groups, p0, l0, st);
  }
}

// Step table of one q: lut[s] is the 32.32 step dt_diff gives for SSD s,
// converted as the ct scans convert it.
__global__ void k_step_lut(double q, int64_t* __restrict__ lut) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s < kStepLutLen) lut[s] = to_fx(dt_diff_ssd(s, q));
}

// Both bounds directions from the guide (bounds_tile_guide), tiles laid out
// as in k_bounds_both: lines are the guide's rows (horizontal) and the
// transposed guide's rows (vertical).
struct BoundsPairG {
  const uchar4* gH;  // [H][W]
  const uchar4* gV;  // [W][H]
  const int64_t* lutH;
  const int64_t* lutV;
  uint32_t* bH;
  uint32_t* bV;
  int* yext;
  int W, H, ext_groups;
  int64_t sfx_h, sfx_v;
  int E_h, E_v, sl_h, sl_v;  // search radii, staged line strides
  int ntiles_h, ntx_h, ntx_v;
  int nb_h;          // narrow horizontal plane
  int vec_h, vec_v;  // lines 16-byte aligned
};

// Eight wavefronts of 4 lines each per 32-line tile: 2 lines staged and 4
// searched at a time keep the kernel within 64 registers, so the four tiles
// per CU that LDS allows run 8 wavefronts per SIMD (1080p: 27.1 vs 30.9 us
// with four wavefronts of 8 lines).
constexpr int kBoundsGThreads = 512;

__global__ __launch_bounds__(kBoundsGThreads) __attribute__((amdgpu_waves_per_eu(8)))
void k_bounds_both_g(BoundsPairG b) {
  extern __shared__ int64_t st[];
  const int id = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  if (id < b.ntiles_h) {
    const int p0 = (id % b.ntx_h) * kBPos, l0 = (id / b.ntx_h) * kBLines;
    if (b.nb_h)
      bounds_tile_guide<false, true, 2, 4, 8>(b.gH, b.W, b.H, b.lutH, b.vec_h != 0, b.sfx_h, b.E_h, b.sl_h,
                                     b.bH, nullptr, 0, p0, l0, st);
    else
      bounds_tile_guide<false, false, 2, 4, 8>(b.gH, b.W, b.H, b.lutH, b.vec_h != 0, b.sfx_h, b.E_h,
                                      b.sl_h, b.bH, nullptr, 0, p0, l0, st);
  } else {
    const int t = id - b.ntiles_h;
    const int p0 = (t % b.ntx_v) * kBPos, l0 = (t / b.ntx_v) * kBLines;
    bounds_tile_guide<true, false, 2, 4, 8>(b.gV, b.H, b.W, b.lutV, b.vec_v != 0, b.sfx_v, b.E_v, b.sl_v,
                                   b.bV, b.yext, b.ext_groups, p0, l0, st);
  }
}

template <int MODE, bool OW, bool STEREO, int NCH, bool NB>
__global__ __launch_bounds__(256) void k_xpass(PassArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int G = (int)gridDim.x;
  const int r = xcd_remap((int)blockIdx.x, G);
  const int nseg = a.ntx * a.H;
  float* V = lds + wid * 2 * 64 * (NCH | 1);
  float* Q = V + 64 * (NCH | 1);
  int s = r * 4 + wid;
  if (s >= nseg) return;
  // Segment s is (line s / ntx, column tile s % ntx); the stride TW = 4 G is
  // split into whole lines and a column remainder on the host.
  int line = s / a.ntx, tx = s - line * a.ntx;
  auto advance = [&](int& ln, int& t) {
    t += a.tw_mod;
    ln += a.tw_div;
    if (t >= a.ntx) {
      t -= a.ntx;
      ln++;
    }
  };
  XBuf<NCH> A, B;
  xseg_load<MODE, OW, NCH, NB>(a, line, tx * kXSeg, lane, A);
  while (true) {
    int l1 = line, t1 = tx;
    advance(l1, t1);
    const bool more = l1 < a.H;
    if (more) xseg_load<MODE, OW, NCH, NB>(a, l1, t1 * kXSeg, lane, B);
    xseg_process<MODE, OW, STEREO, NCH, NB>(a, line, tx * kXSeg, lane, A, V, Q);
    if (!more) break;
    line = l1;
    tx = t1;
    advance(l1, t1);
    const bool more2 = l1 < a.H;
    if (more2) xseg_load<MODE, OW, NCH, NB>(a, l1, t1 * kXSeg, lane, A);
    xseg_process<MODE, OW, STEREO, NCH, NB>(a, line, tx * kXSeg, lane, B, V, Q);
    if (!more2) break;
    line = l1;
    tx = t1;
  }
}

// Wavefronts per single-frame fused tile: 16-row tiles (small frames, where
// each tile's serial chain of staged rows sets the time) take 8, which halves
// each wavefront's rows (320 x 240 depth super-resolution 72 -> 59 us per
// solve); 32-row tiles keep 4, whose smaller LDS keeps four tiles per CU
// resident (1080p: 12.1 vs 12.5 us per iteration with 8); 64-row tiles take
// 8 (two tiles per CU).
constexpr int fused_waves(int yr) { return yr == 32 ? 4 : 8; }

// One Optimize iteration per launch (fused.h): 64 x YR tiles; STEREO with
// the photometric term.
template <int MODE, int NCX, bool NBH, int YR, bool STEREO, bool OW, int NW>
__global__ __launch_bounds__(64 * NW) void k_xypass(PassArgs ax, PassArgs ay) {
  extern __shared__ float lds[];
  const unsigned r = (unsigned)xcd_remap((int)blockIdx.x, ay.ntx * ay.nty);
  xytile<MODE, NCX, NBH, false, YR, STEREO, NW, OW>(
      ax, ay, (int)(r % (unsigned)ay.ntx), (int)(r / (unsigned)ay.ntx), lds);
}

// RMAX > 0: register-staged tiles of kYRowsReg rows; 0: the LDS-loop tile
// of kYRows rows (large windows).  64-row register tiles halve the halo
// re-reads but leave half the blocks, which measured slower at 1080p
// (10.0 vs 8.7 us per pass).
constexpr int kYRowsReg = 32;

template <int MODE, int RMAX, bool NB>
__global__ __launch_bounds__(256) void k_ypass(PassArgs a) {
  extern __shared__ float lds[];
  const int r = xcd_remap((int)blockIdx.x, a.ntx * a.nty);
  if constexpr (RMAX > 0)
    ytile_reg<MODE, (RMAX > 0 ? RMAX : 16), NB, kYRowsReg>(a, r % a.ntx, r / a.ntx, lds);
  else
    ytile<MODE, NB>(a, r % a.ntx, r / a.ntx, lds);
}

// ------------------------------------------- wide windows (any extent)
// Windows wider than the LDS segments of k_xpass and k_ypass: inclusive
// prefix sums S of the pass input along every line, in double precision
// (wide.h), then one thread per pixel forms the window mean
// (S(u) - S(l)) / (u - l) and the update.

// Input of the horizontal pass at element i, after the overwrite (OWP).
template <bool OWP>
__device__ __forceinline__ float wide_in(const PassArgs& a, unsigned i) {
  float v = ld32(a.vin, i);
  if (OWP && ld32(a.c, i) > a.theta) v = ld32(a.t, i);
  return v;
}

// Element functors of the line prefixes (wide.h): the horizontal pass input
// after the overwrite, and a plain plane.
template <bool OWP>
struct WideInX {
  PassArgs a;
  __device__ float operator()(unsigned i) const { return wide_in<OWP>(a, i); }
};
struct PlaneIn {
  const float* p;
  __device__ float operator()(unsigned i) const { return ld32(p, i); }
};

// Horizontal update from the row prefixes (box filter, Optimize X half, stereo term).
template <int MODE, bool OW, bool STEREO, bool NB>
__global__ __launch_bounds__(256) void k_wide_x(PassArgs a, const double* __restrict__ S) {
  constexpr bool kData = (MODE == PM_X_L2 || MODE == PM_X_CH);
  const int x = (int)blockIdx.x * 256 + threadIdx.x, y = (int)blockIdx.y;
  if (x >= a.W) return;
  const unsigned rowoff = (unsigned)y * (unsigned)a.W, i = rowoff + (unsigned)x;
  const uint32_t b = ld_bounds<NB>(a.bnd, i);
  const int l = x - bounds_dl<NB>(b), u = x + bounds_du<NB>(b);  // l >= -1
  const double* Sr = S + rowoff;
  const double sum = Sr[u] - (l >= 0 ? Sr[l] : 0.0);
  const float Bm = (float)(sum / (double)(u - l));
  const float rn = __builtin_amdgcn_rcpf((float)(u - l));
  const float v = wide_in<OW && kData>(a, i);
  const float t = kData ? ld32(a.t, i) : 0.f, c = kData ? ld32(a.c, i) : 0.f;
  st32(a.out, i, x_update<MODE, OW, STEREO>(a, rowoff, x, v, Bm, rn, t, c, true));
}

// Vertical update from the column prefixes (box filter, Optimize Y half); vertical bounds
// planes of a single frame are wide.
template <int MODE>
__global__ __launch_bounds__(256) void k_wide_y(PassArgs a, const double* __restrict__ S) {
  const int x = (int)blockIdx.x * 256 + threadIdx.x, y = (int)blockIdx.y;
  if (x >= a.W) return;
  const unsigned i = (unsigned)y * (unsigned)a.W + (unsigned)x;
  const uint32_t b = ld_bounds<false>(a.bnd, i);
  const int l = y - bounds_dl<false>(b), u = y + bounds_du<false>(b);  // l >= -1
  const double sum = S[(size_t)u * a.W + x] - (l >= 0 ? S[(size_t)l * a.W + x] : 0.0);
  const float Bm = (float)(sum / (double)(u - l));
  const float v = ld32(a.vin, i);
  st32(a.out, i, MODE == PM_FILTER ? Bm : v + a.step * (-(a.lambda * (v - Bm))));
}

// (lower, upper) extent pairs of the vertical-pass tiles.
__global__ void k_fill_ext(int* __restrict__ ext, int n, int lo, int hi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    ext[2 * i] = lo;
    ext[2 * i + 1] = hi;
  }
}

// Calibration stream: out = a + b + c + d over n floats (float4 accesses),
// 20 bytes moved per element.  Gives the practical bandwidth of a plain
// streaming kernel of the same size as a solver pass, and the reference
// point for hardware byte counters.
__global__ __launch_bounds__(256) void k_stream5(const float4* __restrict__ a,
                                                 const float4* __restrict__ b,
                                                 const float4* __restrict__ c,
                                                 const float4* __restrict__ d,
                                                 float4* __restrict__ out, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (size_t)gridDim.x * blockDim.x) {
    const float4 x = a[i], y = b[i], z = c[i], w = d[i];
    out[i] = make_float4(x.x + y.x + z.x + w.x, x.y + y.y + z.y + w.y, x.z + y.z + z.z + w.z,
                         x.w + y.w + z.w + w.w);
  }
}

// ----------------------------------------------------------- confidence
// src/domain_transform_solver.cu:21-44
__global__ void k_confidence(const float* __restrict__ m2, float* __restrict__ conf,
                             int W, int H, float scale, int left_clear) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= W || y >= H) return;
  const size_t i = (size_t)y * W + x;
  float c = 0.f;
  if (x > left_clear) {
    const float m = conf[i];
    const float var = m2[i] - m * m;
    c = (var < 0.f || !isfinite(var)) ? 0.f : expf(-var * scale);
  }
  conf[i] = c;
}

}  // namespace dtsamd

using namespace dtsamd;

struct dts_handle {
  int max_w = 0, max_h = 0, w = 0, h = 0, flags = 0;
  hipStream_t own_stream = nullptr, stream = nullptr;
  uchar4 *color = nullptr, *colorT = nullptr, *rcolor = nullptr;
  float *rdHf = nullptr, *target = nullptr, *conf = nullptr;
  float *vA = nullptr, *vB = nullptr, *csq = nullptr, *csqb = nullptr;
  int64_t *ctH = nullptr, *ctV = nullptr;  // ctV is transposed [x][y]
  uint32_t *bH = nullptr, *bV = nullptr;  // bounds planes (4 bytes per pixel allocated)
  bool bH_narrow = false, bV_narrow = false;  // planes hold 2-byte elements
  double *wide_S = nullptr, *wide_tot = nullptr;  // wide-window prefixes (on first use)
  int* yext = nullptr;  // per vertical-pass tile window extents
  int num_cus = 256;
  dts_filter_params fp{64.f, 64.f, 0.25f};  // solver filter_params_
  dts_filter_params dp{64.f, 64.f, 0.25f};  // params of the next integration
  // ct of the guide for params ctp.  Integration is lazy: ct_lazy means the
  // planes were not formed (an Optimize finds its bounds from the guide and
  // the step tables); whatever reads the planes forms them first, and so
  // does a change of the guide or the dims while ct_lazy is set.
  dts_filter_params ctp{64.f, 64.f, 0.25f};
  bool ct_lazy = false;
  bool ct_of_guide = false;  // the ct (formed or lazy) is that of the current guide
  // Step tables (kStepLutLen int64 each) for q = sigma_x / sigma_r and
  // sigma_y / sigma_r; lut_q[i] is the q table i holds (NaN: none).
  int64_t* lut = nullptr;
  double lut_q[2] = {NAN, NAN};
  bool capturing = false;  // inside a graph capture
  bool fp_set = false;
  float doff = 0.f;
  bool graph_mode = false;
  bool fuse = true;  // one launch per Optimize iteration where it applies
  hipGraphExec_t gexec = nullptr;
  std::vector<char> gkey;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  bool prof = false;
  std::vector<hipEvent_t> xev, yev, xyev;  // pairs: X, Y and fused-iteration launches
  // Differentiable solve: tape of the last checkpointed Optimize.
  float* tape = nullptr;  // (M - 1) checkpoints + C states, max_plane floats each
  size_t tape_planes = 0;
  int tape_K = 0, tape_C = 0;
  bool tape_valid = false;
  dts_optimize_params tape_p{};
  float tape_ow = -1.f;
  int2 *abH = nullptr, *abV = nullptr;  // per-line collapse limits (rows, columns)
  float *adj_g = nullptr, *adj_gb = nullptr, *adj_tb = nullptr, *adj_cb = nullptr;
  // RF-DT: kRfWork work planes, then the tape of the last dts_rf_filter that
  // kept one (the input of each of its 2N passes).
  float* rf_work = nullptr;
  float* rf_tape = nullptr;
  size_t rf_tape_planes = 0;
  bool rf_valid = false;
  int rf_n = 0;
  dts_filter_params rf_p{};
};

namespace {

size_t plane(const dts_handle* h) { return (size_t)h->w * h->h; }
size_t max_plane(const dts_handle* h) { return (size_t)h->max_w * h->max_h; }

int64_t sigma_fx(float s) {
  if (!(s < 1073741824.0f)) return INT64_MAX;
  return (int64_t)std::floor((double)s * 4294967296.0);
}

// Radius bound of the window search along a line of length L.
int sigma_extent(float s, int L) {
  if (!(s >= 0.f)) return 0;  // NaN: the window collapses
  if (s >= (float)L) return L;
  return (int)std::floor(s);
}

int launch_fill(dts_handle* h, float* p, size_t n, float v) {
  if (n == 0) return DTS_OK;
  k_fill_f32<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(p, n, v);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int launch_copy(dts_handle* h, const float* s, float* d, size_t n) {
  if (n == 0 || s == d) return DTS_OK;
  k_copy_f32<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(s, d, n);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int check_params(const dts_filter_params* p) {
  if (!p) {
    set_error("null filter params");
    return DTS_ERR_INVALID_ARG;
  }
  return DTS_OK;
}

// ct_H and ct_V of the current guide for the given filter params, into the
// planes.
int launch_ct_planes(dts_handle* h, const dts_filter_params* p) {
  // src/domain_transform_filter.cu:96-104: the sigma ratios are fp32.
  const float qx = p->sigma_x / p->sigma_r;
  const float qy = p->sigma_y / p->sigma_r;
  const bool quad = h->w % 4 == 0 && h->h % 4 == 0;
  if (quad && h->w <= 2048 && h->h <= 2048) {
    const bool small = std::fabs((double)qx) <= kFxSmallMaxQ && std::fabs((double)qy) <= kFxSmallMaxQ;
    if (small)
      k_ct_quad<true><<<h->h + h->w, 256, 0, h->stream>>>(h->color, h->colorT, h->w, h->h,
                                                          (double)qx, (double)qy, h->ctH, h->ctV);
    else
      k_ct_quad<false><<<h->h + h->w, 256, 0, h->stream>>>(h->color, h->colorT, h->w, h->h,
                                                           (double)qx, (double)qy, h->ctH, h->ctV);
    DTS_LAUNCH_CHECK();
  } else if (h->w <= 2048 && h->h <= 2048) {
    k_ct_both<<<div_up(h->h + h->w, 4), 256, 0, h->stream>>>(
        h->color, h->colorT, h->w, h->h, (double)qx, (double)qy, h->ctH, h->ctV);
    DTS_LAUNCH_CHECK();
  } else {
    k_ct_rows<<<h->h, 256, 0, h->stream>>>(h->color, h->w, (double)qx, h->ctH);
    DTS_LAUNCH_CHECK();
    k_ct_rows<<<h->w, 256, 0, h->stream>>>(h->colorT, h->h, (double)qy, h->ctV);
    DTS_LAUNCH_CHECK();
  }
  return DTS_OK;
}

void drop_graph(dts_handle* h);

// ct planes for params p now (Filter, ComputeConfidence): p also becomes the
// params of the next integration, as the reference's Filter leaves them.
int launch_ct(dts_handle* h, const dts_filter_params* p) {
  h->dp = *p;
  h->ctp = *p;
  h->ct_lazy = false;
  h->ct_of_guide = true;
  return launch_ct_planes(h, p);
}

// Integration with params p, formed when something reads the planes.
void set_ct_lazy(dts_handle* h, const dts_filter_params* p) {
  h->dp = *p;
  h->ctp = *p;
  h->ct_lazy = true;
  h->ct_of_guide = true;
}

// The planes of a lazy integration (params ctp; dp is left as it is).
int ensure_ct(dts_handle* h) {
  if (!h->ct_lazy) return DTS_OK;
  h->ct_lazy = false;
  return launch_ct_planes(h, &h->ctp);
}

// Step tables for the sigma ratios of p: table 0 for q_x, table 1 for q_y
// (table 0 again when they are equal).  Rebuilding a table outside a capture
// drops the cached graph, which may read the old one; graph-mode calls build
// their tables before capturing (lut_prepare), so a replay never rebuilds
// them.
int ensure_lut_for(dts_handle* h, const dts_filter_params* p, const int64_t** lh,
                   const int64_t** lv) {
  if (!h->lut) {
    set_error("no step tables on this handle");
    return DTS_ERR_STATE;
  }
  const double q[2] = {(double)(p->sigma_x / p->sigma_r), (double)(p->sigma_y / p->sigma_r)};
  const int nt = q[1] == q[0] ? 1 : 2;
  for (int i = 0; i < nt; i++) {
    if (h->lut_q[i] == q[i]) continue;  // NaN never matches
    k_step_lut<<<div_up(kStepLutLen, 256), 256, 0, h->stream>>>(
        q[i], h->lut + (size_t)i * kStepLutLen);
    DTS_LAUNCH_CHECK();
    h->lut_q[i] = q[i];
    if (!h->capturing) drop_graph(h);
  }
  *lh = h->lut;
  *lv = h->lut + (size_t)(nt - 1) * kStepLutLen;
  return DTS_OK;
}

int ensure_lut(dts_handle* h, const int64_t** lh, const int64_t** lv) {
  return ensure_lut_for(h, &h->ctp, lh, lv);
}

// Before a graph capture: the step tables the captured bounds will read.
int lut_prepare(dts_handle* h, const dts_filter_params* p) {
  if (!h->lut) return DTS_OK;
  const int64_t *lh, *lv;
  return ensure_lut_for(h, p, &lh, &lv);
}

// Single frames keep vertical bounds planes wide: with the narrow form the
// vertical pass measured slower on cache-resident frames (8.5 vs 8.0 us per
// pass at 1080p), while the horizontal pass gains a little.  The batched
// solve, which streams from HBM, uses narrow planes in both directions.
int launch_bounds(dts_handle* h, float sigma, bool vertical) {
  h->tape_valid = false;  // the tape refers to the Optimize bounds
  if (sigma < 0.f) {
    set_error("sigma must be >= 0 (got %g)", (double)sigma);
    return DTS_ERR_INVALID_ARG;
  }
  int rc;
  if ((rc = ensure_ct(h))) return rc;
  const int64_t sfx = sigma_fx(sigma);
  const int L = vertical ? h->h : h->w, nlines = vertical ? h->w : h->h;
  const int E = sigma_extent(sigma, L);
  const int stage_len = std::min(L, kBPos + 2 * E + 2);
  const size_t lds = std::max(kBoundsResBytes, sizeof(int64_t) * (size_t)kBLines * stage_len);
  const bool tiled = lds <= kBoundsLdsMax;
  const int groups = div_up(h->w, 32);
  if (vertical && !tiled) {
    // Tiled bounds write the actual extents of every (band, group); the
    // global-memory search leaves the caps.
    const int n = groups * div_up(h->h, 32);
    k_fill_ext<<<div_up(n, 256), 256, 0, h->stream>>>(h->yext, n, std::min(E + 1, L),
                                                       std::min(E, L - 1));
    DTS_LAUNCH_CHECK();
  }
  // The global-memory searches write the wide form, and so do vertical
  // planes.
  const bool nb = tiled && E <= kNarrowBoundsMaxE && !vertical;
  (vertical ? h->bV_narrow : h->bH_narrow) = nb;
  if (tiled) {
    const dim3 grid(div_up(L, kBPos), div_up(nlines, kBLines));
    if (vertical)
      k_bounds_tile<true, false><<<grid, 256, lds, h->stream>>>(h->ctV, L, nlines, sfx, E,
                                                                stage_len, h->bV, h->yext, groups);
    else if (nb)
      k_bounds_tile<false, true><<<grid, 256, lds, h->stream>>>(h->ctH, L, nlines, sfx, E,
                                                                stage_len, h->bH, nullptr, 0);
    else
      k_bounds_tile<false, false><<<grid, 256, lds, h->stream>>>(h->ctH, L, nlines, sfx, E,
                                                                 stage_len, h->bH, nullptr, 0);
  } else if (!vertical) {
    k_bounds_rows<<<dim3(div_up(h->w, 256), h->h), 256, 0, h->stream>>>(h->ctH, h->w, sfx, E,
                                                                       h->bH);
  } else {
    k_bounds_cols<<<dim3(div_up(h->h, 64), div_up(h->w, 64)), 256, 0, h->stream>>>(
        h->ctV, h->w, h->h, sfx, E, h->bV);
  }
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

// Bounds of both directions (sigma_h along rows, sigma_v along columns): one
// launch when both fit the LDS tiles, else two.
int launch_bounds2(dts_handle* h, float sigma_h, float sigma_v) {
  h->tape_valid = false;
  if (sigma_h < 0.f || sigma_v < 0.f) {
    set_error("sigma must be >= 0 (got %g, %g)", (double)sigma_h, (double)sigma_v);
    return DTS_ERR_INVALID_ARG;
  }
  BoundsPair b;
  b.ctH = h->ctH;
  b.ctV = h->ctV;
  b.bH = h->bH;
  b.bV = h->bV;
  b.yext = h->yext;
  b.W = h->w;
  b.H = h->h;
  b.ext_groups = div_up(h->w, 32);
  b.sfx_h = sigma_fx(sigma_h);
  b.sfx_v = sigma_fx(sigma_v);
  b.E_h = sigma_extent(sigma_h, h->w);
  b.E_v = sigma_extent(sigma_v, h->h);
  b.stage_h = std::min(h->w, kBPos + 2 * b.E_h + 2);
  b.stage_v = std::min(h->h, kBPos + 2 * b.E_v + 2);
  const size_t lds = std::max(kBoundsResBytes,
                              sizeof(int64_t) * kBLines * (size_t)std::max(b.stage_h, b.stage_v));
  if (lds > kBoundsLdsMax) {
    int rc = launch_bounds(h, sigma_h, false);
    return rc ? rc : launch_bounds(h, sigma_v, true);
  }
  b.ntx_h = div_up(h->w, kBPos);
  b.ntiles_h = b.ntx_h * div_up(h->h, kBLines);
  b.nb_h = b.E_h <= kNarrowBoundsMaxE;
  h->bH_narrow = b.nb_h != 0;
  h->bV_narrow = false;
  b.ntx_v = div_up(h->h, kBPos);
  const int ntiles_v = b.ntx_v * div_up(h->w, kBLines);
  int rc;
  if (h->ct_of_guide && h->lut) {
    // The ct is that of the current guide (formed or not): stage the lines'
    // ct differences from the guide and the step tables (the same int64
    // differences, so the same bounds) instead of reading the planes.
    BoundsPairG g;
    g.gH = h->color;
    g.gV = h->colorT;
    if ((rc = ensure_lut(h, &g.lutH, &g.lutV))) return rc;
    g.bH = h->bH;
    g.bV = h->bV;
    g.yext = h->yext;
    g.W = h->w;
    g.H = h->h;
    g.ext_groups = b.ext_groups;
    g.sfx_h = b.sfx_h;
    g.sfx_v = b.sfx_v;
    g.E_h = b.E_h;
    g.E_v = b.E_v;
    g.sl_h = guide_stage_stride(b.stage_h);
    g.sl_v = guide_stage_stride(b.stage_v);
    g.ntiles_h = b.ntiles_h;
    g.ntx_h = b.ntx_h;
    g.ntx_v = b.ntx_v;
    g.nb_h = b.nb_h;
    // hipMalloc'd planes are 256-byte aligned; lines of a multiple of 4
    // pixels keep every line 16-byte aligned.
    g.vec_h = h->w % 4 == 0;
    g.vec_v = h->h % 4 == 0;
    const size_t lds_g = std::max(kBoundsResBytes,
                                  sizeof(int64_t) * kBLines * (size_t)std::max(g.sl_h, g.sl_v));
    if (lds_g <= kBoundsLdsMax) {
      k_bounds_both_g<<<b.ntiles_h + ntiles_v, kBoundsGThreads, lds_g, h->stream>>>(g);
      DTS_LAUNCH_CHECK();
      return DTS_OK;
    }
  }
  if ((rc = ensure_ct(h))) return rc;
  k_bounds_both<<<b.ntiles_h + ntiles_v, 256, lds, h->stream>>>(b);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

constexpr size_t kLdsBudget = kLdsMax;

// Profiling: ev[0], ev[1] (when non-null) are recorded by the dispatch
// itself at kernel start and end (hipExtLaunchKernel), so their interval is
// the kernel's execution time without queue gaps.
template <typename K, typename... Args>
int launch_kernel_n(dts_handle* h, K kernel, dim3 grid, int threads, size_t lds,
                    hipEvent_t* ev, Args... args) {
  if (ev)
    hipExtLaunchKernelGGL(kernel, grid, dim3(threads), (uint32_t)lds, h->stream, ev[0], ev[1],
                          0, args...);
  else
    kernel<<<grid, threads, lds, h->stream>>>(args...);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

template <typename K, typename... Args>
int launch_kernel(dts_handle* h, K kernel, dim3 grid, size_t lds, hipEvent_t* ev,
                  Args... args) {
  return launch_kernel_n(h, kernel, grid, 256, lds, ev, args...);
}

template <int MODE, bool OW, bool STEREO, int NCH>
int launch_x_n(dts_handle* h, PassArgs a, hipEvent_t* ev) {
  const size_t lds = sizeof(float) * 4 * 2 * 64 * (NCH | 1);
  a.ntx = div_up(a.W, kXSeg);
  a.nty = 1;
  const int nseg = a.ntx * a.H;
  // About three segments per wavefront, at most one resident round.
  const int grid = std::max(1, std::min(div_up(nseg, 4 * 3), h->num_cus * 4));
  a.tw_div = (4 * grid) / a.ntx;
  a.tw_mod = (4 * grid) % a.ntx;
  if (a.bnarrow)
    return launch_kernel(h, k_xpass<MODE, OW, STEREO, NCH, true>, dim3(grid), lds, ev, a);
  return launch_kernel(h, k_xpass<MODE, OW, STEREO, NCH, false>, dim3(grid), lds, ev, a);
}

// Scratch of the wide-window passes, allocated on first use.  dts_solve and
// dts_optimize allocate it before a graph capture (wide_prepare), so no
// capture allocates.
int ensure_wide(dts_handle* h) {
  if (h->wide_S) return DTS_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  DTS_HIP_CHECK(hipStreamIsCapturing(h->stream, &cs));
  if (cs != hipStreamCaptureStatusNone) {
    set_error("wide-window scratch must be allocated before a graph capture");
    return DTS_ERR_STATE;
  }
  const size_t nt = line_prefix_tot(h->max_w, h->max_h);
  DTS_HIP_CHECK(hipMalloc((void**)&h->wide_S, max_plane(h) * sizeof(double)));
  DTS_HIP_CHECK(hipMalloc((void**)&h->wide_tot, nt * sizeof(double)));
  return DTS_OK;
}

template <int MODE, bool OW, bool STEREO>
int launch_x_wide(dts_handle* h, PassArgs a, hipEvent_t* ev) {
  constexpr bool OWP = OW && (MODE == PM_X_L2 || MODE == PM_X_CH);
  int rc;
  if ((rc = ensure_wide(h))) return rc;
  if (ev) DTS_HIP_CHECK(hipEventRecord(ev[0], h->stream));
  const dim3 gp(div_up(a.W, 256), a.H);
  DTS_HIP_CHECK(line_prefix(h->stream, WideInX<OWP>{a}, a.W, a.H, true, h->wide_tot, h->wide_S));
  if (a.bnarrow)
    k_wide_x<MODE, OW, STEREO, true><<<gp, 256, 0, h->stream>>>(a, h->wide_S);
  else
    k_wide_x<MODE, OW, STEREO, false><<<gp, 256, 0, h->stream>>>(a, h->wide_S);
  DTS_LAUNCH_CHECK();
  if (ev) DTS_HIP_CHECK(hipEventRecord(ev[1], h->stream));
  return DTS_OK;
}

template <int MODE>
int launch_y_wide(dts_handle* h, PassArgs a, hipEvent_t* ev) {
  if (a.bnarrow) {
    set_error("vertical bounds planes of a single frame are wide");
    return DTS_ERR_STATE;
  }
  int rc;
  if ((rc = ensure_wide(h))) return rc;
  if (ev) DTS_HIP_CHECK(hipEventRecord(ev[0], h->stream));
  const dim3 gp(div_up(a.W, 256), a.H);
  DTS_HIP_CHECK(line_prefix(h->stream, PlaneIn{a.vin}, a.W, a.H, false, h->wide_tot, h->wide_S));
  k_wide_y<MODE><<<gp, 256, 0, h->stream>>>(a, h->wide_S);
  DTS_LAUNCH_CHECK();
  if (ev) DTS_HIP_CHECK(hipEventRecord(ev[1], h->stream));
  return DTS_OK;
}

template <int MODE, bool OW, bool STEREO>
int launch_x(dts_handle* h, PassArgs a, hipEvent_t* ev) {
  a.seg_stride = kXSeg + a.elo_cap + a.ehi_cap;
  const int nch = div_up(std::min(a.seg_stride - 1, a.W), 64);
  if (nch <= 5) return launch_x_n<MODE, OW, STEREO, 5>(h, a, ev);
  if (nch <= 6) return launch_x_n<MODE, OW, STEREO, 6>(h, a, ev);
  if (nch <= 8) return launch_x_n<MODE, OW, STEREO, 8>(h, a, ev);
  return launch_x_wide<MODE, OW, STEREO>(h, a, ev);
}

template <int MODE>
int launch_y(dts_handle* h, PassArgs a, hipEvent_t* ev) {
  a.yext = h->yext;
  a.ext_groups = div_up(a.W, 32);
  a.ntx = div_up(a.W, kYCols);
  const int rmax = ytile_rmax(kYRowsReg + a.elo_cap + a.ehi_cap);
  const int rows = rmax ? kYRowsReg : kYRows;
  a.seg_stride = rows + a.elo_cap + a.ehi_cap;
  const size_t lds = sizeof(float) * ((size_t)a.seg_stride * kYStride + 4 * kYCols);
  if (lds > kLdsBudget) return launch_y_wide<MODE>(h, a, ev);
  a.nty = div_up(a.H, rows);
  const dim3 grid(a.ntx * a.nty);
  if (a.bnarrow) {
    set_error("vertical bounds planes of a single frame are wide");
    return DTS_ERR_STATE;
  }
  switch (rmax) {
    case 16: return launch_kernel(h, k_ypass<MODE, 16, false>, grid, lds, ev, a);
    case 24: return launch_kernel(h, k_ypass<MODE, 24, false>, grid, lds, ev, a);
    case 32: return launch_kernel(h, k_ypass<MODE, 32, false>, grid, lds, ev, a);
    case 48: return launch_kernel(h, k_ypass<MODE, 48, false>, grid, lds, ev, a);
    default: return launch_kernel(h, k_ypass<MODE, 0, false>, grid, lds, ev, a);
  }
}

// Whether a pass of window radius sigma takes the wide-window path (in
// either direction, current dims); if so, its scratch is allocated now.
int wide_prepare(dts_handle* h, float sigma) {
  const int Ex = sigma_extent(sigma, h->w), Ey = sigma_extent(sigma, h->h);
  const int segx = kXSeg + std::min(Ex + 1, h->w) + std::min(Ex, h->w - 1);
  const bool wx = div_up(std::min(segx - 1, h->w), 64) > 8;
  const int segy = kYRows + std::min(Ey + 1, h->h) + std::min(Ey, h->h - 1);
  const bool wy = sizeof(float) * ((size_t)segy * kYStride + 4 * kYCols) > kLdsBudget;
  return wx || wy ? ensure_wide(h) : DTS_OK;
}

int launch_pass(dts_handle* h, int mode, bool ow, bool stereo, bool vertical, PassArgs a) {
  hipEvent_t evs[2] = {nullptr, nullptr};
  hipEvent_t* ev = nullptr;
  if (h->prof) {
    DTS_HIP_CHECK(hipEventCreate(&evs[0]));
    DTS_HIP_CHECK(hipEventCreate(&evs[1]));
    ev = evs;
  }
  int rc;
  if (vertical) {
    rc = mode == PM_FILTER ? launch_y<PM_FILTER>(h, a, ev) : launch_y<PM_Y>(h, a, ev);
  } else {
    switch (mode) {
      case PM_FILTER: rc = launch_x<PM_FILTER, false, false>(h, a, ev); break;
      case PM_X_L2:
        if (stereo)
          rc = ow ? launch_x<PM_X_L2, true, true>(h, a, ev) : launch_x<PM_X_L2, false, true>(h, a, ev);
        else
          rc = ow ? launch_x<PM_X_L2, true, false>(h, a, ev)
                  : launch_x<PM_X_L2, false, false>(h, a, ev);
        break;
      case PM_X_CH:
        if (stereo)
          rc = ow ? launch_x<PM_X_CH, true, true>(h, a, ev) : launch_x<PM_X_CH, false, true>(h, a, ev);
        else
          rc = ow ? launch_x<PM_X_CH, true, false>(h, a, ev)
                  : launch_x<PM_X_CH, false, false>(h, a, ev);
        break;
      default:
        rc = DTS_ERR_INVALID_ARG;
        set_error("bad pass mode");
    }
  }
  if (ev) {
    auto& v = vertical ? h->yev : h->xev;
    v.push_back(evs[0]);
    v.push_back(evs[1]);
  }
  return rc;
}

PassArgs base_args(dts_handle* h, bool vertical, float sigma) {
  PassArgs a;
  std::memset(&a, 0, sizeof(a));
  const int L = vertical ? h->h : h->w;
  const int E = sigma_extent(sigma, L);
  a.W = h->w;
  a.H = h->h;
  a.bnd = vertical ? h->bV : h->bH;
  a.bnarrow = vertical ? h->bV_narrow : h->bH_narrow;
  a.out = vertical ? h->vA : h->vB;
  a.elo_cap = std::min(E + 1, L);
  a.ehi_cap = std::min(E, L - 1);
  return a;
}

// src/domain_transform_filter.cu:124-188 on device buffers.
int run_filter(dts_handle* h, const dts_filter_params* p, const float* in, float* out,
               int num_iter) {
  // The filter's ct (which the reference leaves in place for a later
  // Optimize) is integrated lazily; its bounds come from the guide.
  int rc;
  if (h->lut)
    set_ct_lazy(h, p);
  else if ((rc = launch_ct(h, p)))
    return rc;
  const float* cur = in;
  for (int it = 0; it < num_iter; it++) {
    float sx = p->sigma_x, sy = p->sigma_y;
    if (it != 0) {
      const float m = (float)(3 * std::pow(2, num_iter - (it + 1)) /
                              std::sqrt(std::pow(4, num_iter) - 1));
      sx *= m;
      sy *= m;
    }
    if ((rc = launch_bounds2(h, sx, sy))) return rc;
    PassArgs ax = base_args(h, false, sx);
    ax.vin = cur;
    if ((rc = launch_pass(h, PM_FILTER, false, false, false, ax))) return rc;
    PassArgs ay = base_args(h, true, sy);
    ay.vin = h->vB;
    if ((rc = launch_pass(h, PM_FILTER, false, false, true, ay))) return rc;
    cur = h->vA;
  }
  if (num_iter <= 0 && (rc = launch_copy(h, in, h->vA, plane(h)))) return rc;
  return launch_copy(h, h->vA, out, plane(h));
}

// Fused iteration launch (fused.h); ncx chunks of the horizontal window span,
// yr rows per tile.
template <int MODE, int YR, bool ST, bool OW, int NW = fused_waves(YR)>
int launch_xy_rows(dts_handle* h, const PassArgs& ax, const PassArgs& ay, int ncx, size_t lds,
                   hipEvent_t* ev) {
  const dim3 grid(ay.ntx * ay.nty);
  const bool nb = ax.bnarrow != 0;
  const int th = 64 * NW;
  switch (ncx) {
    case 2:
      return nb ? launch_kernel_n(h, k_xypass<MODE, 2, true, YR, ST, OW, NW>, grid, th, lds, ev, ax, ay)
                : launch_kernel_n(h, k_xypass<MODE, 2, false, YR, ST, OW, NW>, grid, th, lds, ev, ax, ay);
    case 3:
      return nb ? launch_kernel_n(h, k_xypass<MODE, 3, true, YR, ST, OW, NW>, grid, th, lds, ev, ax, ay)
                : launch_kernel_n(h, k_xypass<MODE, 3, false, YR, ST, OW, NW>, grid, th, lds, ev, ax, ay);
    default:
      return nb ? launch_kernel_n(h, k_xypass<MODE, 4, true, YR, ST, OW, NW>, grid, th, lds, ev, ax, ay)
                : launch_kernel_n(h, k_xypass<MODE, 4, false, YR, ST, OW, NW>, grid, th, lds, ev, ax, ay);
  }
}

template <int MODE, bool ST, bool OW = false>
int launch_xy(dts_handle* h, const PassArgs& ax, const PassArgs& ay, int ncx, int yr, int nw,
              size_t lds, hipEvent_t* ev) {
  // Sixteen wavefronts on 16-row tiles for the plain iteration (no stereo
  // term, no overwrite), on 64-row tiles without the stereo term.
  if constexpr (!ST && !OW)
    if (yr == 16 && nw == 16) return launch_xy_rows<MODE, 16, ST, OW, 16>(h, ax, ay, ncx, lds, ev);
  if constexpr (!ST)
    if (yr == 64 && nw == 16) return launch_xy_rows<MODE, 64, ST, OW, 16>(h, ax, ay, ncx, lds, ev);
  return yr == 16   ? launch_xy_rows<MODE, 16, ST, OW>(h, ax, ay, ncx, lds, ev)
         : yr == 64 ? launch_xy_rows<MODE, 64, ST, OW>(h, ax, ay, ncx, lds, ev)
                    : launch_xy_rows<MODE, kYRows, ST, OW>(h, ax, ay, ncx, lds, ev);
}

// Launch shape of the fused iteration for the current windows.
struct FusedPlan {
  int ncx;     // 64-column chunks of the horizontal window span
  int yr;      // output rows per tile: 16, 32 or 64 by the tiles per CU
  int nw;      // wavefronts per tile
  size_t lds;  // bytes per workgroup
};

// Fills the vertical tile fields of ay and the plan; false when the windows
// do not fit (a horizontal span beyond 4 chunks, vertical rows beyond LDS).
bool fused_plan(dts_handle* h, const PassArgs& ax, PassArgs& ay, bool stereo, bool ow,
                FusedPlan* f) {
  f->ncx = std::max(2, div_up(63 + ax.elo_cap + ax.ehi_cap, 64));
  if (f->ncx > 4 || ay.bnarrow) return false;
  ay.yext = h->yext;
  ay.ext_groups = div_up(ay.W, 32);
  ay.ntx = div_up(ay.W, kYCols);
  // Short tiles double the tiles of small frames: each tile's serial chain
  // of staged rows, not bandwidth, sets the time there (1024 x 768 depth
  // super-resolution: 9.65 -> 7.59 us per iteration; 320 x 240: 7.6 -> 5.7).
  // 1080p measured 13.8 vs 12.0 us per iteration with 16-row tiles.
  f->yr = ay.ntx * div_up(ay.H, kYRows) < 2 * h->num_cus ? 16 : kYRows;
  // 64-row tiles where they still give about two per CU: fewer recomputed
  // halo rows per output row (1080p: 12.2 -> 12.0 us per iteration).
  if (f->yr == kYRows && ay.ntx * div_up(ay.H, 64) * 10 >= 19 * h->num_cus) f->yr = 64;
  ay.nty = div_up(ay.H, f->yr);
  ay.seg_stride = f->yr + ay.elo_cap + ay.ehi_cap;
  // Sixteen wavefronts per tile halve each wavefront's chain of staged rows:
  // on 16-row tiles of frames with fewer tiles than CUs (plain iteration
  // only: 320 x 240 depth super-resolution 56.0 -> 53.6 us; with more tiles
  // than CUs, 1024 x 768, 1.7 us slower per iteration), and on 64-row tiles
  // where two of them still fit a CU's LDS (not with the stereo term; 1080p
  // depth super-resolution 11.85 -> 11.5 us per iteration, colorization
  // with the overwrite 13.8 -> 13.1 us).
  f->nw = fused_waves(f->yr);
  if (f->yr == 16 && !stereo && !ow && ay.ntx * ay.nty < h->num_cus) f->nw = 16;
  if (f->yr == 64 && !stereo && fused_tile_lds(ay.seg_stride, f->ncx, 16) * 2 <= kLdsMax)
    f->nw = 16;
  f->lds = fused_tile_lds(ay.seg_stride, f->ncx, f->nw);
  return f->lds <= kLdsBudget;
}

// One fused iteration from `in` to `out`.  A tile reads rows of `in` that
// other tiles write to `out`, so the two must be different planes.
// Whether an iteration with (ow) / without the target overwrite may run as
// one launch: not together with the stereo term.
bool fuse_ow(const dts_handle*, bool ow, bool stereo) { return !ow || !stereo; }

// ow: the target overwrite in the horizontal half (not with stereo).
int launch_fused(dts_handle* h, int xmode, bool stereo, bool ow, PassArgs ax, PassArgs ay,
                 const FusedPlan& f, const float* in, float* out) {
  if (stereo && ow) {
    set_error("no fused iteration with both the stereo term and the overwrite");
    return DTS_ERR_UNSUPPORTED;
  }
  if (in == out) {
    set_error("fused iteration cannot run in place");
    return DTS_ERR_INVALID_ARG;
  }
  ax.vin = in;
  ay.out = out;
  hipEvent_t evs[2] = {nullptr, nullptr};
  hipEvent_t* ev = nullptr;
  if (h->prof) {
    DTS_HIP_CHECK(hipEventCreate(&evs[0]));
    DTS_HIP_CHECK(hipEventCreate(&evs[1]));
    ev = evs;
  }
  int rc;
  if (stereo)
    rc = xmode == PM_X_L2 ? launch_xy<PM_X_L2, true>(h, ax, ay, f.ncx, f.yr, f.nw, f.lds, ev)
                          : launch_xy<PM_X_CH, true>(h, ax, ay, f.ncx, f.yr, f.nw, f.lds, ev);
  else if (ow)
    rc = xmode == PM_X_L2 ? launch_xy<PM_X_L2, false, true>(h, ax, ay, f.ncx, f.yr, f.nw, f.lds, ev)
                          : launch_xy<PM_X_CH, false, true>(h, ax, ay, f.ncx, f.yr, f.nw, f.lds, ev);
  else
    rc = xmode == PM_X_L2 ? launch_xy<PM_X_L2, false>(h, ax, ay, f.ncx, f.yr, f.nw, f.lds, ev)
                          : launch_xy<PM_X_CH, false>(h, ax, ay, f.ncx, f.yr, f.nw, f.lds, ev);
  if (ev) {
    h->xyev.push_back(evs[0]);
    h->xyev.push_back(evs[1]);
  }
  return rc;
}

// src/domain_transform_solver.cu:100-202 / src/domain_transform_stereo.cu:44-158
int run_optimize(dts_handle* h, const dts_optimize_params* p, float overwrite,
                 bool stereo) {
  if (!p) {
    set_error("null optimize params");
    return DTS_ERR_INVALID_ARG;
  }
  if (p->loss != DTS_LOSS_L1 && p->loss != DTS_LOSS_L2 &&
      p->loss != DTS_LOSS_CHARBONNIER) {
    set_error("unknown loss %d", p->loss);
    return DTS_ERR_INVALID_ARG;
  }
  int rc;
  // L1 launches no kernel in the reference; its ping-pong leaves the target.
  if (p->loss == DTS_LOSS_L1 || p->num_iterations <= 0)
    return launch_copy(h, h->target, h->vA, plane(h));
  const float sigma = p->sigma_z;
  if ((rc = launch_bounds2(h, sigma, sigma))) return rc;
  const bool ow = overwrite > 0.f && overwrite < 1.f;
  const int xmode = p->loss == DTS_LOSS_L2 ? PM_X_L2 : PM_X_CH;
  PassArgs ax = base_args(h, false, sigma);
  ax.t = h->target;
  ax.c = h->conf;
  ax.lambda = p->lambda;
  ax.step = p->step_size;
  ax.theta = overwrite;
  if (stereo) {
    const float q = h->fp.sigma_r / h->fp.sigma_x;
    ax.left = h->color;
    ax.right = h->rcolor;
    ax.rdH = h->rdHf;
    ax.doff = h->doff;
    ax.photo_lambda = p->photo_consis_lambda;
    ax.grad_alpha = q * q;
  }
  PassArgs ay = base_args(h, true, sigma);
  ay.vin = h->vB;
  ay.lambda = p->lambda;
  ay.step = p->step_size;
  // One launch per iteration (depth super-resolution 1080p: 0.203 -> 0.169
  // ms; with the overwrite, whose recomputed halo rows carry target and
  // confidence too, the colorization at 1080p: 0.362 -> 0.315 ms on 64-row
  // tiles, the same time as two passes on 32-row tiles).  Not with both the
  // overwrite and the stereo term.  Iteration k writes vA or vB so that the
  // last one lands in vA.
  FusedPlan f;
  if (h->fuse && fuse_ow(h, ow, stereo) && fused_plan(h, ax, ay, stereo, ow, &f)) {
    const int K = p->num_iterations;
    const float* in = h->target;
    for (int it = 0; it < K; it++) {
      float* out = (K - 1 - it) % 2 == 0 ? h->vA : h->vB;
      if ((rc = launch_fused(h, xmode, stereo, ow, ax, ay, f, in, out))) return rc;
      in = out;
    }
    return DTS_OK;
  }
  for (int it = 0; it < p->num_iterations; it++) {
    ax.vin = it == 0 ? h->target : h->vA;
    if ((rc = launch_pass(h, xmode, ow, stereo, false, ax))) return rc;
    if ((rc = launch_pass(h, PM_Y, false, false, true, ay))) return rc;
  }
  return DTS_OK;
}

// src/domain_transform_solver.cu:206-246
int run_confidence(dts_handle* h, const dts_filter_params* cp, float sigma0,
                   int left_clear) {
  int rc;
  if ((rc = check_params(cp))) return rc;
  if ((rc = run_filter(h, cp, h->target, h->conf, 3))) return rc;
  k_square_f32<<<(unsigned)((plane(h) + 255) / 256), 256, 0, h->stream>>>(h->target, h->csqb,
                                                                         plane(h));
  DTS_LAUNCH_CHECK();
  if ((rc = run_filter(h, cp, h->csqb, h->csq, 3))) return rc;
  const float scale = 0.5f / (sigma0 * sigma0);
  k_confidence<<<dim3(div_up(h->w, 256), h->h), 256, 0, h->stream>>>(h->csq, h->conf, h->w,
                                                                    h->h, scale, left_clear);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int run_solve(dts_handle* h, const dts_solve_options* o, int compute_conf) {
  int rc;
  if ((rc = check_params(&o->filter_options))) return rc;
  h->fp = o->filter_options;
  h->fp_set = true;
  // A handle without step tables (the no-allocation solve) forms the planes.
  if (h->lut)
    set_ct_lazy(h, &o->filter_options);
  else if ((rc = launch_ct(h, &o->filter_options)))
    return rc;
  if (compute_conf &&
      (rc = run_confidence(h, &o->conf_params, o->conf_sigma0, 0)))
    return rc;
  return run_optimize(h, &o->optim_options, o->overwrite_target_above_conf, false);
}

void drop_graph(dts_handle* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  h->gexec = nullptr;
  h->gkey.clear();
}

template <typename F>
int graph_or_direct(dts_handle* h, const void* key_data, size_t key_len, F&& body,
                    bool ct_state_in_key = true) {
  h->tape_valid = false;  // every captured sequence recomputes bounds or state
  if (!h->graph_mode) return body();
  // The captured launches depend on the ct state (formed planes or bounds
  // from the guide, and the params of the ct), so it is part of the key
  // unless the body sets that state itself.
  struct {
    int w, h, lazy, of_guide;
    dts_filter_params ctp;
  } hk;
  std::memset(&hk, 0, sizeof(hk));
  hk.w = h->w;
  hk.h = h->h;
  if (ct_state_in_key) {
    hk.lazy = h->ct_lazy;
    hk.of_guide = h->ct_of_guide;
    hk.ctp = h->ctp;
  }
  std::vector<char> key(sizeof(hk) + key_len);
  std::memcpy(key.data(), &hk, sizeof(hk));
  std::memcpy(key.data() + sizeof(hk), key_data, key_len);
  if (!h->gexec || key != h->gkey) {
    drop_graph(h);
    DTS_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const bool prof = h->prof;
    h->prof = false;
    h->capturing = true;
    const int rc = body();
    h->capturing = false;
    h->prof = prof;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc || e != hipSuccess) {
      // Step tables whose builds were captured were not built.
      h->lut_q[0] = h->lut_q[1] = NAN;
    }
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return DTS_ERR_HIP;
    }
    const hipError_t e2 = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
      h->gexec = nullptr;
      h->lut_q[0] = h->lut_q[1] = NAN;
      set_error("graph instantiate failed: %s", hipGetErrorString(e2));
      return DTS_ERR_HIP;
    }
    h->gkey = key;
  }
  DTS_HIP_CHECK(hipGraphLaunch(h->gexec, h->stream));
  return DTS_OK;
}

int upload(dts_handle* h, void* dst, const void* src, size_t bytes, bool device) {
  if (bytes == 0) return DTS_OK;
  DTS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes,
                               device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                               h->stream));
  return DTS_OK;
}

int upload_guide(dts_handle* h, int cs, const void* rgba, bool device) {
  if (cs < DTS_COLOR_RGB || cs > DTS_COLOR_YYY) {
    set_error("unknown colour space %d", cs);
    return DTS_ERR_INVALID_ARG;
  }
  h->rf_valid = false;
  const size_t n = plane(h);
  int rc;
  // A lazy integration refers to the guide being replaced.
  if ((rc = ensure_ct(h))) return rc;
  h->ct_of_guide = false;
  if ((rc = upload(h, h->color, rgba, n * sizeof(uchar4), device))) return rc;
  if (cs != DTS_COLOR_RGB) {
    k_convert_color<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(h->color, (int)n, cs);
    DTS_LAUNCH_CHECK();
  }
  // The transposed copy lets the vertical differential scan contiguous lines.
  k_transpose_u8x4<<<dim3(div_up(h->w, 64), div_up(h->h, 64)), 256, 0, h->stream>>>(
      h->color, h->w, h->h, h->colorT);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int init_frame_impl(dts_handle* h, int cs, const void* rgba, const float* t,
                    const float* c, bool device) {
  h->tape_valid = false;
  if (!rgba || !t) {
    set_error("null guide or target");
    return DTS_ERR_INVALID_ARG;
  }
  const size_t n = plane(h);
  int rc;
  if ((rc = upload_guide(h, cs, rgba, device))) return rc;
  if ((rc = upload(h, h->target, t, n * sizeof(float), device))) return rc;
  if (c) {
    if ((rc = upload(h, h->conf, c, n * sizeof(float), device))) return rc;
  } else if ((rc = launch_fill(h, h->conf, max_plane(h), 1.0f))) {
    return rc;
  }
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

// ------------------------------------------------ differentiable solve
// Checkpointed forward: the iterations of run_optimize (same kernels, same
// result), with the vertical pass of iteration k writing v_{k+1} into
// checkpoint m when k + 1 = m C, into state i of the last segment when
// k + 1 = (M - 1) C + i, and into vA for the final result.  The adjoint
// replays each earlier segment from its checkpoint (sqrt(K) memory, about
// K (1 - 1/C) - C extra forward iterations) and runs the adjoint passes
// from the last iteration down.
int ensure_buf(dts_handle* h, float** p, size_t floats) {
  if (*p) return DTS_OK;
  DTS_HIP_CHECK(hipMalloc((void**)p, floats * sizeof(float)));
  return DTS_OK;
}

float* tape_ckpt(dts_handle* h, int m) {
  return m == 0 ? h->target : h->tape + (size_t)(m - 1) * max_plane(h);
}

float* tape_state(dts_handle* h, int M, int i) {
  return h->tape + (size_t)(M - 1 + i) * max_plane(h);
}

PassArgs data_pass_args(dts_handle* h, const dts_optimize_params* p, float overwrite) {
  PassArgs ax = base_args(h, false, p->sigma_z);
  ax.t = h->target;
  ax.c = h->conf;
  ax.lambda = p->lambda;
  ax.step = p->step_size;
  ax.theta = overwrite;
  return ax;
}

// The pass arguments of one forward iteration (horizontal output to vB).
void iteration_args(dts_handle* h, const dts_optimize_params* p, float overwrite, PassArgs* ax,
                    PassArgs* ay) {
  *ax = data_pass_args(h, p, overwrite);
  ax->out = h->vB;
  *ay = base_args(h, true, p->sigma_z);
  ay->vin = h->vB;
  ay->lambda = p->lambda;
  ay->step = p->step_size;
}

// Whether run_optimize takes one launch per iteration for these options.
bool iteration_fused(dts_handle* h, const dts_optimize_params* p, float overwrite) {
  const bool ow = overwrite > 0.f && overwrite < 1.f;
  if (!h->fuse || !fuse_ow(h, ow, false)) return false;
  PassArgs ax, ay;
  iteration_args(h, p, overwrite, &ax, &ay);
  FusedPlan f;
  return fused_plan(h, ax, ay, false, ow, &f);
}

// One forward iteration from `in`, vertical output to `out`: the launches
// run_optimize makes for it, so the states equal its iterates bit for bit.
// The two-pass form may run in place (in == out; vB holds the horizontal
// output), the fused form may not.
int forward_iteration(dts_handle* h, const dts_optimize_params* p, float overwrite,
                      const float* in, float* out) {
  const bool ow = overwrite > 0.f && overwrite < 1.f;
  const int xmode = p->loss == DTS_LOSS_L2 ? PM_X_L2 : PM_X_CH;
  PassArgs ax, ay;
  iteration_args(h, p, overwrite, &ax, &ay);
  ax.vin = in;
  ay.out = out;
  FusedPlan f;
  if (h->fuse && fuse_ow(h, ow, false) && fused_plan(h, ax, ay, false, ow, &f))
    return launch_fused(h, xmode, false, ow, ax, ay, f, in, out);
  int rc;
  if ((rc = launch_pass(h, xmode, ow, false, false, ax))) return rc;
  return launch_pass(h, PM_Y, false, false, true, ay);
}

int run_optimize_ckpt(dts_handle* h, const dts_optimize_params* p, float overwrite,
                      int ckpt_every) {
  if (!p) {
    set_error("null optimize params");
    return DTS_ERR_INVALID_ARG;
  }
  if (h->flags & DTS_FLAG_STEREO) {
    set_error("the checkpointed solve has no photometric term; use a plain handle");
    return DTS_ERR_UNSUPPORTED;
  }
  if (p->loss != DTS_LOSS_L1 && p->loss != DTS_LOSS_L2 && p->loss != DTS_LOSS_CHARBONNIER) {
    set_error("unknown loss %d", p->loss);
    return DTS_ERR_INVALID_ARG;
  }
  int rc;
  const size_t P = max_plane(h);
  if ((rc = ensure_buf(h, &h->adj_g, P)) || (rc = ensure_buf(h, &h->adj_gb, P)) ||
      (rc = ensure_buf(h, &h->adj_tb, P)) || (rc = ensure_buf(h, &h->adj_cb, P)))
    return rc;
  const int K = p->num_iterations;
  if (p->loss == DTS_LOSS_L1 || K <= 0) {
    if ((rc = launch_copy(h, h->target, h->vA, plane(h)))) return rc;
    h->tape_p = *p;
    h->tape_ow = overwrite;
    h->tape_K = 0;
    h->tape_C = 1;
    h->tape_valid = true;
    return DTS_OK;
  }
  const float sigma = p->sigma_z;
  if ((rc = launch_bounds2(h, sigma, sigma))) return rc;
  if (!h->abH) DTS_HIP_CHECK(hipMalloc((void**)&h->abH, sizeof(int2) * h->max_h));
  if (!h->abV) DTS_HIP_CHECK(hipMalloc((void**)&h->abV, sizeof(int2) * h->max_w));
  const int64_t sfx = sigma_fx(sigma);
  if ((rc = ensure_ct(h))) return rc;
  if ((rc = adj_collapse2(h->stream, h->ctH, h->ctV, h->w, h->h, sfx,
                          sigma_extent(sigma, h->w), sigma_extent(sigma, h->h), h->abH,
                          h->abV)))
    return rc;
  int C = ckpt_every > 0 ? ckpt_every : (int)std::ceil(std::sqrt((double)K));
  C = std::max(1, std::min(C, K));
  const int M = (K + C - 1) / C;
  const size_t planes = (size_t)(M - 1) + C;
  if (planes > h->tape_planes) {
    DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (h->tape) DTS_HIP_CHECK(hipFree(h->tape));
    h->tape = nullptr;
    h->tape_planes = 0;
    DTS_HIP_CHECK(hipMalloc((void**)&h->tape, planes * P * sizeof(float)));
    h->tape_planes = planes;
  }
  const int last0 = (M - 1) * C;
  // Iterates kept nowhere go to vA, in place, for the two-pass form; the
  // fused form alternates vB and vA (the final result lands in vA).
  const bool fused = iteration_fused(h, p, overwrite);
  const float* cur = h->target;
  for (int k = 0; k < K; k++) {
    const int n = k + 1;
    float* dst;
    if (n == K)
      dst = cur == h->vA ? h->vB : h->vA;
    else if (n % C == 0)
      dst = tape_ckpt(h, n / C);
    else if (n > last0)
      dst = tape_state(h, M, n - last0);
    else
      dst = fused && cur == h->vA ? h->vB : h->vA;
    if ((rc = forward_iteration(h, p, overwrite, cur, dst))) return rc;
    cur = dst;
  }
  if (cur != h->vA && (rc = launch_copy(h, cur, h->vA, plane(h)))) return rc;
  h->tape_p = *p;
  h->tape_ow = overwrite;
  h->tape_K = K;
  h->tape_C = C;
  h->tape_valid = true;
  return DTS_OK;
}

int run_optimize_adjoint(dts_handle* h, const float* ybar, float* tbar, float* cbar,
                         float* gbar, bool device) {
  if (!h->tape_valid) {
    set_error("no checkpointed Optimize to differentiate (run dts_optimize_checkpointed "
              "after the last change of frame, dims or bounds)");
    return DTS_ERR_STATE;
  }
  if (!ybar || !tbar || !cbar) {
    set_error("null ybar, tbar or cbar");
    return DTS_ERR_INVALID_ARG;
  }
  const size_t n = plane(h);
  const hipMemcpyKind down = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  int rc;
  // Device calls read ybar in place and accumulate into the caller's tbar
  // and cbar; host calls go through the handle's planes.
  const float* yb = device ? ybar : h->adj_g;
  float* tb = device ? tbar : h->adj_tb;
  float* cb = device ? cbar : h->adj_cb;
  if (!device && (rc = upload(h, h->adj_g, ybar, n * sizeof(float), false))) return rc;
  const dts_optimize_params* p = &h->tape_p;
  const int K = h->tape_K, C = h->tape_C;
  if (K > 0) {
    const int M = (K + C - 1) / C;
    const float sigma = p->sigma_z;
    const bool ow = h->tape_ow > 0.f && h->tape_ow < 1.f;
    AdjArgs ay{};
    {
      const PassArgs b = base_args(h, true, sigma);
      ay.W = h->w;
      ay.H = h->h;
      ay.elo = b.elo_cap;
      ay.ehi = b.ehi_cap;
    }
    ay.gin = h->adj_g;
    ay.bnd = h->bV;
    ay.bnarrow = h->bV_narrow;
    ay.ab = h->abV;
    ay.vout = h->adj_gb;
    ay.lambda = p->lambda;
    ay.step = p->step_size;
    AdjArgs ax{};
    {
      const PassArgs b = base_args(h, false, sigma);
      ax.W = h->w;
      ax.H = h->h;
      ax.elo = b.elo_cap;
      ax.ehi = b.ehi_cap;
    }
    ax.gin = h->adj_gb;
    ax.bnd = h->bH;
    ax.bnarrow = h->bH_narrow;
    ax.ab = h->abH;
    ax.t = h->target;
    ax.c = h->conf;
    ax.vout = h->adj_g;
    ax.tbar = tb;
    ax.cbar = cb;
    ax.lambda = p->lambda;
    ax.step = p->step_size;
    ax.theta = h->tape_ow;
    ax.ow = ow ? 1 : 0;
    ax.loss = p->loss;
    // Windows beyond the LDS segments take the double-precision path.
    if ((rc = wide_prepare(h, sigma))) return rc;
    ax.wS = ay.wS = h->wide_S;
    ax.wtot = ay.wtot = h->wide_tot;
    ay.yext = h->yext;
    ay.ext_groups = div_up(h->w, 32);
    // Segment m's states 1 .. C-1, replayed from its checkpoint (the last
    // segment's come from the forward).
    auto replay = [&](int m) -> int {
      const int k0 = m * C, k1 = std::min(K, k0 + C);
      const float* cur = tape_ckpt(h, m);
      for (int i = 1; i < k1 - k0; i++) {
        float* dst = tape_state(h, M, i);
        int r;
        if ((r = forward_iteration(h, p, h->tape_ow, cur, dst))) return r;
        cur = dst;
      }
      return DTS_OK;
    };
    auto state_of = [&](int k) {
      const int m = k / C, i = k - m * C;
      return i == 0 ? (const float*)tape_ckpt(h, m) : (const float*)tape_state(h, M, i);
    };
    if (adj_xy_fits(ax, ay)) {
      // The vertical adjoint of the last iteration, then the horizontal
      // adjoint of iteration k with the vertical adjoint of iteration k - 1
      // as one launch (k = K-1 .. 1), then the horizontal adjoint of
      // iteration 0.  A segment is replayed just before its last iteration's
      // horizontal adjoint, the first to read its states.
      float* cur = h->adj_gb;
      float* nxt = h->adj_g;
      ay.gin = yb;
      ay.vout = cur;
      if ((rc = adj_ypass(h->stream, ay))) return rc;
      for (int k = K - 1; k >= 0; k--) {
        const int m = k / C;
        if (m != M - 1 && k == m * C + C - 1 && C > 1 && (rc = replay(m))) return rc;
        ax.gin = cur;
        ax.state = state_of(k);
        ax.acc_first = k == K - 1;
        ax.acc_last = k == 0;
        if (k == 0) {
          ax.vout = nxt;
          if ((rc = adj_xpass(h->stream, ax))) return rc;
        } else {
          ay.vout = nxt;
          if ((rc = adj_xypass(h->stream, ax, ay))) return rc;
          std::swap(cur, nxt);
        }
      }
    } else {
    for (int m = M - 1; m >= 0; m--) {
      const int k0 = m * C, k1 = std::min(K, k0 + C);
      if (m != M - 1 && (rc = replay(m))) return rc;
      for (int k = k1 - 1; k >= k0; k--) {
        const int i = k - k0;
        ay.gin = k == K - 1 ? yb : h->adj_g;
        if ((rc = adj_ypass(h->stream, ay))) return rc;
        ax.state = i == 0 ? tape_ckpt(h, m) : tape_state(h, M, i);
        // The first pass starts d/dt and d/dc at zero; the last adds the
        // cotangent of v_0 (the target) to d/dt.
        ax.acc_first = k == K - 1;
        ax.acc_last = k == 0;
        if ((rc = adj_xpass(h->stream, ax))) return rc;
      }
    }
    }
  } else {
    // L1 / no iterations: the result is the target.
    if (device) {
      DTS_HIP_CHECK(hipMemcpyAsync(tbar, ybar, n * sizeof(float), hipMemcpyDeviceToDevice,
                                   h->stream));
    } else {
      std::memcpy(tbar, ybar, n * sizeof(float));
    }
    if ((rc = launch_fill(h, cb, n, 0.f))) return rc;
  }
  if (!device) {
    if (K > 0) DTS_HIP_CHECK(hipMemcpyAsync(tbar, tb, n * sizeof(float), down, h->stream));
    DTS_HIP_CHECK(hipMemcpyAsync(cbar, cb, n * sizeof(float), down, h->stream));
  }
  if (gbar) {
    // The guide moves only the integer window bounds: zero almost everywhere.
    if (device) {
      if ((rc = launch_fill(h, gbar, 4 * n, 0.f))) return rc;
    } else {
      std::memset(gbar, 0, 4 * n * sizeof(float));
    }
  }
  // Device results are stream-ordered on the handle's stream; host copies
  // are complete on return.
  if (!device) DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

// ------------------------------------------------------------ RF-DT
// Work planes (max plane each): d along rows [H][W], d along columns [W][H],
// their cotangents, normal and transposed work planes, two sweep scratch
// planes, and four planes for the guide cotangent (float4 per pixel).
enum { kRfDH, kRfDV, kRfDbH, kRfDbV, kRfP, kRfT, kRfY, kRfG, kRfGuide, kRfWork = kRfGuide + 4 };
constexpr int kRfMaxIter = 16;

float* rf_plane(dts_handle* h, int i) { return h->rf_work + (size_t)i * max_plane(h); }
float* rf_tape_plane(dts_handle* h, int i) { return h->rf_tape + (size_t)i * max_plane(h); }

// sigma_i = sigma sqrt(3) 2^(N-1-i) / sqrt(4^N - 1); returns -sqrt(2) / sigma_i.
double rf_dlna(float sigma, int i, int n) {
  const double si = (double)sigma * std::sqrt(3.0) * std::ldexp(1.0, n - 1 - i) /
                    std::sqrt(std::ldexp(1.0, 2 * n) - 1.0);
  return -std::sqrt(2.0) / si;
}

int run_rf_filter(dts_handle* h, const dts_filter_params* p, const float* input, float* output,
                  int n, bool keep, bool device) {
  if (!p || !input || !output) return DTS_ERR_INVALID_ARG;
  if (n < 1 || n > kRfMaxIter) {
    set_error("RF iterations must be in [1, %d] (got %d)", kRfMaxIter, n);
    return DTS_ERR_INVALID_ARG;
  }
  if (!(p->sigma_x > 0.f) || !(p->sigma_y > 0.f) || !(p->sigma_r > 0.f)) {
    set_error("RF sigmas must be > 0");
    return DTS_ERR_INVALID_ARG;
  }
  h->rf_valid = false;
  const size_t MP = max_plane(h), P = plane(h);
  if (!h->rf_work) DTS_HIP_CHECK(hipMalloc((void**)&h->rf_work, kRfWork * MP * sizeof(float)));
  if (keep && h->rf_tape_planes < (size_t)(2 * n)) {
    DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (h->rf_tape) DTS_HIP_CHECK(hipFree(h->rf_tape));
    h->rf_tape = nullptr;
    h->rf_tape_planes = 0;
    DTS_HIP_CHECK(hipMalloc((void**)&h->rf_tape, (size_t)2 * n * MP * sizeof(float)));
    h->rf_tape_planes = (size_t)2 * n;
  }
  const int W = h->w, H = h->h;
  hipStream_t s = h->stream;
  int rc;
  float *dH = rf_plane(h, kRfDH), *dV = rf_plane(h, kRfDV);
  float *Pz = rf_plane(h, kRfP), *Tz = rf_plane(h, kRfT);
  if ((rc = rf_diff(s, h->color, W, H, p->sigma_x / p->sigma_r, dH))) return rc;
  if ((rc = rf_diff(s, h->colorT, H, W, p->sigma_y / p->sigma_r, dV))) return rc;
  float* cur = keep ? rf_tape_plane(h, 0) : Pz;
  DTS_HIP_CHECK(hipMemcpyAsync(cur, input, P * sizeof(float),
                               device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  const double log2e = 1.4426950408889634;
  for (int i = 0; i < n; i++) {
    const float cx = (float)(rf_dlna(p->sigma_x, i, n) * log2e);
    const float cy = (float)(rf_dlna(p->sigma_y, i, n) * log2e);
    if ((rc = rf_rows(s, cur, Pz, dH, W, H, cx))) return rc;
    float* tv = keep ? rf_tape_plane(h, 2 * i + 1) : Tz;
    if ((rc = rf_transpose(s, Pz, tv, W, H))) return rc;
    if ((rc = rf_rows(s, tv, Tz, dV, H, W, cy))) return rc;
    cur = keep && i + 1 < n ? rf_tape_plane(h, 2 * i + 2) : Pz;
    if ((rc = rf_transpose(s, Tz, cur, H, W))) return rc;
  }
  DTS_HIP_CHECK(hipMemcpyAsync(output, cur, P * sizeof(float),
                               device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  if (!device) DTS_HIP_CHECK(hipStreamSynchronize(s));
  if (keep) {
    h->rf_valid = true;
    h->rf_n = n;
    h->rf_p = *p;
  }
  return DTS_OK;
}

int run_rf_adjoint(dts_handle* h, const float* ybar, float* xbar, float* gbar, bool device) {
  if (!ybar || !xbar) return DTS_ERR_INVALID_ARG;
  if (!h->rf_valid) {
    set_error("no RF tape: run dts_rf_filter with keep_tape on the current frame first");
    return DTS_ERR_STATE;
  }
  const int W = h->w, H = h->h, n = h->rf_n;
  const size_t P = plane(h);
  const dts_filter_params& p = h->rf_p;
  hipStream_t s = h->stream;
  int rc;
  float *dH = rf_plane(h, kRfDH), *dV = rf_plane(h, kRfDV);
  float *dbH = rf_plane(h, kRfDbH), *dbV = rf_plane(h, kRfDbV);
  float *Pb = rf_plane(h, kRfP), *Tb = rf_plane(h, kRfT);
  float *yt = rf_plane(h, kRfY), *gt = rf_plane(h, kRfG);
  DTS_HIP_CHECK(hipMemcpyAsync(Pb, ybar, P * sizeof(float),
                               device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  if ((rc = launch_fill(h, dbH, P, 0.f))) return rc;
  if ((rc = launch_fill(h, dbV, P, 0.f))) return rc;
  const double log2e = 1.4426950408889634;
  for (int i = n - 1; i >= 0; i--) {
    const double lx = rf_dlna(p.sigma_x, i, n), ly = rf_dlna(p.sigma_y, i, n);
    if ((rc = rf_transpose(s, Pb, Tb, W, H))) return rc;
    if ((rc = rf_rows_adj(s, rf_tape_plane(h, 2 * i + 1), Tb, yt, gt, dV, dbV, H, W,
                          (float)(ly * log2e), (float)ly)))
      return rc;
    if ((rc = rf_transpose(s, Tb, Pb, H, W))) return rc;
    if ((rc = rf_rows_adj(s, rf_tape_plane(h, 2 * i), Pb, yt, gt, dH, dbH, W, H,
                          (float)(lx * log2e), (float)lx)))
      return rc;
  }
  DTS_HIP_CHECK(hipMemcpyAsync(xbar, Pb, P * sizeof(float),
                               device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  if (gbar) {
    // d = sqrt(1 + q^2 K S) with the fp32 ratio q, as the differentials
    const float qx = p.sigma_x / p.sigma_r, qy = p.sigma_y / p.sigma_r;
    const float kx = (float)((double)qx * qx * (double)kINV255SQ_F);
    const float ky = (float)((double)qy * qy * (double)kINV255SQ_F);
    float* g4 = rf_plane(h, kRfGuide);
    // the column planes back in row-major order, so every access coalesces
    if ((rc = rf_transpose(s, dV, yt, H, W))) return rc;
    if ((rc = rf_transpose(s, dbV, gt, H, W))) return rc;
    if ((rc = rf_guide_grad(s, h->color, W, H, dH, dbH, yt, gt, kx, ky, g4))) return rc;
    DTS_HIP_CHECK(hipMemcpyAsync(gbar, g4, 4 * P * sizeof(float),
                                 device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  }
  if (!device) DTS_HIP_CHECK(hipStreamSynchronize(s));
  return DTS_OK;
}

}  // namespace

// ================================================================== C ABI
extern "C" {

const char* dts_last_error(void) { return last_error(); }
int dts_version(void) { return 1; }

int dts_create(int max_w, int max_h, int flags, dts_handle** out) {
  if (!out) {
    set_error("null out");
    return DTS_ERR_INVALID_ARG;
  }
  *out = nullptr;
  if (max_w <= 0 || max_h <= 0 || max_w > kMaxDim || max_h > kMaxDim) {
    set_error("max dims %dx%d outside 1..%d", max_w, max_h, kMaxDim);
    return DTS_ERR_INVALID_DIMS;
  }
  dts_handle* h = new dts_handle();
  h->max_w = h->w = max_w;
  h->max_h = h->h = max_h;
  h->flags = flags;
  const size_t n = (size_t)max_w * max_h;
  auto fail = [&](hipError_t e) {
    set_error("allocation failed: %s", hipGetErrorString(e));
    dts_destroy(h);
    return DTS_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking))) return fail(e);
  h->stream = h->own_stream;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
      h->num_cus = cus;
  }
#define DTS_ALLOC(ptr, bytes) \
  if ((e = hipMalloc((void**)&(ptr), (bytes)))) return fail(e);
  DTS_ALLOC(h->color, n * sizeof(uchar4));
  DTS_ALLOC(h->colorT, n * sizeof(uchar4));
  DTS_ALLOC(h->target, n * sizeof(float));
  DTS_ALLOC(h->conf, n * sizeof(float));
  DTS_ALLOC(h->vA, n * sizeof(float));
  DTS_ALLOC(h->vB, n * sizeof(float));
  DTS_ALLOC(h->csq, n * sizeof(float));
  DTS_ALLOC(h->csqb, n * sizeof(float));
  DTS_ALLOC(h->ctH, n * sizeof(int64_t));
  DTS_ALLOC(h->ctV, n * sizeof(int64_t));
  DTS_ALLOC(h->lut, 2 * sizeof(int64_t) * (size_t)kStepLutLen);
  DTS_ALLOC(h->bH, n * sizeof(uint32_t));
  DTS_ALLOC(h->bV, n * sizeof(uint32_t));
  DTS_ALLOC(h->yext, 2 * sizeof(int) * (size_t)div_up(max_w, 32) * div_up(max_h, 32));
  if (flags & DTS_FLAG_STEREO) {
    DTS_ALLOC(h->rcolor, n * sizeof(uchar4));
    DTS_ALLOC(h->rdHf, n * sizeof(float));
  }
#undef DTS_ALLOC
  if ((e = hipEventCreate(&h->t0))) return fail(e);
  if ((e = hipEventCreate(&h->t1))) return fail(e);
  const int rc = dts_clear_all(h);
  if (rc) {
    dts_destroy(h);
    return rc;
  }
  *out = h;
  return DTS_OK;
}

int dts_destroy(dts_handle* h) {
  if (!h) return DTS_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  for (auto* v : {&h->xev, &h->yev, &h->xyev})
    for (auto e : *v) (void)hipEventDestroy(e);
  void* ptrs[] = {h->color, h->colorT, h->rcolor, h->rdHf, h->target, h->conf,
                  h->vA,    h->vB,     h->csq,    h->csqb, h->ctH,    h->ctV,
                  h->bH,    h->bV,     h->yext,   h->tape, h->abH,    h->abV,
                  h->adj_g, h->adj_gb, h->adj_tb, h->adj_cb, h->rf_work, h->rf_tape,
                  h->wide_S, h->wide_tot, h->lut};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->t0) (void)hipEventDestroy(h->t0);
  if (h->t1) (void)hipEventDestroy(h->t1);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
  return DTS_OK;
}

int dts_get_stream(dts_handle* h, void** s) {
  if (!h || !s) return DTS_ERR_INVALID_ARG;
  *s = (void*)h->stream;
  return DTS_OK;
}

int dts_set_stream(dts_handle* h, void* s) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->stream = s ? (hipStream_t)s : h->own_stream;
  return DTS_OK;
}

int dts_synchronize(dts_handle* h) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int dts_set_image_dim(dts_handle* h, int w, int hh) {
  if (!h) return DTS_ERR_INVALID_ARG;
  if (w > 0 && w <= h->max_w && hh > 0 && hh <= h->max_h) {
    h->tape_valid = false;
    h->rf_valid = false;
    int rc;
    if ((rc = ensure_ct(h))) return rc;  // a lazy integration at the old dims
    h->ct_of_guide = false;
    h->w = w;
    h->h = hh;
    return DTS_OK;
  }
  set_error("Invalid image dims WxH: %d x %d", w, hh);
  return DTS_ERR_INVALID_DIMS;
}

int dts_get_image_dim(dts_handle* h, int* w, int* hh) {
  if (!h || !w || !hh) return DTS_ERR_INVALID_ARG;
  *w = h->w;
  *hh = h->h;
  return DTS_OK;
}

int dts_clear_all(dts_handle* h) {
  if (!h) return DTS_ERR_INVALID_ARG;
  h->tape_valid = false;
  h->rf_valid = false;
  const size_t n = max_plane(h);
  void* f4[] = {h->target, h->conf, h->vA, h->vB, h->csq, h->csqb, h->bH, h->bV, h->rdHf};
  for (void* p : f4)
    if (p) DTS_HIP_CHECK(hipMemsetAsync(p, 0, n * 4, h->stream));
  DTS_HIP_CHECK(hipMemsetAsync(h->color, 0, n * sizeof(uchar4), h->stream));
  DTS_HIP_CHECK(hipMemsetAsync(h->colorT, 0, n * sizeof(uchar4), h->stream));
  if (h->rcolor) DTS_HIP_CHECK(hipMemsetAsync(h->rcolor, 0, n * sizeof(uchar4), h->stream));
  DTS_HIP_CHECK(hipMemsetAsync(h->ctH, 0, n * sizeof(int64_t), h->stream));
  DTS_HIP_CHECK(hipMemsetAsync(h->ctV, 0, n * sizeof(int64_t), h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  h->ct_lazy = false;
  h->ct_of_guide = false;
  return DTS_OK;
}

int dts_init_frame(dts_handle* h, int cs, const void* rgba, const float* t,
                   const float* c) {
  if (!h) return DTS_ERR_INVALID_ARG;
  return init_frame_impl(h, cs, rgba, t, c, false);
}

int dts_init_guide(dts_handle* h, int cs, const void* rgba) {
  if (!h || !rgba) return DTS_ERR_INVALID_ARG;
  const int rc = upload_guide(h, cs, rgba, false);
  if (rc) return rc;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int dts_init_frame_device(dts_handle* h, int cs, const void* rgba, const float* t,
                          const float* c) {
  if (!h) return DTS_ERR_INVALID_ARG;
  return init_frame_impl(h, cs, rgba, t, c, true);
}

int dts_init_right_frame(dts_handle* h, float doff, int cs, const void* rgba) {
  if (!h || !rgba) return DTS_ERR_INVALID_ARG;
  if (!(h->flags & DTS_FLAG_STEREO)) {
    set_error("handle was created without DTS_FLAG_STEREO");
    return DTS_ERR_STATE;
  }
  if (cs < DTS_COLOR_RGB || cs > DTS_COLOR_YYY) {
    set_error("unknown colour space %d", cs);
    return DTS_ERR_INVALID_ARG;
  }
  h->doff = doff;
  const size_t n = plane(h);
  int rc;
  if ((rc = upload(h, h->rcolor, rgba, n * sizeof(uchar4), false))) return rc;
  if (cs != DTS_COLOR_RGB) {
    k_convert_color<<<(unsigned)((n + 255) / 256), 256, 0, h->stream>>>(h->rcolor, (int)n, cs);
    DTS_LAUNCH_CHECK();
  }
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int dts_process_right_frame(dts_handle* h) {
  if (!h) return DTS_ERR_INVALID_ARG;
  if (!(h->flags & DTS_FLAG_STEREO)) {
    set_error("handle was created without DTS_FLAG_STEREO");
    return DTS_ERR_STATE;
  }
  // src/domain_transform_stereo.cu:24-27: right dH with the solver's params.
  const float qx = h->fp.sigma_x / h->fp.sigma_r;
  k_diff_h<<<dim3(div_up(h->w, 256), h->h), 256, 0, h->stream>>>(h->rcolor, h->w, h->h,
                                                                (double)qx, h->rdHf);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int dts_compute_differential(dts_handle* h, const dts_filter_params* p) {
  const TraceRange trace_("dts_compute_differential");
  if (!h) return DTS_ERR_INVALID_ARG;
  int rc;
  if ((rc = check_params(p))) return rc;
  // The differential is folded into the ct scans; record the params.
  h->fp = *p;
  h->fp_set = true;
  h->dp = *p;
  return DTS_OK;
}

int dts_integrate(dts_handle* h) {
  const TraceRange trace_("dts_integrate");
  if (!h) return DTS_ERR_INVALID_ARG;
  if (!h->lut) return launch_ct(h, &h->dp);
  set_ct_lazy(h, &h->dp);
  return DTS_OK;
}

int dts_filter_device(dts_handle* h, const dts_filter_params* p, const float* in,
                      float* out, int num_iter) {
  if (!h || !in || !out) return DTS_ERR_INVALID_ARG;
  int rc;
  if ((rc = check_params(p))) return rc;
  return run_filter(h, p, in, out, num_iter);
}

int dts_filter(dts_handle* h, const dts_filter_params* p, const float* in, float* out,
               int num_iter) {
  const TraceRange trace_("dts_filter");
  if (!h || !in || !out) return DTS_ERR_INVALID_ARG;
  int rc;
  if ((rc = check_params(p))) return rc;
  const size_t n = plane(h);
  // Stage the input in csqb (a scratch plane), filter into vA.
  if ((rc = upload(h, h->csqb, in, n * sizeof(float), false))) return rc;
  if ((rc = run_filter(h, p, h->csqb, h->vA, num_iter))) return rc;
  DTS_HIP_CHECK(hipMemcpyAsync(out, h->vA, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int dts_optimize(dts_handle* h, const dts_optimize_params* p, float overwrite) {
  const TraceRange trace_("dts_optimize");
  if (!h || !p) return DTS_ERR_INVALID_ARG;
  struct {
    dts_optimize_params p;
    float ow;
    int kind;
  } key{*p, overwrite, 1};
  int rc;
  if (h->graph_mode && ((rc = wide_prepare(h, p->sigma_z)) ||
                        (h->ct_of_guide && (rc = lut_prepare(h, &h->ctp)))))
    return rc;
  return graph_or_direct(h, &key, sizeof(key),
                         [&]() { return run_optimize(h, p, overwrite, false); });
}

int dts_optimize_checkpointed(dts_handle* h, const dts_optimize_params* p,
                              float overwrite_target_above_conf, int ckpt_every) {
  const TraceRange trace_("dts_optimize_checkpointed");
  if (!h) return DTS_ERR_INVALID_ARG;
  return run_optimize_ckpt(h, p, overwrite_target_above_conf, ckpt_every);
}

int dts_optimize_adjoint(dts_handle* h, const float* ybar, float* tbar, float* cbar,
                         float* gbar, int device_ptrs) {
  const TraceRange trace_("dts_optimize_adjoint");
  if (!h) return DTS_ERR_INVALID_ARG;
  return run_optimize_adjoint(h, ybar, tbar, cbar, gbar, device_ptrs != 0);
}

int dts_rf_filter(dts_handle* h, const dts_filter_params* p, const float* input, float* output,
                  int num_iter, int keep_tape, int device_ptrs) {
  const TraceRange trace_("dts_rf_filter");
  if (!h) return DTS_ERR_INVALID_ARG;
  return run_rf_filter(h, p, input, output, num_iter, keep_tape != 0, device_ptrs != 0);
}

int dts_rf_filter_adjoint(dts_handle* h, const float* ybar, float* xbar, float* gbar,
                          int device_ptrs) {
  const TraceRange trace_("dts_rf_filter_adjoint");
  if (!h) return DTS_ERR_INVALID_ARG;
  return run_rf_adjoint(h, ybar, xbar, gbar, device_ptrs != 0);
}

int dts_optimize_stereo(dts_handle* h, const dts_optimize_params* p, float overwrite) {
  const TraceRange trace_("dts_optimize_stereo");
  if (!h || !p) return DTS_ERR_INVALID_ARG;
  if (!(h->flags & DTS_FLAG_STEREO)) {
    set_error("handle was created without DTS_FLAG_STEREO");
    return DTS_ERR_STATE;
  }
  return run_optimize(h, p, overwrite, true);
}

int dts_compute_confidence(dts_handle* h, const dts_filter_params* cp, float sigma0,
                           int left_clear) {
  const TraceRange trace_("dts_compute_confidence");
  if (!h) return DTS_ERR_INVALID_ARG;
  return run_confidence(h, cp, sigma0, left_clear);
}

int dts_solve(dts_handle* h, const dts_solve_options* o, int compute_conf) {
  const TraceRange trace_("dts_solve");
  if (!h || !o) return DTS_ERR_INVALID_ARG;
  struct {
    dts_solve_options o;
    int cc;
    int kind;
  } key;
  std::memset(&key, 0, sizeof(key));
  key.o = *o;
  key.cc = compute_conf;
  key.kind = 2;
  if (h->graph_mode) {
    int rc;
    if ((rc = wide_prepare(h, o->optim_options.sigma_z))) return rc;
    // The Optimize's bounds use the ct of the confidence params when the
    // confidence is computed (its filter leaves them), else the filter
    // options'.
    if ((rc = lut_prepare(h, compute_conf ? &o->conf_params : &o->filter_options))) return rc;
    if (compute_conf &&
        (rc = wide_prepare(h, std::max(o->conf_params.sigma_x, o->conf_params.sigma_y))))
      return rc;
  }
  // run_solve sets the ct state itself (a lazy integration with the filter
  // options), so the state before the call does not select the launches.
  return graph_or_direct(
      h, &key, sizeof(key), [&]() { return run_solve(h, o, compute_conf); }, false);
}

int dts_download(dts_handle* h, int type, void* out) {
  if (!h || !out) return DTS_ERR_INVALID_ARG;
  const size_t n = plane(h);
  const void* src = nullptr;
  size_t bytes = n * sizeof(float);
  switch (type) {
    case DTS_IMAGE_COLOR_IMAGE:
      src = h->color;
      bytes = n * sizeof(uchar4);
      break;
    case DTS_IMAGE_DIFFERENTIAL: {
      const float qx = h->dp.sigma_x / h->dp.sigma_r;
      k_diff_h<<<dim3(div_up(h->w, 256), h->h), 256, 0, h->stream>>>(h->color, h->w, h->h,
                                                                    (double)qx, h->csqb);
      DTS_LAUNCH_CHECK();
      src = h->csqb;
      break;
    }
    case DTS_IMAGE_INTEGRAL: {
      int rc;
      if ((rc = ensure_ct(h))) return rc;
      std::vector<int64_t> tmp(n);
      DTS_HIP_CHECK(hipMemcpyAsync(tmp.data(), h->ctH, n * sizeof(int64_t),
                                   hipMemcpyDeviceToHost, h->stream));
      DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
      float* o = (float*)out;
      for (size_t i = 0; i < n; i++) o[i] = (float)((double)tmp[i] * (1.0 / 4294967296.0));
      return DTS_OK;
    }
    case DTS_IMAGE_TARGET: src = h->target; break;
    case DTS_IMAGE_CONFIDENCE: src = h->conf; break;
    case DTS_IMAGE_OPTIMIZED_QUANTITY: src = h->vA; break;
    default:
      set_error("unsupported image type %d", type);
      return DTS_ERR_INVALID_ARG;
  }
  DTS_HIP_CHECK(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return DTS_OK;
}

int dts_result_device(dts_handle* h, const float** out) {
  if (!h || !out) return DTS_ERR_INVALID_ARG;
  *out = h->vA;
  return DTS_OK;
}

int dts_debug_bounds(dts_handle* h, float sigma, int vertical, int* lower, int* upper) {
  if (!h || !lower || !upper) return DTS_ERR_INVALID_ARG;
  int rc;
  // A lazy integration finds the bounds from the guide (both directions, as
  // an Optimize does); formed ct planes are searched directly.
  if ((rc = h->ct_lazy ? launch_bounds2(h, sigma, sigma) : launch_bounds(h, sigma, vertical != 0)))
    return rc;
  const size_t n = plane(h);
  const bool nb = vertical ? h->bV_narrow : h->bH_narrow;
  std::vector<uint32_t> b(n);
  std::vector<uint16_t> b16(nb ? n : 0);
  if (nb)
    DTS_HIP_CHECK(hipMemcpyAsync(b16.data(), vertical ? h->bV : h->bH, n * 2,
                                 hipMemcpyDeviceToHost, h->stream));
  else
    DTS_HIP_CHECK(hipMemcpyAsync(b.data(), vertical ? h->bV : h->bH, n * 4,
                                 hipMemcpyDeviceToHost, h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  const int sh = nb ? 8 : 16;
  const uint32_t mask = nb ? 0xffu : 0xffffu;
  for (int y = 0; y < h->h; y++)
    for (int x = 0; x < h->w; x++) {
      const size_t o = (size_t)y * h->w + x;
      const uint32_t v = nb ? b16[o] : b[o];
      const int p = vertical ? y : x;
      lower[o] = p - (int)(v & mask);
      upper[o] = p + (int)(v >> sh);
    }
  return DTS_OK;
}

int dts_debug_ct(dts_handle* h, int vertical, int64_t* out) {
  if (!h || !out) return DTS_ERR_INVALID_ARG;
  int rc;
  if ((rc = ensure_ct(h))) return rc;
  const size_t n = plane(h);
  if (!vertical) {
    DTS_HIP_CHECK(hipMemcpyAsync(out, h->ctH, n * sizeof(int64_t), hipMemcpyDeviceToHost,
                                 h->stream));
    DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
    return DTS_OK;
  }
  std::vector<int64_t> t(n);
  DTS_HIP_CHECK(hipMemcpyAsync(t.data(), h->ctV, n * sizeof(int64_t), hipMemcpyDeviceToHost,
                               h->stream));
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  for (int y = 0; y < h->h; y++)
    for (int x = 0; x < h->w; x++) out[(size_t)y * h->w + x] = t[(size_t)x * h->h + y];
  return DTS_OK;
}

int dts_debug_stream(const float* a, const float* b, const float* c, const float* d,
                     float* out, size_t n, int reps, float* ms) {
  if (!a || !b || !c || !d || !out || (n & 3) || reps <= 0 || !ms) return DTS_ERR_INVALID_ARG;
  hipEvent_t e0, e1;
  DTS_HIP_CHECK(hipEventCreate(&e0));
  DTS_HIP_CHECK(hipEventCreate(&e1));
  const size_t n4 = n / 4;
  const unsigned grid = (unsigned)std::min<size_t>((n4 + 255) / 256, 256 * 32);
  DTS_HIP_CHECK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++) {
    k_stream5<<<grid, 256, 0, nullptr>>>((const float4*)a, (const float4*)b, (const float4*)c,
                                        (const float4*)d, (float4*)out, n4);
    DTS_LAUNCH_CHECK();
  }
  DTS_HIP_CHECK(hipEventRecord(e1, nullptr));
  DTS_HIP_CHECK(hipEventSynchronize(e1));
  DTS_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
  *ms /= reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return DTS_OK;
}

int dts_set_graph_mode(dts_handle* h, int enable) {
  if (!h) return DTS_ERR_INVALID_ARG;
  h->graph_mode = enable != 0;
  if (!h->graph_mode) drop_graph(h);
  return DTS_OK;
}

int dts_timer_start(dts_handle* h) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipEventRecord(h->t0, h->stream));
  return DTS_OK;
}

int dts_timer_stop(dts_handle* h, float* ms) {
  if (!h || !ms) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipEventRecord(h->t1, h->stream));
  DTS_HIP_CHECK(hipEventSynchronize(h->t1));
  DTS_HIP_CHECK(hipEventElapsedTime(ms, h->t0, h->t1));
  return DTS_OK;
}

int dts_profile_passes(dts_handle* h, int enable) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  for (auto* v : {&h->xev, &h->yev, &h->xyev}) {
    for (auto e : *v) (void)hipEventDestroy(e);
    v->clear();
  }
  h->prof = enable != 0;
  return DTS_OK;
}

static int sum_event_pairs(const std::vector<hipEvent_t>& v, double* ms, int* cnt) {
  double s = 0;
  for (size_t i = 0; i + 1 < v.size(); i += 2) {
    float t = 0;
    DTS_HIP_CHECK(hipEventElapsedTime(&t, v[i], v[i + 1]));
    s += t;
  }
  if (ms) *ms = s;
  if (cnt) *cnt = (int)(v.size() / 2);
  return DTS_OK;
}

int dts_profile_read(dts_handle* h, double* x_ms, int* x_count, double* y_ms, int* y_count) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  int rc;
  if ((rc = sum_event_pairs(h->xev, x_ms, x_count))) return rc;
  return sum_event_pairs(h->yev, y_ms, y_count);
}

int dts_profile_read_fused(dts_handle* h, double* ms, int* count) {
  if (!h) return DTS_ERR_INVALID_ARG;
  DTS_HIP_CHECK(hipStreamSynchronize(h->stream));
  return sum_event_pairs(h->xyev, ms, count);
}

int dts_set_fused(dts_handle* h, int enable) {
  if (!h) return DTS_ERR_INVALID_ARG;
  if (h->fuse != (enable != 0)) drop_graph(h);
  h->fuse = enable != 0;
  return DTS_OK;
}

}  // extern "C"

// ------------------------------------------------ solve without allocation
// dts_solve_noalloc: the dts_solve sequence (colour conversion, ct, bounds,
// optional confidence, Optimize) on caller-provided device buffers, with
// every intermediate plane carved from one file-scope device array, so the
// call allocates nothing.  Frames up to kNoAllocMaxPixels pixels, sides up
// to kNoAllocMaxSide.  Each device holds its own copy of the array (at its
// own address, found once per device with hipGetSymbolAddress); calls that
// overlap on one device share it and race.
namespace dtsamd {

constexpr size_t kNoAllocMaxPixels = size_t(1) << 20;
constexpr int kNoAllocMaxSide = 4096;
constexpr size_t kNoAllocAlign = 256;
// color, colorT, conf, vB, csq, csqb (4 B each), ctH, ctV (8 B each), bH, bV
// (4 B each): 48 bytes per pixel, each plane 256-byte aligned; then the
// (lower, upper) extent pairs of the vertical tiles,
// ceil(W/32) * ceil(H/32) <= P/1024 + W/32 + H/32 + 1 pairs.
constexpr size_t kNoAllocPlaneBytes(size_t bpp) {
  return (kNoAllocMaxPixels * bpp + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign;
}
constexpr size_t kNoAllocExtBytes =
    2 * sizeof(int) * (kNoAllocMaxPixels / 1024 + 2 * (kNoAllocMaxSide / 32) + 1);
constexpr size_t kNoAllocScratchBytes =
    8 * kNoAllocPlaneBytes(4) + 2 * kNoAllocPlaneBytes(8) +
    (kNoAllocExtBytes + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign;

}  // namespace dtsamd

static __device__ __align__(256) unsigned char scratch[dtsamd::kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};
int g_noalloc_cus[kNoAllocMaxDevices] = {};

// The current device's copy of `scratch` (its device address) and CU count,
// looked up on the first call on that device.
int noalloc_scratch(unsigned char** ptr, int* cus) {
  int dev = 0;
  DTS_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error("device index %d beyond %d", dev, kNoAllocMaxDevices);
    return DTS_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    DTS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    int n = 0;
    DTS_HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    if (!p) {
      set_error("no device address for the scratch array");
      return DTS_ERR_HIP;
    }
    g_noalloc_ptr[dev] = p;
    g_noalloc_cus[dev] = n;
  }
  *ptr = static_cast<unsigned char*>(g_noalloc_ptr[dev]);
  *cus = g_noalloc_cus[dev];
  return DTS_OK;
}

// Whether a window radius sigma needs the wide-window path (which allocates
// its prefix planes) on a w x h frame; the same test as wide_prepare.
bool needs_wide(int w, int h, float sigma) {
  const int Ex = sigma_extent(sigma, w), Ey = sigma_extent(sigma, h);
  const int segx = kXSeg + std::min(Ex + 1, w) + std::min(Ex, w - 1);
  const bool wx = div_up(std::min(segx - 1, w), 64) > 8;
  const int segy = kYRows + std::min(Ey + 1, h) + std::min(Ey, h - 1);
  const bool wy = sizeof(float) * ((size_t)segy * kYStride + 4 * kYCols) > kLdsBudget;
  return wx || wy;
}

}  // namespace

extern "C" {

int dts_solve_noalloc(int w, int h, int cs, const void* rgba_dev, const float* target_dev,
                      const float* conf_dev, const dts_solve_options* o, int compute_conf,
                      float* out_dev, void* stream) {
  const TraceRange trace_("dts_solve_noalloc");
  if (!rgba_dev || !target_dev || !o || !out_dev) {
    set_error("null guide, target, options or output");
    return DTS_ERR_INVALID_ARG;
  }
  if (w <= 0 || h <= 0 || w > kNoAllocMaxSide || h > kNoAllocMaxSide ||
      (size_t)w * h > kNoAllocMaxPixels) {
    set_error("dts_solve_noalloc: %d x %d exceeds its scratch (at most %zu pixels, sides <= %d)",
              w, h, kNoAllocMaxPixels, kNoAllocMaxSide);
    return DTS_ERR_INVALID_DIMS;
  }
  if (needs_wide(w, h, o->optim_options.sigma_z) ||
      (compute_conf && needs_wide(w, h, std::max(o->conf_params.sigma_x, o->conf_params.sigma_y)))) {
    set_error("dts_solve_noalloc: windows wider than the LDS segments need the allocating "
              "wide-window path; use dts_solve");
    return DTS_ERR_UNSUPPORTED;
  }
  unsigned char* base = nullptr;
  int cus = 0;
  int rc;
  if ((rc = noalloc_scratch(&base, &cus))) return rc;
  // A handle whose planes are views: caller buffers for the target and the
  // result, the scratch array for the rest.  Nothing below allocates.
  dts_handle hs;
  hs.max_w = hs.w = w;
  hs.max_h = hs.h = h;
  hs.stream = static_cast<hipStream_t>(stream);
  hs.num_cus = cus;
  const size_t n = (size_t)w * h;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    void* p = base + off;
    off += (bytes + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign;
    return p;
  };
  hs.color = static_cast<uchar4*>(carve(n * 4));
  hs.colorT = static_cast<uchar4*>(carve(n * 4));
  hs.conf = static_cast<float*>(carve(n * 4));
  hs.vB = static_cast<float*>(carve(n * 4));
  hs.csq = static_cast<float*>(carve(n * 4));
  hs.csqb = static_cast<float*>(carve(n * 4));
  hs.ctH = static_cast<int64_t*>(carve(n * 8));
  hs.ctV = static_cast<int64_t*>(carve(n * 8));
  hs.bH = static_cast<uint32_t*>(carve(n * 4));
  hs.bV = static_cast<uint32_t*>(carve(n * 4));
  hs.yext = static_cast<int*>(carve(2 * sizeof(int) * (size_t)div_up(w, 32) * div_up(h, 32)));
  if (off > kNoAllocScratchBytes) {
    set_error("dts_solve_noalloc: scratch layout overflow");
    return DTS_ERR_STATE;
  }
  hs.target = const_cast<float*>(target_dev);
  hs.vA = out_dev;
  if ((rc = upload_guide(&hs, cs, rgba_dev, true))) return rc;
  if (conf_dev)
    rc = launch_copy(&hs, conf_dev, hs.conf, n);
  else
    rc = launch_fill(&hs, hs.conf, n, 1.0f);
  if (rc) return rc;
  return run_solve(&hs, o, compute_conf);
}

}  // extern "C"
