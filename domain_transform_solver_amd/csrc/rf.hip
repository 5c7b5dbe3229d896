// Recursive-filter domain transform (RF-DT) kernels and their adjoint; see
// rf.h for the recursion.  One wavefront per line: each 64-position chunk
// of a recursion is one DPP scan of affine maps (wave_linrec_scan_dpp) plus
// the carry of the previous chunk, so a line costs ceil(L / 64) dependent
// scans instead of L dependent steps.  Lines of up to 2048 positions stay in
// registers; longer ones make two sweeps through global memory.  Columns run
// as rows of transposed planes (rf_transpose), so every access is a
// contiguous line segment.
#include "common.h"
#include "dts_kernels.h"
#include "rf.h"

namespace dtsamd {

namespace {

__device__ inline float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Orders the global stores of the workgroup's first sweep before the loads
// of its second (the lanes of a wavefront read each other's positions).
__device__ inline void sweep_sync() {
  __threadfence_block();
  __syncthreads();
}

}  // namespace

// ---- lines of up to 64 NC positions held in registers ---------------------
// Chunk c holds positions 64 c + lane.  All loads of the line are issued
// before the first scan; the causal recursion scans chunks upwards, the
// anticausal one downwards with the lanes reversed (ds_bpermute), so both
// are forward scans of affine maps.  No global round trip between sweeps.

// The value at position 64 c + lane + 1 (lane 63: lane 0 of chunk c + 1).
template <int NC>
__device__ __forceinline__ float next_pos(const float (&v)[NC], int c, int lane) {
  const float n = __shfl_down(v[c], 1, 64);
  const float n0 = c + 1 < NC ? readlane_f(v[c + 1 < NC ? c + 1 : c], 0) : 0.f;
  return lane == 63 ? n0 : n;
}

// The value at position 64 c + lane - 1 (lane 0: lane 63 of chunk c - 1).
template <int NC>
__device__ __forceinline__ float prev_pos(const float (&v)[NC], int c, int lane) {
  const float p = __shfl_up(v[c], 1, 64);
  const float p63 = c > 0 ? readlane_f(v[c > 0 ? c - 1 : 0], 63) : 0.f;
  return lane == 0 ? p63 : p;
}

__device__ __forceinline__ float rev(float v, int lane) { return __shfl(v, 63 - lane, 64); }

template <int NC>
__device__ __forceinline__ void rf_line_reg(const float* xi, float* zo,
                                            const float* __restrict__ dl, int L, float c2,
                                            int lane) {
  float v[NC], w[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int k = 64 * c + lane;
    v[c] = k < L ? xi[k] : 0.f;
    w[c] = k < L ? dl[k] : 0.f;
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int k = 64 * c + lane;
    w[c] = k > 0 && k < L ? exp2f(c2 * w[c]) : 0.f;
  }
  float carry = 0.f;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    if (64 * c >= L) break;
    const int k = 64 * c + lane;
    float a = k < L ? w[c] : 1.f;
    float b = k < L ? (k > 0 ? (1.f - w[c]) * v[c] : v[c]) : 0.f;
    wave_linrec_scan_dpp(a, b);
    v[c] = a * carry + b;
    carry = readlane_f(v[c], 63);
  }
  carry = 0.f;
#pragma unroll
  for (int c = NC - 1; c >= 0; c--) {
    if (64 * c >= L) continue;
    const int k = 64 * c + lane;
    const float wn = next_pos<NC>(w, c, lane);
    const bool last = k == L - 1;
    float a = k < L ? (last ? 0.f : wn) : 1.f;
    float b = k < L ? (last ? v[c] : (1.f - wn) * v[c]) : 0.f;
    a = rev(a, lane);
    b = rev(b, lane);
    wave_linrec_scan_dpp(a, b);
    const float zr = a * carry + b;
    carry = readlane_f(zr, 63);
    v[c] = rev(zr, lane);
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int k = 64 * c + lane;
    if (k < L) zo[k] = v[c];
  }
}

template <int NC>
__device__ __forceinline__ void rf_line_adj_reg(const float* __restrict__ xl,
                                                float* __restrict__ gl,
                                                const float* __restrict__ dl,
                                                float* __restrict__ bl, int L, float c2,
                                                float dlna, int lane) {
  float x[NC], w[NC], y[NC], gp[NC], dbv[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int k = 64 * c + lane;
    x[c] = k < L ? xl[k] : 0.f;
    w[c] = k < L ? dl[k] : 0.f;
    gp[c] = k < L ? gl[k] : 0.f;
    dbv[c] = k < L ? bl[k] : 0.f;
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int k = 64 * c + lane;
    w[c] = k > 0 && k < L ? exp2f(c2 * w[c]) : 0.f;
  }
  // forward sweep: y = causal(x), g'_k = g_k + w_k g'_(k-1)
  float cy = 0.f, cg = 0.f;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    if (64 * c >= L) break;
    const int k = 64 * c + lane;
    float a = k < L ? w[c] : 1.f;
    float b = k < L ? (k > 0 ? (1.f - w[c]) * x[c] : x[c]) : 0.f;
    float b2 = k < L ? gp[c] : 0.f;
    wave_linrec_scan2_dpp(a, b, b2);
    y[c] = a * cy + b;
    gp[c] = a * cg + b2;
    cy = readlane_f(y[c], 63);
    cg = readlane_f(gp[c], 63);
  }
  // backward sweep (lanes reversed): z = anticausal(y), y' from ybar
  float cz = 0.f, cyp = 0.f;
#pragma unroll
  for (int c = NC - 1; c >= 0; c--) {
    if (64 * c >= L) continue;
    const int k = 64 * c + lane;
    const bool ok = k < L, last = k == L - 1, first = k == 0;
    const float w1 = next_pos<NC>(w, c, lane);
    const float ym1 = prev_pos<NC>(y, c, lane);
    const float gm1 = prev_pos<NC>(gp, c, lane);
    float a = ok ? (last ? 0.f : w1) : 1.f;
    float b = ok ? (last ? y[c] : (1.f - w1) * y[c]) : 0.f;
    const float ybar = last ? gp[c] : (1.f - w1) * gp[c];
    float b2 = ok ? ybar : 0.f;
    a = rev(a, lane);
    b = rev(b, lane);
    b2 = rev(b2, lane);
    wave_linrec_scan2_dpp(a, b, b2);
    const float zr = a * cz + b, ypr = a * cyp + b2;
    cz = readlane_f(zr, 63);
    cyp = readlane_f(ypr, 63);
    const float z = rev(zr, lane), yp = rev(ypr, lane);
    if (ok) gl[k] = first ? yp : (1.f - w[c]) * yp;
    if (!first) dbv[c] += (gm1 * (z - ym1) + yp * (ym1 - x[c])) * w[c] * dlna;
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int k = 64 * c + lane;
    if (k < L) bl[k] = dbv[c];
  }
}

template <int NC>
__global__ __launch_bounds__(256) void k_rf_rows_reg(const float* __restrict__ in,
                                                     float* __restrict__ out,
                                                     const float* __restrict__ d, int L,
                                                     int nlines, float c2) {
  const int line = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (line >= nlines) return;
  const size_t o = (size_t)line * (size_t)L;
  rf_line_reg<NC>(in + o, out + o, d + o, L, c2, threadIdx.x & 63);
}

template <int NC>
__global__ __launch_bounds__(256) void k_rf_rows_adj_reg(const float* __restrict__ x,
                                                         float* __restrict__ gz,
                                                         const float* __restrict__ d,
                                                         float* __restrict__ db, int L,
                                                         int nlines, float c2, float dlna) {
  const int line = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (line >= nlines) return;
  const size_t o = (size_t)line * (size_t)L;
  rf_line_adj_reg<NC>(x + o, gz + o, d + o, db + o, L, c2, dlna, threadIdx.x & 63);
}

// d of the transformed domain between positions k - 1 and k of every line,
// in double as the box transform's differential, rounded to fp32.
__global__ __launch_bounds__(256) void k_rf_diff(const uchar4* __restrict__ guide, int L,
                                                 int nlines, double q, float* __restrict__ d) {
  const size_t n = (size_t)L * nlines;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % (size_t)L);
    d[i] = k == 0 ? 1.f : (float)dt_diff(guide[i], guide[i - 1], q);
  }
}

// One RF pass over lines of length L: causal sweep into out, then the
// anticausal sweep in place.  Four lines per workgroup, one per wavefront.
__global__ __launch_bounds__(256) void k_rf_rows(const float* __restrict__ in,
                                                 float* __restrict__ out,
                                                 const float* __restrict__ d, int L, int nlines,
                                                 float c2) {
  const int lane = threadIdx.x & 63;
  const int line = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const bool live = line < nlines;  // every wavefront reaches the barrier
  const size_t o = (size_t)(live ? line : 0) * (size_t)L;
  const float* xi = in + o;
  const float* dl = d + o;
  float* zo = out + o;
  if (live) {
    float carry = 0.f;
    for (int c0 = 0; c0 < L; c0 += 64) {
      const int k = c0 + lane;
      float a = 1.f, b = 0.f;
      if (k < L) {
        const float x = xi[k];
        const float w = k > 0 ? exp2f(c2 * dl[k]) : 0.f;
        a = w;
        b = k > 0 ? (1.f - w) * x : x;
      }
      wave_linrec_scan_dpp(a, b);
      const float y = a * carry + b;
      if (k < L) zo[k] = y;
      carry = readlane_f(y, 63);
    }
  }
  sweep_sync();
  if (live) {
    float carry = 0.f;
    for (int c0 = 0; c0 < L; c0 += 64) {
      const int k = L - 1 - (c0 + lane);  // descending positions
      float a = 1.f, b = 0.f;
      if (k >= 0) {
        const float y = zo[k];
        const float w = k < L - 1 ? exp2f(c2 * dl[k + 1]) : 0.f;
        a = w;
        b = k < L - 1 ? (1.f - w) * y : y;
      }
      wave_linrec_scan_dpp(a, b);
      const float z = a * carry + b;
      if (k >= 0) zo[k] = z;
      carry = readlane_f(z, 63);
    }
  }
}

// Reverse mode of one RF pass for its input x.  With the cotangent g of the
// pass output z:
//   forward sweep:  y = causal(x) and g'_k = g_k + w_k g'_(k-1);
//   backward sweep: z = anticausal(y) again, ybar_k = (1 - w_(k+1)) g'_k
//   (g'_(L-1) at the end), y'_k = ybar_k + w_(k+1) y'_(k+1),
//   xbar_k = (1 - w_k) y'_k (y'_0 at 0), and the weight cotangent
//   wbar_k = g'_(k-1) (z_k - y_(k-1)) + y'_k (y_(k-1) - x_k), which enters
//   db_k += wbar_k w_k dlna (dw/dd = w ln a).
__global__ __launch_bounds__(256) void k_rf_rows_adj(const float* __restrict__ x,
                                                     float* __restrict__ gz,
                                                     float* __restrict__ ytmp,
                                                     float* __restrict__ gtmp,
                                                     const float* __restrict__ d,
                                                     float* __restrict__ db, int L, int nlines,
                                                     float c2, float dlna) {
  const int lane = threadIdx.x & 63;
  const int line = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const bool live = line < nlines;
  const size_t o = (size_t)(live ? line : 0) * (size_t)L;
  const float* xl = x + o;
  const float* dl = d + o;
  float* gl = gz + o;
  float* yl = ytmp + o;
  float* pl = gtmp + o;
  float* bl = db + o;
  if (live) {
    float cy = 0.f, cg = 0.f;
    for (int c0 = 0; c0 < L; c0 += 64) {
      const int k = c0 + lane;
      float a = 1.f, b = 0.f, a2 = 1.f, b2 = 0.f;
      if (k < L) {
        const float xv = xl[k];
        const float w = k > 0 ? exp2f(c2 * dl[k]) : 0.f;
        a = w;
        b = k > 0 ? (1.f - w) * xv : xv;
        a2 = w;
        b2 = gl[k];
      }
      wave_linrec_scan_dpp(a, b);
      wave_linrec_scan_dpp(a2, b2);
      const float y = a * cy + b, gp = a2 * cg + b2;
      if (k < L) {
        yl[k] = y;
        pl[k] = gp;
      }
      cy = readlane_f(y, 63);
      cg = readlane_f(gp, 63);
    }
  }
  sweep_sync();
  if (live) {
    float cz = 0.f, cyp = 0.f;
    for (int c0 = 0; c0 < L; c0 += 64) {
      const int k = L - 1 - (c0 + lane);
      const bool ok = k >= 0;
      const int kc = ok ? k : 0;
      const float xv = xl[kc], yv = yl[kc], gp = pl[kc];
      const bool last = kc == L - 1, first = kc == 0;
      const float w1 = last ? 0.f : exp2f(c2 * dl[kc + 1]);
      const float w0 = first ? 0.f : exp2f(c2 * dl[kc]);
      const float ym1 = first ? 0.f : yl[kc - 1];
      const float gm1 = first ? 0.f : pl[kc - 1];
      float a = ok ? w1 : 1.f, b = ok ? (last ? yv : (1.f - w1) * yv) : 0.f;
      const float ybar = last ? gp : (1.f - w1) * gp;
      float a2 = ok ? w1 : 1.f, b2 = ok ? ybar : 0.f;
      wave_linrec_scan_dpp(a, b);
      wave_linrec_scan_dpp(a2, b2);
      const float z = a * cz + b, yp = a2 * cyp + b2;
      if (ok) {
        gl[k] = first ? yp : (1.f - w0) * yp;
        if (!first) {
          const float wbar = gm1 * (z - ym1) + yp * (ym1 - xv);
          bl[k] += wbar * w0 * dlna;
        }
      }
      cz = readlane_f(z, 63);
      cyp = readlane_f(yp, 63);
    }
  }
}

// 32 x 32 tiles through LDS (padded rows), 8 rows of threads: enough
// workgroups to fill the GPU at 1080p, 128-byte row segments either way.
__global__ __launch_bounds__(256) void k_rf_transpose(const float* __restrict__ in,
                                                      float* __restrict__ out, int W, int H) {
  __shared__ float tile[32][33];
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int r = ty; r < 32; r += 8) {
    const int y = y0 + r, x = x0 + tx;
    if (y < H && x < W) tile[r][tx] = in[(size_t)y * W + x];
  }
  __syncthreads();
#pragma unroll
  for (int r = ty; r < 32; r += 8) {
    const int x = x0 + r, y = y0 + tx;
    if (x < W && y < H) out[(size_t)x * H + y] = tile[tx][r];
  }
}

__global__ __launch_bounds__(256) void k_rf_guide_grad(const uchar4* __restrict__ guide, int W,
                                                       int H, const float* __restrict__ d_h,
                                                       const float* __restrict__ db_h,
                                                       const float* __restrict__ d_v,
                                                       const float* __restrict__ db_v, float kx,
                                                       float ky, float4* __restrict__ gbar) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= W || y >= H) return;
  const size_t i = (size_t)y * W + x;
  auto f4 = [](uchar4 p) { return make_float4(p.x, p.y, p.z, p.w); };
  const float4 g = f4(guide[i]);
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
  auto add = [&](float4 other, float s) {  // s * (g - other)
    r.x += s * (g.x - other.x);
    r.y += s * (g.y - other.y);
    r.z += s * (g.z - other.z);
    r.w += s * (g.w - other.w);
  };
  // d(k) joins k - 1 and k: dd/dG(k) = k q^2 K (G(k) - G(k-1)) / d(k), and
  // dd(k+1)/dG(k) = k q^2 K (G(k) - G(k+1)) / d(k+1).
  if (x > 0) add(f4(guide[i - 1]), db_h[i] * kx / d_h[i]);
  if (x + 1 < W) add(f4(guide[i + 1]), db_h[i + 1] * kx / d_h[i + 1]);
  if (y > 0) add(f4(guide[i - W]), db_v[i] * ky / d_v[i]);
  if (y + 1 < H) add(f4(guide[i + W]), db_v[i + W] * ky / d_v[i + W]);
  gbar[i] = r;
}

int rf_diff(hipStream_t s, const uchar4* guide, int L, int nlines, float q, float* d) {
  const size_t n = (size_t)L * nlines;
  const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
  k_rf_diff<<<grid, 256, 0, s>>>(guide, L, nlines, (double)q, d);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int rf_rows(hipStream_t s, const float* in, float* out, const float* d, int L, int nlines,
            float c2) {
  const int g = div_up(nlines, 4);
  if (L <= 256)
    k_rf_rows_reg<4><<<g, 256, 0, s>>>(in, out, d, L, nlines, c2);
  else if (L <= 512)
    k_rf_rows_reg<8><<<g, 256, 0, s>>>(in, out, d, L, nlines, c2);
  else if (L <= 1024)
    k_rf_rows_reg<16><<<g, 256, 0, s>>>(in, out, d, L, nlines, c2);
  else if (L <= 2048)
    k_rf_rows_reg<32><<<g, 256, 0, s>>>(in, out, d, L, nlines, c2);
  else
    k_rf_rows<<<g, 256, 0, s>>>(in, out, d, L, nlines, c2);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int rf_rows_adj(hipStream_t s, const float* x, float* gz, float* ytmp, float* gtmp,
                const float* d, float* db, int L, int nlines, float c2, float dlna) {
  const int g = div_up(nlines, 4);
  if (L <= 256)
    k_rf_rows_adj_reg<4><<<g, 256, 0, s>>>(x, gz, d, db, L, nlines, c2, dlna);
  else if (L <= 512)
    k_rf_rows_adj_reg<8><<<g, 256, 0, s>>>(x, gz, d, db, L, nlines, c2, dlna);
  else if (L <= 1024)
    k_rf_rows_adj_reg<16><<<g, 256, 0, s>>>(x, gz, d, db, L, nlines, c2, dlna);
  else if (L <= 2048)
    k_rf_rows_adj_reg<32><<<g, 256, 0, s>>>(x, gz, d, db, L, nlines, c2, dlna);
  else
    k_rf_rows_adj<<<g, 256, 0, s>>>(x, gz, ytmp, gtmp, d, db, L, nlines, c2, dlna);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int rf_transpose(hipStream_t s, const float* in, float* out, int W, int H) {
  k_rf_transpose<<<dim3(div_up(W, 32), div_up(H, 32)), 256, 0, s>>>(in, out, W, H);
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

int rf_guide_grad(hipStream_t s, const uchar4* guide, int W, int H, const float* d_h,
                  const float* db_h, const float* d_v, const float* db_v, float kx, float ky,
                  float* gbar) {
  k_rf_guide_grad<<<dim3(div_up(W, 64), div_up(H, 4)), 256, 0, s>>>(
      guide, W, H, d_h, db_h, d_v, db_v, kx, ky, reinterpret_cast<float4*>(gbar));
  DTS_LAUNCH_CHECK();
  return DTS_OK;
}

}  // namespace dtsamd
