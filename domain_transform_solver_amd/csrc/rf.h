// Recursive-filter domain transform (RF-DT) and its adjoint: launchers over
// plain device pointers, used by the handle code in dts.hip.  Not in the
// reference (which has the box domain transform only): the recursive variant
// of Gastal and Oliveira (2011) on the reference's differential.
//
// Per iteration i of N, rows then columns: weights w = exp(-sqrt(2) d /
// sigma_i) with d the transformed-domain step between neighbours and
// sigma_i = sigma sqrt(3) 2^(N-1-i) / sqrt(4^N - 1); per line the causal
// recursion y_k = (1 - w_k) x_k + w_k y_(k-1) (y_0 = x_0), then the
// anticausal z_k = (1 - w_(k+1)) y_k + w_(k+1) z_(k+1) (z_(L-1) = y_(L-1)).
// Columns run as rows of transposed planes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtsamd {

// d of every line of a line-major guide ([nlines][L] uchar4): entry k joins
// k - 1 and k (entry 0 is unused and set to 1); q = sigma_s / sigma_r.
int rf_diff(hipStream_t s, const uchar4* guide, int L, int nlines, float q, float* d);

// One pass over every line of a line-major plane: in -> out (may alias).
// c2 = -sqrt(2) log2(e) / sigma_i, so that w = exp2(c2 d).
int rf_rows(hipStream_t s, const float* in, float* out, const float* d, int L, int nlines,
            float c2);

// Reverse mode of rf_rows for the pass input x: the cotangent of the pass
// output in gz becomes the cotangent of x (in place), and db += d<gz, out>/dd.
// ytmp, gtmp: scratch planes of the same shape.  dlna = -sqrt(2) / sigma_i.
int rf_rows_adj(hipStream_t s, const float* x, float* gz, float* ytmp, float* gtmp,
                const float* d, float* db, int L, int nlines, float c2, float dlna);

// out[x][y] = in[y][x] for a W x H plane.
int rf_transpose(hipStream_t s, const float* in, float* out, int W, int H);

// Guide cotangent (4 floats per pixel of the W x H guide) from the weight
// cotangents and steps along rows (db_h, d_h) and columns (db_v, d_v), all
// [H][W]; kx = qx^2 kINV255Sq, ky = qy^2 kINV255Sq.
int rf_guide_grad(hipStream_t s, const uchar4* guide, int W, int H, const float* d_h,
                  const float* db_h, const float* d_v, const float* db_v, float kx, float ky,
                  float* gbar);

}  // namespace dtsamd
