/*
 * FP64 oracle of the recursive-filter domain transform (RF-DT) and its
 * adjoint.  TEST INFRASTRUCTURE ONLY: the product library never links or
 * calls this file.
 *
 * RF-DT is not in the reference, which filters with the box (normalized
 * convolution) domain transform only.  It is the recursive variant of the
 * domain transform (Gastal and Oliveira, "Domain Transform for Edge-Aware
 * Image and Video Processing", SIGGRAPH 2011, section 4 and eq. 21) built on
 * the reference's own differential (src/domain_transform_common_kernels.cuh:
 * 71-116, fp32 sigma ratio of src/domain_transform_filter.cu:96-97):
 *
 *   d(x)   = sqrt(1 + q^2 * kINV255Sq * sum_c (G_c(x) - G_c(x-1))^2),
 *            the step of the transformed domain between x-1 and x;
 *   sigma_i = sigma * sqrt(3) * 2^(N-1-i) / sqrt(4^N - 1), i = 0 .. N-1;
 *   w(x)   = a_i^d(x) = exp(-sqrt(2) d(x) / sigma_i);
 *   per line, causal then anticausal:
 *     y_0 = x_0,          y_k = (1 - w_k) x_k + w_k y_{k-1};
 *     z_{L-1} = y_{L-1},  z_k = (1 - w_{k+1}) y_k + w_{k+1} z_{k+1};
 *   iteration i filters every row with sigma_x (d along x), then every
 *   column with sigma_y (d along y).
 *
 * The guide G is taken as real RGBA values (the colour-converted guide of
 * InitFrame for 8-bit input), so that its derivative is defined: through
 * w = exp(-sqrt(2) d / sigma_i) and d, every weight depends smoothly on the
 * two pixels it joins.  The adjoint returns d<ybar, out>/d input and
 * d<ybar, out>/d G.
 */
#include "dts_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const float kINV255SQ_RF = 0.0000153787f; /* as kINV255SQ_F */

double orc_rf_sigma_i(double sigma, int i, int n) {
  return sigma * sqrt(3.0) * pow(2.0, (double)(n - 1 - i)) / sqrt(pow(4.0, (double)n) - 1.0);
}

/* d along rows (dH, x >= 1) and columns (dV, y >= 1); entries at x = 0 /
 * y = 0 are unused and set to 1. */
static void rf_diffs(const double* g, int w, int h, float sx, float sy, float sr,
                     double* dH, double* dV) {
  const float qx_f = sx / sr, qy_f = sy / sr;
  const double qx = qx_f, qy = qy_f, K = kINV255SQ_RF;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const double* c = g + 4 * ((size_t)y * w + x);
      double sh = 0, sv = 0;
      for (int k = 0; k < 4; k++) {
        if (x > 0) sh += (c[k] - c[k - 4]) * (c[k] - c[k - 4]);
        if (y > 0) sv += (c[k] - c[k - 4 * (size_t)w]) * (c[k] - c[k - 4 * (size_t)w]);
      }
      dH[(size_t)y * w + x] = x > 0 ? sqrt(sh * K * qx * qx + 1) : 1.0;
      dV[(size_t)y * w + x] = y > 0 ? sqrt(sv * K * qy * qy + 1) : 1.0;
    }
}

/* One line: x (input), wt (weights, wt[0] unused) -> z; y is scratch. */
static void rf_line(const double* x, const double* wt, double* y, double* z, int L) {
  y[0] = x[0];
  for (int k = 1; k < L; k++) y[k] = (1 - wt[k]) * x[k] + wt[k] * y[k - 1];
  z[L - 1] = y[L - 1];
  for (int k = L - 2; k >= 0; k--) z[k] = (1 - wt[k + 1]) * y[k] + wt[k + 1] * z[k + 1];
}

/* Reverse mode of rf_line: zb (cotangent of z) -> xb, and wb += d<zb, z>/dw.
 * y, z, yb, zc are scratch of length L. */
static void rf_line_adj(const double* x, const double* wt, const double* zbar, double* xb,
                        double* wb, double* y, double* z, double* yb, double* zc, int L) {
  rf_line(x, wt, y, z, L);
  memcpy(zc, zbar, sizeof(double) * (size_t)L);
  for (int k = 0; k < L; k++) yb[k] = 0;
  /* anticausal recursion, in reverse of its computation order */
  for (int k = 0; k <= L - 2; k++) {
    yb[k] += (1 - wt[k + 1]) * zc[k];
    wb[k + 1] += zc[k] * (z[k + 1] - y[k]);
    zc[k + 1] += wt[k + 1] * zc[k];
  }
  yb[L - 1] += zc[L - 1];
  /* causal recursion */
  for (int k = 0; k < L; k++) xb[k] = 0;
  for (int k = L - 1; k >= 1; k--) {
    xb[k] += (1 - wt[k]) * yb[k];
    wb[k] += yb[k] * (y[k - 1] - x[k]);
    yb[k - 1] += wt[k] * yb[k];
  }
  xb[0] += yb[0];
}

void orc_rf_line(const double* x, const double* wt, double* z, int L) {
  double* y = malloc(sizeof(double) * (size_t)L);
  if (!y) return;
  rf_line(x, wt, y, z, L);
  free(y);
}

typedef struct {
  int w, h, n;
  double *dH, *dV;
  double *wl, *xl, *zl, *y, *z, *yb, *zc, *xbl, *wbl; /* line scratch */
} rf_ctx;

static int rf_ctx_init(rf_ctx* c, const double* g, int w, int h, float sx, float sy, float sr,
                       int n) {
  const int L = w > h ? w : h;
  c->w = w;
  c->h = h;
  c->n = n;
  c->dH = malloc(sizeof(double) * (size_t)w * h);
  c->dV = malloc(sizeof(double) * (size_t)w * h);
  double* s = malloc(sizeof(double) * (size_t)L * 9);
  if (!c->dH || !c->dV || !s) return -1;
  c->wl = s;
  c->xl = s + L;
  c->zl = s + 2 * L;
  c->y = s + 3 * L;
  c->z = s + 4 * L;
  c->yb = s + 5 * L;
  c->zc = s + 6 * L;
  c->xbl = s + 7 * L;
  c->wbl = s + 8 * L;
  rf_diffs(g, w, h, sx, sy, sr, c->dH, c->dV);
  return 0;
}

static void rf_ctx_free(rf_ctx* c) {
  free(c->dH);
  free(c->dV);
  free(c->wl);
}

/* One pass of iteration i over all rows (vertical = 0) or columns, in place. */
static void rf_pass(rf_ctx* c, double* v, int vertical, double sigma_i) {
  const int w = c->w, h = c->h;
  const int nl = vertical ? w : h, L = vertical ? h : w;
  const size_t st = vertical ? (size_t)w : 1, ls = vertical ? 1 : (size_t)w;
  const double* d = vertical ? c->dV : c->dH;
  for (int l = 0; l < nl; l++) {
    for (int k = 0; k < L; k++) {
      const size_t i = (size_t)l * ls + (size_t)k * st;
      c->xl[k] = v[i];
      c->wl[k] = exp(-sqrt(2.0) * d[i] / sigma_i);
    }
    rf_line(c->xl, c->wl, c->y, c->zl, L);
    for (int k = 0; k < L; k++) v[(size_t)l * ls + (size_t)k * st] = c->zl[k];
  }
}

/* Reverse mode of rf_pass: vb (cotangent of the pass output) becomes the
 * cotangent of its input x; db += d<vb, out>/dd. */
static void rf_pass_adj(rf_ctx* c, const double* x, double* vb, double* db, int vertical,
                        double sigma_i) {
  const int w = c->w, h = c->h;
  const int nl = vertical ? w : h, L = vertical ? h : w;
  const size_t st = vertical ? (size_t)w : 1, ls = vertical ? 1 : (size_t)w;
  const double* d = vertical ? c->dV : c->dH;
  const double dlna = -sqrt(2.0) / sigma_i; /* dw/dd = w * dlna */
  for (int l = 0; l < nl; l++) {
    for (int k = 0; k < L; k++) {
      const size_t i = (size_t)l * ls + (size_t)k * st;
      c->xl[k] = x[i];
      c->wl[k] = exp(dlna * d[i]);
      c->zl[k] = vb[i];
      c->wbl[k] = 0;
    }
    rf_line_adj(c->xl, c->wl, c->zl, c->xbl, c->wbl, c->y, c->z, c->yb, c->zc, L);
    for (int k = 0; k < L; k++) {
      const size_t i = (size_t)l * ls + (size_t)k * st;
      vb[i] = c->xbl[k];
      if (k > 0) db[i] += c->wbl[k] * c->wl[k] * dlna;
    }
  }
}

int orc_rf_filter(const double* guide, int w, int h, float sx, float sy, float sr, int n,
                  const double* in, double* out) {
  if (w <= 0 || h <= 0 || n < 1) return -1;
  rf_ctx c;
  if (rf_ctx_init(&c, guide, w, h, sx, sy, sr, n)) return -2;
  memcpy(out, in, sizeof(double) * (size_t)w * h);
  for (int i = 0; i < n; i++) {
    rf_pass(&c, out, 0, orc_rf_sigma_i(sx, i, n));
    rf_pass(&c, out, 1, orc_rf_sigma_i(sy, i, n));
  }
  rf_ctx_free(&c);
  return 0;
}

int orc_rf_filter_adjoint(const double* guide, int w, int h, float sx, float sy, float sr,
                          int n, const double* in, const double* ybar, double* xbar,
                          double* gbar) {
  if (w <= 0 || h <= 0 || n < 1) return -1;
  const size_t P = (size_t)w * h;
  rf_ctx c;
  if (rf_ctx_init(&c, guide, w, h, sx, sy, sr, n)) return -2;
  /* forward, keeping the input of every pass */
  double* tape = malloc(sizeof(double) * P * 2 * (size_t)n);
  double* dbH = calloc(P, sizeof(double));
  double* dbV = calloc(P, sizeof(double));
  double* v = malloc(sizeof(double) * P);
  if (!tape || !dbH || !dbV || !v) return -2;
  memcpy(v, in, sizeof(double) * P);
  for (int i = 0; i < n; i++) {
    memcpy(tape + P * (2 * (size_t)i), v, sizeof(double) * P);
    rf_pass(&c, v, 0, orc_rf_sigma_i(sx, i, n));
    memcpy(tape + P * (2 * (size_t)i + 1), v, sizeof(double) * P);
    rf_pass(&c, v, 1, orc_rf_sigma_i(sy, i, n));
  }
  memcpy(xbar, ybar, sizeof(double) * P);
  for (int i = n - 1; i >= 0; i--) {
    rf_pass_adj(&c, tape + P * (2 * (size_t)i + 1), xbar, dbV, 1, orc_rf_sigma_i(sy, i, n));
    rf_pass_adj(&c, tape + P * (2 * (size_t)i), xbar, dbH, 0, orc_rf_sigma_i(sx, i, n));
  }
  if (gbar) {
    /* d = sqrt(1 + q^2 K S): dd/dG_c(x) = q^2 K (G_c(x) - G_c(x-1)) / d, and
     * the negative of it for G_c(x-1). */
    const float qx_f = sx / sr, qy_f = sy / sr;
    const double kx = (double)qx_f * (double)qx_f * (double)kINV255SQ_RF;
    const double ky = (double)qy_f * (double)qy_f * (double)kINV255SQ_RF;
    memset(gbar, 0, sizeof(double) * 4 * P);
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        const size_t i = (size_t)y * w + x;
        const double* g = guide + 4 * i;
        for (int k = 0; k < 4; k++) {
          if (x > 0) {
            const double t = dbH[i] * kx * (g[k] - g[k - 4]) / c.dH[i];
            gbar[4 * i + k] += t;
            gbar[4 * (i - 1) + k] -= t;
          }
          if (y > 0) {
            const double t = dbV[i] * ky * (g[k] - g[k - 4 * (size_t)w]) / c.dV[i];
            gbar[4 * i + k] += t;
            gbar[4 * (i - (size_t)w) + k] -= t;
          }
        }
      }
  }
  free(tape);
  free(dbH);
  free(dbV);
  free(v);
  rf_ctx_free(&c);
  return 0;
}
