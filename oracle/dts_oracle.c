/*
 * FP64 CPU oracle for the DTS / HFBS hot path.  TEST INFRASTRUCTURE ONLY:
 * the product library never links or calls this file.  See dts_oracle.h for
 * the boundary contract shared with the GPU kernels.
 *
 * Arithmetic: continuous quantities are double; reference constants that are
 * fp32 literals are taken as their fp32 values promoted to double.  Quantities
 * that select discrete indices (sigma ratios, grid coordinates) follow the
 * reference's fp32 arithmetic so that the chosen indices are the same.
 * Compile with -ffp-contract=off: the colour conversion and the
 * differentials must round exactly like the GPU kernels, which also disable
 * contraction for them.
 */
#include "dts_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/domain_transform_common_kernels.cuh:9-10 (fp32 literals). */
static const float kINV255_F = 0.00392156863f;
static const float kINV255SQ_F = 0.0000153787f;

#define FX_ONE 4294967296.0 /* 2^32 */

void orc_dts_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int64_t orc_to_fx(double d) { return (int64_t)nearbyint(d * FX_ONE); }

/* The bound predicate ct(i) - ct(p) > sigma evaluated on 32.32 integers:
 * D > sigma*2^32  <=>  D > floor(sigma*2^32) for integer D.  NaN and huge
 * sigma never satisfy it (the window collapses, as in the reference). */
int64_t orc_sigma_fx(float sigma) {
  if (!(sigma < 1073741824.0f)) return INT64_MAX; /* NaN, inf, >= 2^30 */
  return (int64_t)floor((double)sigma * FX_ONE);
}

/* ------------------------------------------------------------------ state */
struct orc_dts {
  int max_w, max_h, w, h;
  uint8_t* color;
  uint8_t* right_color;
  double *dH, *dV, *right_dH;
  int64_t *ctH, *ctV, *right_ctH, *right_ctV;
  double *target, *conf, *var, *conf_sq, *conf_sq_buf;
  float fp_sx, fp_sy, fp_sr; /* filter_params_ of the solver */
  float doff;
};

orc_dts* orc_dts_create(int max_w, int max_h) {
  if (max_w <= 0 || max_h <= 0) return NULL;
  orc_dts* s = (orc_dts*)calloc(1, sizeof(orc_dts));
  size_t n = (size_t)max_w * max_h;
  s->max_w = s->w = max_w;
  s->max_h = s->h = max_h;
  s->color = (uint8_t*)calloc(n * 4, 1);
  s->right_color = (uint8_t*)calloc(n * 4, 1);
  s->dH = (double*)calloc(n, sizeof(double));
  s->dV = (double*)calloc(n, sizeof(double));
  s->right_dH = (double*)calloc(n, sizeof(double));
  s->ctH = (int64_t*)calloc(n, sizeof(int64_t));
  s->ctV = (int64_t*)calloc(n, sizeof(int64_t));
  s->right_ctH = (int64_t*)calloc(n, sizeof(int64_t));
  s->right_ctV = (int64_t*)calloc(n, sizeof(int64_t));
  s->target = (double*)calloc(n, sizeof(double));
  s->conf = (double*)calloc(n, sizeof(double));
  s->var = (double*)calloc(n, sizeof(double));
  s->conf_sq = (double*)calloc(n, sizeof(double));
  s->conf_sq_buf = (double*)calloc(n, sizeof(double));
  return s;
}

void orc_dts_destroy(orc_dts* s) {
  if (!s) return;
  free(s->color); free(s->right_color); free(s->dH); free(s->dV);
  free(s->right_dH); free(s->ctH); free(s->ctV); free(s->right_ctH);
  free(s->right_ctV); free(s->target); free(s->conf); free(s->var);
  free(s->conf_sq); free(s->conf_sq_buf);
  free(s);
}

/* src/domain_transform_filter.cu:114-122: invalid dims leave state as is. */
int orc_dts_set_image_dim(orc_dts* s, int w, int h) {
  if (w > 0 && w <= s->max_w && h > 0 && h <= s->max_h) {
    s->w = w;
    s->h = h;
    return 0;
  }
  return 1;
}

/* src/domain_transform_filter.cu:202-216, src/domain_transform_solver.cu:248-255 */
int orc_dts_clear_all(orc_dts* s) {
  size_t n = (size_t)s->max_w * s->max_h;
  memset(s->color, 0, n * 4);
  memset(s->right_color, 0, n * 4);
  memset(s->dH, 0, n * sizeof(double));
  memset(s->dV, 0, n * sizeof(double));
  memset(s->right_dH, 0, n * sizeof(double));
  memset(s->ctH, 0, n * sizeof(int64_t));
  memset(s->ctV, 0, n * sizeof(int64_t));
  memset(s->right_ctH, 0, n * sizeof(int64_t));
  memset(s->right_ctV, 0, n * sizeof(int64_t));
  memset(s->target, 0, n * sizeof(double));
  memset(s->conf, 0, n * sizeof(double));
  memset(s->var, 0, n * sizeof(double));
  memset(s->conf_sq, 0, n * sizeof(double));
  memset(s->conf_sq_buf, 0, n * sizeof(double));
  return 0;
}

/* -------------------------------------------------------------- colour */
/* src/color_space.cuh:11-40: double arithmetic, truncation to uchar, alpha
 * kept.  Evaluation order follows the reference expressions. */
void orc_convert_color(uint8_t* rgba, int w, int h, int cs) {
  if (cs == ORC_RGB) return;
  long n = (long)w * h;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    uint8_t* p = rgba + 4 * i;
    double r = p[0], g = p[1], b = p[2];
    double yv = r * 0.299 + 0.587 * g + 0.114 * b;
    if (cs == ORC_YCBCR) {
      double cb = 128 - r * 0.168736 - 0.331264 * g + 0.5 * b;
      double cr = 128 + r * 0.5 - 0.418688 * g - 0.081312 * b;
      p[0] = (uint8_t)yv;
      p[1] = (uint8_t)cb;
      p[2] = (uint8_t)cr;
    } else {
      uint8_t y8 = (uint8_t)yv;
      p[0] = y8;
      p[1] = y8;
      p[2] = y8;
    }
  }
}

int orc_dts_init_frame(orc_dts* s, int cs, const uint8_t* rgba,
                       const float* target, const float* confidence) {
  /* src/domain_transform_filter.cu:18-49, src/domain_transform_solver.cu:46-61 */
  int w = s->w, h = s->h;
  size_t n = (size_t)w * h;
  memcpy(s->color, rgba, n * 4);
  orc_convert_color(s->color, w, h, cs);
  for (size_t i = 0; i < n; i++) s->target[i] = target[i];
  if (confidence) {
    for (size_t i = 0; i < n; i++) s->conf[i] = confidence[i];
  } else {
    size_t nmax = (size_t)s->max_w * s->max_h;
    for (size_t i = 0; i < nmax; i++) s->conf[i] = 1.0;
  }
  return 0;
}

/* Replaces target and confidence with double values (finite-difference
 * checks perturb them below fp32 resolution). */
int orc_dts_set_target_conf_f64(orc_dts* s, const double* target,
                                const double* confidence) {
  const size_t n = (size_t)s->w * s->h;
  memcpy(s->target, target, n * sizeof(double));
  memcpy(s->conf, confidence, n * sizeof(double));
  return 0;
}

int orc_dts_init_right_frame(orc_dts* s, float doff, int cs,
                             const uint8_t* rgba) {
  /* src/domain_transform_stereo.cu:17-22 */
  s->doff = doff;
  memcpy(s->right_color, rgba, (size_t)s->w * s->h * 4);
  orc_convert_color(s->right_color, s->w, s->h, cs);
  return 0;
}

/* ------------------------------------------------------- differentials */
/* src/domain_transform_common_kernels.cuh:71-116.  Neighbours outside the
 * image read as (0,0,0,0). */
static inline double ssd4(const uint8_t* a, const uint8_t* b) {
  int s = 0;
  for (int k = 0; k < 4; k++) {
    int d = (int)a[k] - (int)b[k];
    s += d * d;
  }
  return (double)s;
}

static void differentials(const uint8_t* color, int w, int h, float sx,
                          float sy, float sr, double* dH, double* dV) {
  /* src/domain_transform_filter.cu:92-110: the ratios are fp32. */
  const float qx_f = sx / sr;
  const float qy_f = sy / sr;
  const double qx = qx_f, qy = qy_f, K = kINV255SQ_F;
  static const uint8_t zero[4] = {0, 0, 0, 0};
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      const uint8_t* c = color + 4 * ((size_t)y * w + x);
      const uint8_t* l = x > 0 ? c - 4 : zero;
      const uint8_t* u = y > 0 ? c - 4 * (size_t)w : zero;
      double a = ssd4(c, l) * K;
      a = a * qx;
      a = a * qx;
      dH[(size_t)y * w + x] = sqrt(a + 1);
      double b = ssd4(c, u) * K;
      b = b * qy;
      b = b * qy;
      dV[(size_t)y * w + x] = sqrt(b + 1);
    }
  }
}

static void integrate(const double* dH, const double* dV, int w, int h, int64_t* ctH,
                      int64_t* ctV);

/* Differentials and ct of a real-valued guide (4 doubles per pixel), the
 * same formula as differentials() and integrate() with real channel values, for
 * finite-difference checks of the guide derivative.  For integer-valued
 * guides the result equals orc_dts_compute_differential + orc_dts_integrate. */
int orc_dts_set_guide_f64(orc_dts* s, const double* g, float sx, float sy, float sr) {
  const int w = s->w, h = s->h;
  const float qx_f = sx / sr, qy_f = sy / sr;
  const double qx = qx_f, qy = qy_f, K = kINV255SQ_F;
  s->fp_sx = sx;
  s->fp_sy = sy;
  s->fp_sr = sr;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const double* c = g + 4 * ((size_t)y * w + x);
      double sh = 0, sv = 0;
      for (int k = 0; k < 4; k++) {
        const double l = x > 0 ? c[k - 4] : 0.0, u = y > 0 ? c[k - 4 * (size_t)w] : 0.0;
        sh += (c[k] - l) * (c[k] - l);
        sv += (c[k] - u) * (c[k] - u);
      }
      double a = sh * K;
      a = a * qx;
      a = a * qx;
      s->dH[(size_t)y * w + x] = sqrt(a + 1);
      double b = sv * K;
      b = b * qy;
      b = b * qy;
      s->dV[(size_t)y * w + x] = sqrt(b + 1);
    }
  integrate(s->dH, s->dV, w, h, s->ctH, s->ctV);
  return 0;
}

int orc_dts_compute_differential(orc_dts* s, float sx, float sy, float sr) {
  /* src/domain_transform_solver.cu:89-93 stores the params for the stereo
   * right frame, then runs the filter's differential. */
  s->fp_sx = sx;
  s->fp_sy = sy;
  s->fp_sr = sr;
  differentials(s->color, s->w, s->h, sx, sy, sr, s->dH, s->dV);
  return 0;
}

/* ----------------------------------------------------------- integrate */
static void integrate(const double* dH, const double* dV, int w, int h,
                      int64_t* ctH, int64_t* ctV) {
  /* src/domain_transform_filter.cu:85-90: inclusive scans along rows (ct_H)
   * and columns (ct_V). */
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    int64_t acc = 0;
    for (int x = 0; x < w; x++) {
      acc += orc_to_fx(dH[(size_t)y * w + x]);
      ctH[(size_t)y * w + x] = acc;
    }
  }
#pragma omp parallel for schedule(static)
  for (int x = 0; x < w; x++) {
    int64_t acc = 0;
    for (int y = 0; y < h; y++) {
      acc += orc_to_fx(dV[(size_t)y * w + x]);
      ctV[(size_t)y * w + x] = acc;
    }
  }
}

int orc_dts_integrate(orc_dts* s) {
  integrate(s->dH, s->dV, s->w, s->h, s->ctH, s->ctV);
  return 0;
}

int orc_dts_process_right_frame(orc_dts* s) {
  /* src/domain_transform_stereo.cu:24-27: right differential with the
   * solver's stored filter params. */
  double* tmpV = (double*)malloc(sizeof(double) * (size_t)s->w * s->h);
  differentials(s->right_color, s->w, s->h, s->fp_sx, s->fp_sy, s->fp_sr,
                s->right_dH, tmpV);
  free(tmpV);
  return 0;
}

/* -------------------------------------------------------------- bounds */
/* src/domain_transform_filter.cuh:7-28, linear search, same break rules.
 * ct(p) - sigma > ct(i)  <=>  ct(p) - ct(i) > sigma. */
void orc_line_bounds(const int64_t* ct, int L, float sigma, int* lower,
                     int* upper) {
  const int64_t sfx = orc_sigma_fx(sigma);
  for (int p = 0; p < L; p++) {
    int lo = p - 1, up = p;
    for (int i = p; i >= 0; i--) {
      if (ct[p] - ct[i] > sfx) {
        lo = i;
        break;
      }
    }
    for (int i = p; i < L; i++) {
      if (ct[i] - ct[p] > sfx) {
        up = i - 1;
        break;
      }
    }
    lower[p] = lo;
    upper[p] = up;
  }
}

static void bounds_h(const int64_t* ctH, int w, int h, float sigma, int* lo,
                     int* up) {
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++)
    orc_line_bounds(ctH + (size_t)y * w, w, sigma, lo + (size_t)y * w,
                    up + (size_t)y * w);
}

static void bounds_v(const int64_t* ctV, int w, int h, float sigma, int* lo,
                     int* up) {
#pragma omp parallel
  {
    int64_t* line = (int64_t*)malloc(sizeof(int64_t) * h);
    int* l = (int*)malloc(sizeof(int) * h);
    int* u = (int*)malloc(sizeof(int) * h);
#pragma omp for schedule(static)
    for (int x = 0; x < w; x++) {
      for (int y = 0; y < h; y++) line[y] = ctV[(size_t)y * w + x];
      orc_line_bounds(line, h, sigma, l, u);
      for (int y = 0; y < h; y++) {
        lo[(size_t)y * w + x] = l[y];
        up[(size_t)y * w + x] = u[y];
      }
    }
    free(line);
    free(l);
    free(u);
  }
}

int orc_dts_bounds(const orc_dts* s, float sigma, int vertical, int* lo,
                   int* up) {
  if (vertical)
    bounds_v(s->ctV, s->w, s->h, sigma, lo, up);
  else
    bounds_h(s->ctH, s->w, s->h, sigma, lo, up);
  return 0;
}

/* -------------------------------------------------------------- passes */
/* Box mean B(p) = (S(u) - S(l)) / (u - l) with S the inclusive prefix sum of
 * the line and S(-1) = 0 (src/domain_transform_filter.cuh:30-71,
 * src/domain_transform_solver.cuh:53-58,100-102). */

typedef struct {
  int mode; /* 0: filter (out = B), 1: optimize X, 2: optimize Y */
  int loss;
  double lambda, step;
  /* overwrite after the pass (Optimize), when 0 < theta < 1 */
  int overwrite;
  double theta;
  const double *target, *conf;
  /* stereo (src/domain_transform_loss_term.cuh:20-39) */
  int stereo;
  double doff, photo_lambda, grad_alpha;
  const uint8_t *left, *right;
  const double* right_dH;
} pass_args;

static inline double data_grad(int loss, double c, double v, double t) {
  /* src/domain_transform_solver.cuh:12-23 */
  if (loss == ORC_L2) return 2 * c * (v - t);
  double x = v - t;
  return c * x / sqrt(x * x + 0.000001);
}

static double stereo_term(const pass_args* a, int w, int x, int y, double v) {
  /* src/domain_transform_loss_term.cuh:20-39; the valid range is the image
   * width. */
  double xi = a->doff + x - v;
  if (!(xi >= 0 && xi < w)) return 0;
  int xr = (int)xi;
  const uint8_t* l = a->left + 4 * ((size_t)y * w + x);
  const uint8_t* r = a->right + 4 * ((size_t)y * w + xr);
  double sad = 0;
  for (int k = 0; k < 4; k++) sad += fabs((double)l[k] - (double)r[k]);
  sad *= (double)kINV255_F;
  double g = a->right_dH[(size_t)y * w + xr];
  return a->photo_lambda * 2 * sad * ((g * g + (-1.0)) * a->grad_alpha);
}

/* One pass over all lines.  horizontal: lines are rows; vertical: columns. */
static void run_pass(int w, int h, int vertical, const double* in,
                     double* out, const int* lo, const int* up,
                     const pass_args* a) {
  const int nlines = vertical ? w : h;
  const int L = vertical ? h : w;
  const size_t sline = vertical ? 1 : (size_t)w;
  const size_t spos = vertical ? (size_t)w : 1;
#pragma omp parallel
  {
    double* S = (double*)malloc(sizeof(double) * (L + 1));
#pragma omp for schedule(static)
    for (int ln = 0; ln < nlines; ln++) {
      const size_t base = ln * sline;
      S[0] = 0; /* S[k+1] = prefix through k; S[0] is S(-1) = 0 */
      double acc = 0;
      for (int k = 0; k < L; k++) {
        acc += in[base + k * spos];
        S[k + 1] = acc;
      }
      for (int p = 0; p < L; p++) {
        const size_t i = base + p * spos;
        const int l = lo[i], u = up[i];
        const double n = (double)(u - l);
        const double B = (S[u + 1] - S[l + 1]) / n;
        double r;
        if (a->mode == 0) {
          r = B;
        } else {
          const double v = in[i];
          double g = 0;
          if (a->mode == 1) {
            g = data_grad(a->loss, a->conf[i] / n, v, a->target[i]);
            if (a->stereo) {
              int x = vertical ? ln : p, y = vertical ? p : ln;
              g += stereo_term(a, w, x, y, v);
            }
          }
          r = v + a->step * (-(a->lambda * (v - B) + 0.5 * g));
        }
        if (a->overwrite && a->conf[i] > a->theta) r = a->target[i];
        out[i] = r;
      }
    }
    free(S);
  }
}

/* -------------------------------------------------------------- filter */
/* src/domain_transform_filter.cu:124-188.  The differential and ct are
 * rebuilt with the filter params (a side effect that persists). */
static float filter_multiplier(int iter, int num_iter) {
  if (iter == 0) return 1.0f;
  return (float)(3 * pow(2, num_iter - (iter + 1)) /
                 sqrt(pow(4, num_iter) - 1));
}

static void filter_internal(orc_dts* s, float sx, float sy, float sr,
                            const double* in, double* out, int num_iter) {
  const int w = s->w, h = s->h;
  const size_t n = (size_t)w * h;
  differentials(s->color, w, h, sx, sy, sr, s->dH, s->dV);
  integrate(s->dH, s->dV, w, h, s->ctH, s->ctV);
  double* v = (double*)malloc(sizeof(double) * n);
  double* t = (double*)malloc(sizeof(double) * n);
  int* lo = (int*)malloc(sizeof(int) * n);
  int* up = (int*)malloc(sizeof(int) * n);
  memcpy(v, in, sizeof(double) * n);
  pass_args a;
  memset(&a, 0, sizeof(a));
  a.mode = 0;
  for (int it = 0; it < num_iter; it++) {
    float sxi = sx, syi = sy;
    if (it != 0) {
      const float m = filter_multiplier(it, num_iter);
      sxi *= m;
      syi *= m;
    }
    bounds_h(s->ctH, w, h, sxi, lo, up);
    run_pass(w, h, 0, v, t, lo, up, &a);
    bounds_v(s->ctV, w, h, syi, lo, up);
    run_pass(w, h, 1, t, v, lo, up, &a);
  }
  memcpy(out, v, sizeof(double) * n);
  free(v); free(t); free(lo); free(up);
}

int orc_dts_filter(orc_dts* s, float sx, float sy, float sr, const double* in,
                   double* out, int num_iter) {
  /* src/domain_transform_filter.cu:190-200: the result also lands in var. */
  filter_internal(s, sx, sy, sr, in, s->var, num_iter);
  memcpy(out, s->var, sizeof(double) * (size_t)s->w * s->h);
  return 0;
}

/* ------------------------------------------------------------ optimize */
static int optimize_impl(orc_dts* s, int K, int loss, float lambda,
                         float sigma_z, float step, float overwrite,
                         int stereo, float photo_lambda) {
  /* src/domain_transform_solver.cu:100-202 and
   * src/domain_transform_stereo.cu:44-158. */
  const int w = s->w, h = s->h;
  const size_t n = (size_t)w * h;
  memcpy(s->var, s->target, sizeof(double) * n);
  /* RobustLoss::L1 launches no kernel; the ping-pong then leaves var equal
   * to the target (src/domain_transform_solver.cu:133-153). */
  if (loss == ORC_L1 || K <= 0) return 0;
  const int ow = overwrite > 0 && overwrite < 1;
  int* loH = (int*)malloc(sizeof(int) * n);
  int* upH = (int*)malloc(sizeof(int) * n);
  int* loV = (int*)malloc(sizeof(int) * n);
  int* upV = (int*)malloc(sizeof(int) * n);
  double* t = (double*)malloc(sizeof(double) * n);
  /* The bounds depend only on ct and sigma_z, which are constant during
   * Optimize, so they are found once. */
  bounds_h(s->ctH, w, h, sigma_z, loH, upH);
  bounds_v(s->ctV, w, h, sigma_z, loV, upV);
  pass_args ax, ay;
  memset(&ax, 0, sizeof(ax));
  ax.mode = 1;
  ax.loss = loss;
  ax.lambda = lambda;
  ax.step = step;
  ax.overwrite = ow;
  ax.theta = overwrite;
  ax.target = s->target;
  ax.conf = s->conf;
  if (stereo) {
    /* src/domain_transform_stereo.cu:77-87 */
    const float q = s->fp_sr / s->fp_sx;
    ax.stereo = 1;
    ax.doff = s->doff;
    ax.photo_lambda = photo_lambda;
    ax.grad_alpha = (double)(q * q);
    ax.left = s->color;
    ax.right = s->right_color;
    ax.right_dH = s->right_dH;
  }
  ay = ax;
  ay.mode = 2;
  ay.stereo = 0;
  ay.overwrite = 0;
  for (int it = 0; it < K; it++) {
    if (ow)
      for (size_t i = 0; i < n; i++)
        if (s->conf[i] > overwrite) s->var[i] = s->target[i];
    run_pass(w, h, 0, s->var, t, loH, upH, &ax); /* includes overwrite */
    run_pass(w, h, 1, t, s->var, loV, upV, &ay);
  }
  free(loH); free(upH); free(loV); free(upV); free(t);
  return 0;
}

int orc_dts_optimize(orc_dts* s, int K, int loss, float lambda, float sigma_z,
                     float step, float overwrite) {
  return optimize_impl(s, K, loss, lambda, sigma_z, step, overwrite, 0, 0.f);
}

int orc_dts_optimize_stereo(orc_dts* s, int K, int loss, float lambda,
                            float sigma_z, float step, float photo_lambda,
                            float overwrite) {
  return optimize_impl(s, K, loss, lambda, sigma_z, step, overwrite, 1,
                       photo_lambda);
}

/* ------------------------------------------------- optimize adjoint */
/* Reverse-mode derivative of orc_dts_optimize (non-stereo) with respect to
 * the target and the confidence, for a cotangent ybar of the result.  No
 * reference counterpart: the reference has no gradients.  Every forward
 * state is kept (no checkpointing), and the box adjoint is the plain
 * scatter of each window's share, sum over p of ybar(p)/n(p) into
 * (l(p), u(p)], so it does not rely on any symmetry of the windows.  The
 * overwrite masks c > theta are piecewise constant in c and contribute no
 * derivative; the guide enters only through the integer bounds, so its
 * derivative is zero almost everywhere and is not formed. */
static void data_grad_partials(int loss, double c, double n, double v, double t,
                               double* dv, double* dt, double* dc) {
  const double x = v - t;
  if (loss == ORC_L2) {
    *dv = 2 * c / n;
    *dt = -2 * c / n;
    *dc = 2 * x / n;
  } else {
    const double den = x * x + 0.000001;
    const double rs = 1.0 / sqrt(den);
    *dv = (c / n) * 0.000001 * rs * rs * rs;
    *dt = -*dv;
    *dc = x * rs / n;
  }
}

/* Backward of run_pass in modes 1 (X, data term, overwrite) and 2 (Y).
 * obar: cotangent of out; in: the pass input; adds into vbar (zeroed by the
 * caller), tbar, cbar. */
static void run_pass_adjoint(int w, int h, int vertical, const double* in,
                             const double* obar, double* vbar, double* tbar,
                             double* cbar, const int* lo, const int* up,
                             const pass_args* a) {
  const int nlines = vertical ? w : h;
  const int L = vertical ? h : w;
  const size_t sline = vertical ? 1 : (size_t)w;
  const size_t spos = vertical ? (size_t)w : 1;
#pragma omp parallel
  {
    double* D = (double*)malloc(sizeof(double) * (L + 1));
#pragma omp for schedule(static)
    for (int ln = 0; ln < nlines; ln++) {
      const size_t base = ln * sline;
      for (int k = 0; k <= L; k++) D[k] = 0;
      for (int p = 0; p < L; p++) {
        const size_t i = base + p * spos;
        double rb = obar[i];
        if (a->overwrite && a->conf[i] > a->theta) {
          tbar[i] += rb;
          rb = 0;
        }
        const int l = lo[i], u = up[i];
        const double n = (double)(u - l);
        double dv = 0, dt = 0, dc = 0;
        if (a->mode == 1)
          data_grad_partials(a->loss, a->conf[i], n, in[i], a->target[i], &dv, &dt, &dc);
        vbar[i] += rb * (1 - a->step * a->lambda) - 0.5 * a->step * rb * dv;
        if (a->mode == 1) {
          tbar[i] += -0.5 * a->step * rb * dt;
          cbar[i] += -0.5 * a->step * rb * dc;
        }
        /* B(p) = sum_{l < k <= u} in(k) / n: scatter step*lambda*rb/n. */
        const double share = a->step * a->lambda * rb / n;
        for (int k = l + 1; k <= u; k++) D[k] += share;
      }
      for (int k = 0; k < L; k++) vbar[base + k * spos] += D[k];
    }
    free(D);
  }
}

int orc_dts_optimize_adjoint(orc_dts* s, int K, int loss, float lambda,
                             float sigma_z, float step, float overwrite,
                             const double* ybar, double* tbar, double* cbar) {
  const int w = s->w, h = s->h;
  const size_t n = (size_t)w * h;
  memset(cbar, 0, sizeof(double) * n);
  if (loss == ORC_L1 || K <= 0) {
    memcpy(s->var, s->target, sizeof(double) * n);
    memcpy(tbar, ybar, sizeof(double) * n);
    return 0;
  }
  memset(tbar, 0, sizeof(double) * n);
  const int ow = overwrite > 0 && overwrite < 1;
  int* loH = (int*)malloc(sizeof(int) * n);
  int* upH = (int*)malloc(sizeof(int) * n);
  int* loV = (int*)malloc(sizeof(int) * n);
  int* upV = (int*)malloc(sizeof(int) * n);
  /* states: W[k] = X-pass input of iteration k (after the pre-pass
   * overwrite), X[k] = X-pass output (Y-pass input). */
  double* Wst = (double*)malloc(sizeof(double) * n * K);
  double* Xst = (double*)malloc(sizeof(double) * n * K);
  double* g = (double*)malloc(sizeof(double) * n);
  double* gb = (double*)malloc(sizeof(double) * n);
  bounds_h(s->ctH, w, h, sigma_z, loH, upH);
  bounds_v(s->ctV, w, h, sigma_z, loV, upV);
  pass_args ax, ay;
  memset(&ax, 0, sizeof(ax));
  ax.mode = 1;
  ax.loss = loss;
  ax.lambda = lambda;
  ax.step = step;
  ax.overwrite = ow;
  ax.theta = overwrite;
  ax.target = s->target;
  ax.conf = s->conf;
  ay = ax;
  ay.mode = 2;
  ay.overwrite = 0;
  memcpy(s->var, s->target, sizeof(double) * n);
  for (int it = 0; it < K; it++) {
    double* wk = Wst + n * it;
    memcpy(wk, s->var, sizeof(double) * n);
    if (ow)
      for (size_t i = 0; i < n; i++)
        if (s->conf[i] > overwrite) wk[i] = s->target[i];
    run_pass(w, h, 0, wk, Xst + n * it, loH, upH, &ax);
    run_pass(w, h, 1, Xst + n * it, s->var, loV, upV, &ay);
  }
  memcpy(g, ybar, sizeof(double) * n);
  for (int it = K - 1; it >= 0; it--) {
    /* Y pass: g is the cotangent of its output. */
    memset(gb, 0, sizeof(double) * n);
    run_pass_adjoint(w, h, 1, Xst + n * it, g, gb, tbar, cbar, loV, upV, &ay);
    /* X pass (with its trailing overwrite). */
    memset(g, 0, sizeof(double) * n);
    run_pass_adjoint(w, h, 0, Wst + n * it, gb, g, tbar, cbar, loH, upH, &ax);
    /* pre-pass overwrite of iteration it */
    if (ow)
      for (size_t i = 0; i < n; i++)
        if (s->conf[i] > overwrite) {
          tbar[i] += g[i];
          g[i] = 0;
        }
  }
  for (size_t i = 0; i < n; i++) tbar[i] += g[i]; /* v^0 = target */
  free(loH); free(upH); free(loV); free(upV);
  free(Wst); free(Xst); free(g); free(gb);
  return 0;
}

/* ---------------------------------------------------------- confidence */
int orc_dts_compute_confidence(orc_dts* s, float sx, float sy, float sr,
                               float sigma0, int left_clear) {
  /* src/domain_transform_solver.cu:206-246 and :21-44 */
  const int w = s->w, h = s->h;
  const size_t n = (size_t)w * h;
  filter_internal(s, sx, sy, sr, s->target, s->conf, 3);
  for (size_t i = 0; i < n; i++) s->conf_sq_buf[i] = s->target[i] * s->target[i];
  filter_internal(s, sx, sy, sr, s->conf_sq_buf, s->conf_sq, 3);
  /* Filter leaves its result in var (src/domain_transform_filter.cu:136-187). */
  memcpy(s->var, s->conf_sq, sizeof(double) * n);
  const float scale_f = 0.5f / (sigma0 * sigma0);
  const double scale = scale_f;
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      const size_t i = (size_t)y * w + x;
      double c = 0;
      if (x > left_clear) {
        const double m = s->conf[i];
        const double var = s->conf_sq[i] - m * m;
        c = (var < 0 || !isfinite(var)) ? 0 : exp(-var * scale);
      }
      s->conf[i] = c;
    }
  }
  return 0;
}

const double* orc_dts_result(const orc_dts* s) { return s->var; }
const double* orc_dts_confidence(const orc_dts* s) { return s->conf; }
const double* orc_dts_target(const orc_dts* s) { return s->target; }
const double* orc_dts_dHdx(const orc_dts* s) { return s->dH; }
const double* orc_dts_dVdy(const orc_dts* s) { return s->dV; }
const int64_t* orc_dts_ctH_fx(const orc_dts* s) { return s->ctH; }
const int64_t* orc_dts_ctV_fx(const orc_dts* s) { return s->ctV; }
const uint8_t* orc_dts_color(const orc_dts* s) { return s->color; }

/* =================================================================== HFBS */
/* src/hfbs.cu:586-743.  Grid layout [gl][gy][gx]; reads outside the grid
 * are 0 (contract). */

typedef struct {
  int gx, gy, gl;
} gdim;

static inline size_t gidx(const gdim* g, int x, int y, int z) {
  return ((size_t)z * g->gy + y) * g->gx + x;
}

static inline double gget(const double* a, const gdim* g, int x, int y,
                          int z) {
  if (x < 0 || y < 0 || z < 0 || x >= g->gx || y >= g->gy || z >= g->gl)
    return 0.0;
  return a[gidx(g, x, y, z)];
}

void orc_laplace_line(const double* src, double* dst, int n, int stride,
                      double beta) {
  /* src/hfbs.cu:357-379: forward then backward exponential sweep. */
  double run = src[0];
  dst[0] = run;
  for (int i = 1; i < n; i++) {
    run = src[(size_t)i * stride] + beta * run;
    dst[(size_t)i * stride] = run;
  }
  run = src[(size_t)(n - 1) * stride];
  for (int i = n - 2; i >= 0; i--) {
    run = beta * run;
    dst[(size_t)i * stride] += run;
    run = run + src[(size_t)i * stride];
  }
}

static void laplace3(const gdim* g, double* a, double* tmp, double beta) {
  /* X (a -> tmp), Y (tmp -> a), Z (a -> tmp), then copy tmp -> a
   * (src/hfbs.cu:679-695 and :698-711). */
  const int gx = g->gx, gy = g->gy, gl = g->gl;
  for (int z = 0; z < gl; z++)
    for (int y = 0; y < gy; y++)
      orc_laplace_line(a + gidx(g, 0, y, z), tmp + gidx(g, 0, y, z), gx, 1,
                       beta);
  for (int z = 0; z < gl; z++)
    for (int x = 0; x < gx; x++)
      orc_laplace_line(tmp + gidx(g, x, 0, z), a + gidx(g, x, 0, z), gy, gx,
                       beta);
  for (int y = 0; y < gy; y++)
    for (int x = 0; x < gx; x++)
      orc_laplace_line(a + gidx(g, x, y, 0), tmp + gidx(g, x, y, 0), gl,
                       gx * gy, beta);
  memcpy(a, tmp, sizeof(double) * (size_t)gx * gy * gl);
}

/* Grid coordinates in fp32, as the reference computes them
 * (src/hfbs.cu:82-86,256-260); clamped into the grid (contract). */
static inline float gcoord(int v, float sigma, int gmax) {
  float g = (float)v * 1.0f / sigma;
  if (g > (float)(gmax - 1)) g = (float)(gmax - 1);
  return g;
}

/* K heavy-ball steps from Y0 (src/hfbs.cu:200-239, driven as in
 * :720-732): the ping-pong grids are Y0 and Y1, h holds the momentum (zero on
 * entry).  Returns the grid holding the last iterate. */
static const double* hfbs_solve(const gdim* gp, const orc_hfbs_params* p,
                                const double* b, const double* nrm,
                                const double* invA, double* Y0, double* Y1,
                                double* hb) {
  const gdim g = *gp;
  const double lam = p->lambda;
  double* ys = Y0;
  double* yd = Y1;
  const double alpha = p->alpha, beta = p->beta;
  for (int it = 0; it < p->optim_iter_max; it++) {
    for (int z = 0; z < g.gl; z++)
      for (int y = 0; y < g.gy; y++)
        for (int x = 0; x < g.gx; x++) {
          const double dif =
              gget(ys, &g, x - 1, y, z) * gget(nrm, &g, x - 1, y, z) +
              gget(ys, &g, x + 1, y, z) * gget(nrm, &g, x + 1, y, z) +
              gget(ys, &g, x, y - 1, z) * gget(nrm, &g, x, y - 1, z) +
              gget(ys, &g, x, y + 1, z) * gget(nrm, &g, x, y + 1, z) +
              gget(ys, &g, x, y, z - 1) * gget(nrm, &g, x, y, z - 1) +
              gget(ys, &g, x, y, z + 1) * gget(nrm, &g, x, y, z + 1);
          const size_t v = gidx(&g, x, y, z);
          const double q = (b[v] + lam * nrm[v] * dif) * invA[v];
          const double yv = ys[v];
          const double grad = yv - q;
          const double gc = fabs(grad) > 1.345 ? copysign(1.0, grad) : grad;
          const double hn = beta * hb[v] + gc;
          hb[v] = hn;
          yd[v] = yv - alpha * hn;
        }
    double* t = ys;
    ys = yd;
    yd = t;
  }
  return ys;
}

/* Trilinear slice of Y weighted by the splat count S (src/hfbs.cu:244-330)
 * at the fp32 grid coordinates of each pixel and its (converted) luma. */
static void hfbs_slice(const gdim* gp, const orc_hfbs_params* p,
                       const uint8_t* rgba, int w, int h, const double* S,
                       const double* ys, double* out) {
  const gdim g = *gp;
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      const size_t i = (size_t)y * w + x;
      const float fx = gcoord(x, p->sigma_x, g.gx);
      const float fy = gcoord(y, p->sigma_y, g.gy);
      const float fz = gcoord(rgba[4 * i], p->sigma_l, g.gl);
      const int x0 = (int)floorf(fx), y0 = (int)floorf(fy),
                z0 = (int)floorf(fz);
      const double ax = (double)fx - x0, ay = (double)fy - y0,
                   az = (double)fz - z0;
#define SY(X, Y, Z) (gget(S, &g, X, Y, Z) * gget(ys, &g, X, Y, Z))
#define SS(X, Y, Z) gget(S, &g, X, Y, Z)
      const double c00 = SY(x0, y0, z0) * (1 - ax) + SY(x0 + 1, y0, z0) * ax;
      const double c01 =
          SY(x0, y0, z0 + 1) * (1 - ax) + SY(x0 + 1, y0, z0 + 1) * ax;
      const double c10 =
          SY(x0, y0 + 1, z0) * (1 - ax) + SY(x0 + 1, y0 + 1, z0) * ax;
      const double c11 =
          SY(x0, y0 + 1, z0 + 1) * (1 - ax) + SY(x0 + 1, y0 + 1, z0 + 1) * ax;
      const double c0 = c00 * (1 - ay) + c10 * ay;
      const double c1 = c01 * (1 - ay) + c11 * ay;
      const double m00 = SS(x0, y0, z0) * (1 - ax) + SS(x0 + 1, y0, z0) * ax;
      const double m01 =
          SS(x0, y0, z0 + 1) * (1 - ax) + SS(x0 + 1, y0, z0 + 1) * ax;
      const double m10 =
          SS(x0, y0 + 1, z0) * (1 - ax) + SS(x0 + 1, y0 + 1, z0) * ax;
      const double m11 =
          SS(x0, y0 + 1, z0 + 1) * (1 - ax) + SS(x0 + 1, y0 + 1, z0 + 1) * ax;
      const double m0 = m00 * (1 - ay) + m10 * ay;
      const double m1 = m01 * (1 - ay) + m11 * ay;
#undef SY
#undef SS
      out[i] = (c0 * (1 - az) + c1 * az) / (m0 * (1 - az) + m1 * az);
    }
  }
}

static int hfbs_run(const uint8_t* rgba, int w, int h, const float* target,
                    const float* confidence, const orc_hfbs_params* p,
                    int max_gx, int max_gy, int max_gl, double* out,
                    int* grid_dims, double* o_b, double* o_c, double* o_s,
                    double* o_n, double* o_a, double* o_z0, double* o_z) {
  gdim g;
  /* src/hfbs.cu:589-591 */
  g.gx = (int)ceilf((float)w / p->sigma_x);
  g.gy = (int)ceilf((float)h / p->sigma_y);
  g.gl = (int)ceil(255.0 / (double)p->sigma_l);
  if (grid_dims) {
    grid_dims[0] = g.gx;
    grid_dims[1] = g.gy;
    grid_dims[2] = g.gl;
  }
  const size_t np = (size_t)w * h;
  if (g.gx > max_gx || g.gy > max_gy || g.gl > max_gl || g.gx <= 0 ||
      g.gy <= 0 || g.gl <= 0) {
    /* src/hfbs.cu:594-611 */
    if (out)
      for (size_t i = 0; i < np; i++) out[i] = target[i];
    return 1;
  }
  const size_t nv = (size_t)g.gx * g.gy * g.gl;
  double* b = (double*)calloc(nv, sizeof(double));
  double* cg = (double*)calloc(nv, sizeof(double));
  double* S = (double*)calloc(nv, sizeof(double));
  double* invA = (double*)calloc(nv, sizeof(double));
  double* n0 = (double*)malloc(nv * sizeof(double));
  double* n1 = (double*)malloc(nv * sizeof(double));
  double* tmp = (double*)malloc(nv * sizeof(double));
  double* Y0 = (double*)malloc(nv * sizeof(double));
  double* Y1 = (double*)malloc(nv * sizeof(double));
  double* hb = (double*)calloc(nv, sizeof(double));
  double* zb = (double*)malloc(nv * sizeof(double));

  /* Splat (src/hfbs.cu:71-120): b += c*t, c_g += c, S += 1, invA += c. */
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      const size_t i = (size_t)y * w + x;
      const int ix = (int)floorf(gcoord(x, p->sigma_x, g.gx));
      const int iy = (int)floorf(gcoord(y, p->sigma_y, g.gy));
      const int iz = (int)floorf(gcoord(rgba[4 * i], p->sigma_l, g.gl));
      const size_t v = gidx(&g, ix, iy, iz);
      const double c = confidence ? confidence[i] : 1.0;
      b[v] += c * (double)target[i];
      cg[v] += c;
      S[v] += 1;
      invA[v] += c;
    }
  }
  if (o_b) memcpy(o_b, b, nv * sizeof(double));
  if (o_s) memcpy(o_s, S, nv * sizeof(double));

  /* Bistochastization (src/hfbs.cu:125-155, :644-666); n starts at 1. */
  for (size_t v = 0; v < nv; v++) n0[v] = 1.0;
  double* src = n0;
  double* dst = n1;
  const double eps = (double)0.00001f;
  for (int it = 0; it < p->bistoch_iter_max; it++) {
    for (int z = 0; z < g.gl; z++)
      for (int y = 0; y < g.gy; y++)
        for (int x = 0; x < g.gx; x++) {
          const double d = gget(src, &g, x - 1, y, z) +
                           gget(src, &g, x + 1, y, z) +
                           gget(src, &g, x, y - 1, z) +
                           gget(src, &g, x, y + 1, z) +
                           gget(src, &g, x, y, z - 1) +
                           gget(src, &g, x, y, z + 1);
          const size_t v = gidx(&g, x, y, z);
          const double nv0 = src[v];
          const double bl = 2 * nv0 + d;
          dst[v] = sqrt(nv0 * (S[v] + eps) / (bl + eps));
        }
    double* t = src;
    src = dst;
    dst = t;
  }
  double* nrm = src;
  if (o_n) memcpy(o_n, nrm, nv * sizeof(double));

  /* Inverse diagonal (src/hfbs.cu:158-195). */
  const double lam = p->lambda;
  for (size_t v = 0; v < nv; v++) {
    const double diag = lam * (S[v] - 2 * nrm[v] * nrm[v]) + invA[v];
    invA[v] = 1.0 / (diag > 1e-3 ? diag : 1e-3);
  }
  if (o_a) memcpy(o_a, invA, nv * sizeof(double));

  /* Laplacian initialisation (src/hfbs.cu:677-717). */
  const double beta_b = p->neg_exp_beta;
  laplace3(&g, cg, tmp, beta_b);
  memcpy(zb, b, nv * sizeof(double));
  laplace3(&g, zb, tmp, beta_b);
  for (size_t v = 0; v < nv; v++) Y0[v] = zb[v] / (cg[v] + 1e-7);
  if (o_c) memcpy(o_c, cg, nv * sizeof(double));
  if (o_z0) memcpy(o_z0, Y0, nv * sizeof(double));

  /* Heavy-ball solve (src/hfbs.cu:200-239, :720-732). */
  const double* ys = hfbs_solve(&g, p, b, nrm, invA, Y0, Y1, hb);
  if (o_z) memcpy(o_z, ys, nv * sizeof(double));

  /* Trilinear slice weighted by the splat count (src/hfbs.cu:244-330). */
  if (out) hfbs_slice(&g, p, rgba, w, h, S, ys, out);
  free(b); free(cg); free(S); free(invA); free(n0); free(n1); free(tmp);
  free(Y0); free(Y1); free(hb); free(zb);
  return 0;
}

int orc_hfbs_optimize(const uint8_t* rgba, int w, int h, const float* target,
                      const float* confidence, const orc_hfbs_params* p,
                      int max_gx, int max_gy, int max_gl, double* out,
                      int* grid_dims) {
  return hfbs_run(rgba, w, h, target, confidence, p, max_gx, max_gy, max_gl,
                  out, grid_dims, NULL, NULL, NULL, NULL, NULL, NULL, NULL);
}

int orc_hfbs_grids(const uint8_t* rgba, int w, int h, const float* target,
                   const float* confidence, const orc_hfbs_params* p,
                   double* b, double* splat_c, double* splat_sum,
                   double* splat_norm, double* inv_diag_a, double* z_init,
                   double* z_final) {
  return hfbs_run(rgba, w, h, target, confidence, p, 1 << 30, 1 << 30,
                  1 << 30, NULL, NULL, b, splat_c, splat_sum, splat_norm,
                  inv_diag_a, z_init, z_final);
}

int orc_hfbs_solve(int gx, int gy, int gl, const orc_hfbs_params* p,
                   const double* b, const double* splat_norm,
                   const double* inv_diag_a, const double* z_init,
                   double* z_final) {
  const gdim g = {gx, gy, gl};
  if (gx <= 0 || gy <= 0 || gl <= 0) return 1;
  const size_t nv = (size_t)gx * gy * gl;
  double* Y0 = (double*)malloc(nv * sizeof(double));
  double* Y1 = (double*)malloc(nv * sizeof(double));
  double* hb = (double*)calloc(nv, sizeof(double));
  memcpy(Y0, z_init, nv * sizeof(double));
  const double* ys = hfbs_solve(&g, p, b, splat_norm, inv_diag_a, Y0, Y1, hb);
  memcpy(z_final, ys, nv * sizeof(double));
  free(Y0); free(Y1); free(hb);
  return 0;
}

int orc_hfbs_slice(const uint8_t* rgba_converted, int w, int h,
                   const orc_hfbs_params* p, int gx, int gy, int gl,
                   const double* splat_sum, const double* z, double* out) {
  const gdim g = {gx, gy, gl};
  if (gx <= 0 || gy <= 0 || gl <= 0) return 1;
  hfbs_slice(&g, p, rgba_converted, w, h, splat_sum, z, out);
  return 0;
}
