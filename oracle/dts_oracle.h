/*
 * FP64 CPU oracle for the Domain Transform Solver (DTS) and HFBS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library's results are compared against
 * it; nothing in the product library links, loads or calls it.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark use it.
 *
 * It restates the reference algorithm (akashbapat/domain_transform_solver)
 * in double precision with the reference's parameters, iteration order and
 * buffer ping-pong.  Every function cites the reference file:line it follows.
 * Where the reference reads out of bounds or relies on unspecified container
 * behaviour, the oracle follows one written contract (the same one the GPU
 * kernels follow):
 *   - guide pixels outside the image read as (0,0,0,0);
 *   - the prefix sum S(-1) of a line is 0;
 *   - the vertical window search runs over the column length H (the
 *     reference loops to the image width);
 *   - the transformed-domain coordinate ct is accumulated as 32.32 fixed
 *     point (int64, d rounded to 2^-32), so window bounds do not depend on
 *     the summation order;
 *   - HFBS grid reads outside [0,grid) are 0 and the luma bin is clamped to
 *     grid_l-1.
 *
 * Planes are row-major, tightly packed, W*H elements.
 */
#ifndef DTS_ORACLE_H_
#define DTS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Colour spaces, image types and losses use the reference's enum order
 * (src/domain_transform_common.h:20-32, src/domain_transform_optimize.h:8). */
enum { ORC_RGB = 0, ORC_YCBCR = 1, ORC_YYY = 2 };
enum { ORC_L1 = 0, ORC_L2 = 1, ORC_CHARBONNIER = 2 };

typedef struct orc_dts orc_dts;

/* State object mirroring DomainTransformFilter/Solver/Stereo
 * (src/domain_transform_filter.h:19-70, src/domain_transform_solver.h:34-85,
 * src/domain_transform_stereo.h:36-57). */
orc_dts* orc_dts_create(int max_w, int max_h);
void orc_dts_destroy(orc_dts* s);
int orc_dts_set_image_dim(orc_dts* s, int w, int h);
void orc_dts_set_threads(int n);
int orc_dts_clear_all(orc_dts* s);
int orc_dts_init_frame(orc_dts* s, int color_space, const uint8_t* rgba,
                       const float* target, const float* confidence);
int orc_dts_init_right_frame(orc_dts* s, float doff, int color_space,
                             const uint8_t* rgba);
int orc_dts_compute_differential(orc_dts* s, float sigma_x, float sigma_y,
                                 float sigma_r);
int orc_dts_integrate(orc_dts* s);
int orc_dts_process_right_frame(orc_dts* s);
int orc_dts_filter(orc_dts* s, float sigma_x, float sigma_y, float sigma_r,
                   const double* in, double* out, int num_iter);
int orc_dts_optimize(orc_dts* s, int num_iter, int loss, float lambda,
                     float sigma_z, float step_size, float overwrite);
int orc_dts_set_target_conf_f64(orc_dts* s, const double* target,
                                const double* confidence);
/* Reverse-mode derivative of orc_dts_optimize (non-stereo) for a cotangent
 * ybar of the result: tbar = d<ybar, v_K>/dt, cbar = d<ybar, v_K>/dc.
 * Re-runs the forward (the result is left in the state). */
int orc_dts_optimize_adjoint(orc_dts* s, int num_iter, int loss, float lambda,
                             float sigma_z, float step, float overwrite,
                             const double* ybar, double* tbar, double* cbar);
int orc_dts_optimize_stereo(orc_dts* s, int num_iter, int loss, float lambda,
                            float sigma_z, float step_size,
                            float photo_lambda, float overwrite);
int orc_dts_compute_confidence(orc_dts* s, float sigma_x, float sigma_y,
                               float sigma_r, float sigma0, int left_clear);

/* Accessors (double planes of the current image size). */
const double* orc_dts_result(const orc_dts* s);     /* OPTIMIZED_QUANTITY */
const double* orc_dts_confidence(const orc_dts* s); /* CONFIDENCE */
const double* orc_dts_target(const orc_dts* s);     /* TARGET */
const double* orc_dts_dHdx(const orc_dts* s);       /* DIFFERENTIAL */
const double* orc_dts_dVdy(const orc_dts* s);
const int64_t* orc_dts_ctH_fx(const orc_dts* s);    /* INTEGRAL, 32.32 */
const int64_t* orc_dts_ctV_fx(const orc_dts* s);    /* row-major [y][x] */
const uint8_t* orc_dts_color(const orc_dts* s);     /* COLOR_IMAGE */

/* Window bounds of one line: lower[p], upper[p] as absolute indices. */
void orc_line_bounds(const int64_t* ct_fx, int L, float sigma, int* lower,
                     int* upper);
/* All bounds of the current frame for one sigma; lower/upper are W*H planes
 * in image layout (horizontal: along rows; vertical: along columns). */
int orc_dts_bounds(const orc_dts* s, float sigma, int vertical, int* lower,
                   int* upper);

/* 32.32 fixed-point helpers shared with the tests. */
int64_t orc_to_fx(double d);
int64_t orc_sigma_fx(float sigma);

/* ---------------------------------------------------------------- HFBS */
typedef struct {
  float sigma_x, sigma_y, sigma_l, lambda, alpha, beta;
  int bistoch_iter_max, optim_iter_max;
  float neg_exp_beta;
} orc_hfbs_params;

/* One full HFBS::Optimize (src/hfbs.cu:586-743) on a frame whose guide has
 * already been colour-converted.  out gets W*H values.  When grid_* are not
 * NULL they receive (gx,gy,gl).  Returns 1 when the grid exceeds the
 * max_grid limits (output = target, src/hfbs.cu:594-611), 0 on success. */
int orc_hfbs_optimize(const uint8_t* rgba_converted, int w, int h,
                      const float* target, const float* confidence,
                      const orc_hfbs_params* p, int max_gx, int max_gy,
                      int max_gl, double* out, int* grid_dims);
/* Intermediate grids for the tests (layout [gl][gy][gx]); any may be NULL. */
int orc_hfbs_grids(const uint8_t* rgba_converted, int w, int h,
                   const float* target, const float* confidence,
                   const orc_hfbs_params* p, double* b, double* splat_c,
                   double* splat_sum, double* splat_norm, double* inv_diag_a,
                   double* z_init, double* z_final);
/* The heavy-ball solve alone (src/hfbs.cu:200-239, :720-732): p->optim_iter_max
 * steps from z_init with zero momentum; grids [gl][gy][gx]. */
int orc_hfbs_solve(int gx, int gy, int gl, const orc_hfbs_params* p,
                   const double* b, const double* splat_norm,
                   const double* inv_diag_a, const double* z_init,
                   double* z_final);
/* The trilinear slice alone (src/hfbs.cu:244-330) of grid z weighted by
 * splat_sum, at the pixels of a colour-converted guide. */
int orc_hfbs_slice(const uint8_t* rgba_converted, int w, int h,
                   const orc_hfbs_params* p, int gx, int gy, int gl,
                   const double* splat_sum, const double* z, double* out);
/* Symmetric exponential sweep along one line (src/hfbs.cu:357-379). */
void orc_laplace_line(const double* src, double* dst, int n, int stride,
                      double beta);
void orc_convert_color(uint8_t* rgba, int w, int h, int color_space);

#ifdef __cplusplus
}
#endif

/* Differentials and ct of a real-valued guide (4*w*h doubles, RGBA), for
 * finite-difference checks of the guide derivative. */
int orc_dts_set_guide_f64(orc_dts* s, const double* guide, float sx, float sy, float sr);

/* ---- RF-DT (recursive-filter domain transform), rf_oracle.c ----------------
 * Not in the reference: the recursive variant of the domain transform
 * (Gastal and Oliveira 2011) on the reference's differential.  guide: real
 * RGBA, 4*w*h doubles; planes w*h doubles; n iterations (rows with sigma_x,
 * then columns with sigma_y, per iteration). */
double orc_rf_sigma_i(double sigma, int i, int n);
/* One line, causal then anticausal: wt[k] joins k-1 and k (wt[0] unused). */
void orc_rf_line(const double* x, const double* wt, double* z, int L);
int orc_rf_filter(const double* guide, int w, int h, float sx, float sy, float sr, int n,
                  const double* in, double* out);
/* Reverse mode: xbar = d<ybar, out>/d in (w*h), gbar = d<ybar, out>/d guide
 * (4*w*h, may be NULL). */
int orc_rf_filter_adjoint(const double* guide, int w, int h, float sx, float sy, float sr,
                          int n, const double* in, const double* ybar, double* xbar,
                          double* gbar);

#endif /* DTS_ORACLE_H_ */
