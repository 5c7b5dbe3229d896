/*
 * C ABI of the MI355X (gfx950) Domain Transform Solver / HFBS library
 * (libdts_amd.so).  Plain pointers and sizes only; every function returns an
 * int status (DTS_OK = 0) and never exits the process.  dts_last_error()
 * returns a thread-local description of the last failure.
 *
 * Each entry point replaces one member of the reference's C++ class API; the
 * reference has no C ABI of its own, so the header-compatible C++ classes in
 * domain_transform_filter.h / domain_transform_solver.h /
 * domain_transform_stereo.h / hfbs.h (this directory) are implemented on top
 * of these calls, and Python reaches them through ctypes.
 *
 * Host buffers are tightly packed row-major planes of exactly W*H elements of
 * the current image size (uchar4 guide, fp32 planes).  Device work runs on the
 * handle's own HIP stream (dts_get_stream / dts_set_stream); compute calls are
 * asynchronous, uploads and downloads synchronise, as in the reference.  Calls
 * with device-pointer outputs (device_ptrs != 0) are stream-ordered: they
 * return without waiting, and their results are ready for later work on the
 * handle's stream or after dts_synchronize.
 */
#ifndef DTS_CAPI_H_
#define DTS_CAPI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define DTS_OK 0
#define DTS_ERR_INVALID_ARG 1
#define DTS_ERR_INVALID_DIMS 2
#define DTS_ERR_HIP 3
#define DTS_ERR_STATE 4
#define DTS_ERR_UNSUPPORTED 5

/* Enum values follow src/domain_transform_common.h:20-32 and
 * src/domain_transform_optimize.h:8. */
#define DTS_COLOR_RGB 0
#define DTS_COLOR_YCBCR 1
#define DTS_COLOR_YYY 2

#define DTS_IMAGE_COLOR_IMAGE 0
#define DTS_IMAGE_DIFFERENTIAL 1
#define DTS_IMAGE_INTEGRAL 2
#define DTS_IMAGE_TARGET 3
#define DTS_IMAGE_CONFIDENCE 4
#define DTS_IMAGE_OPTIMIZED_QUANTITY 5

#define DTS_LOSS_L1 0
#define DTS_LOSS_L2 1
#define DTS_LOSS_CHARBONNIER 2

/* Handle flags. */
#define DTS_FLAG_STEREO 1 /* allocate the right-frame planes */

/* Mirrors DomainFilterParams (src/domain_transform_filter_params.h:8-31). */
typedef struct {
  float sigma_x;
  float sigma_y;
  float sigma_r;
} dts_filter_params;

/* Mirrors DomainOptimizeParams (src/domain_transform_optimize.h:30-50). */
typedef struct {
  int num_iterations;
  int loss;
  float lambda;
  float sigma_z;
  float step_size;
  float photo_consis_lambda;
  float DT_lambda;
} dts_optimize_params;

/* Mirrors DtsOptions minus the image dims (src/edge_aware_solver.h:13-21). */
typedef struct {
  dts_filter_params filter_options;
  dts_optimize_params optim_options;
  dts_filter_params conf_params; /* DomainFilterParamsVec<1> sigma_x/y/r */
  float conf_sigma0;             /* DomainFilterParamsVec<1>::sigmas[0] */
  float overwrite_target_above_conf;
} dts_solve_options;

typedef struct dts_handle dts_handle;

const char* dts_last_error(void);
int dts_version(void);

/* DomainTransformFilter(max_image_dims) / DomainTransformSolver(...) /
 * DomainTransformStereo(...) constructors
 * (src/domain_transform_filter.cu:76-83, src/domain_transform_solver.cu:95-98,
 * src/domain_transform_stereo.cu:40-42). */
int dts_create(int max_width, int max_height, int flags, dts_handle** out);
int dts_destroy(dts_handle* h);

/* Stream control: by default the handle owns a non-blocking stream. */
int dts_get_stream(dts_handle* h, void** stream);
int dts_set_stream(dts_handle* h, void* stream); /* NULL: back to own stream */
int dts_synchronize(dts_handle* h);

/* SetImageDim (src/domain_transform_filter.cu:114-122): invalid dims return
 * DTS_ERR_INVALID_DIMS and keep the previous size. */
int dts_set_image_dim(dts_handle* h, int width, int height);
int dts_get_image_dim(dts_handle* h, int* width, int* height);

/* ClearAll (src/domain_transform_solver.cu:248-255). */
int dts_clear_all(dts_handle* h);

/* InitFrame (src/domain_transform_solver.cu:46-61): guide upload plus
 * in-place colour conversion; confidence NULL fills 1. */
int dts_init_frame(dts_handle* h, int color_space, const void* rgba,
                   const float* target, const float* confidence);
/* DomainTransformFilter::InitFrame (src/domain_transform_filter.cu:18-49):
 * the guide alone; target and confidence planes are left as they are. */
int dts_init_guide(dts_handle* h, int color_space, const void* rgba);
/* Same from device memory (no host copies). */
int dts_init_frame_device(dts_handle* h, int color_space, const void* rgba_dev,
                          const float* target_dev, const float* conf_dev);

/* DomainTransformStereo::InitRightFrame / ProcessRightFrame
 * (src/domain_transform_stereo.cu:17-27). */
int dts_init_right_frame(dts_handle* h, float doff, int color_space,
                         const void* rgba);
int dts_process_right_frame(dts_handle* h);

/* ComputeColorSpaceDifferential / IntegrateColorDifferentials
 * (src/domain_transform_solver.cu:89-93, src/domain_transform_filter.cu:85-110). */
int dts_compute_differential(dts_handle* h, const dts_filter_params* p);
int dts_integrate(dts_handle* h);

/* Filter with host pointers (src/domain_transform_filter.cu:190-200). */
int dts_filter(dts_handle* h, const dts_filter_params* p, const float* input,
               float* output, int num_iter);
/* Filter on device pointers (src/domain_transform_filter.cu:124-188). */
int dts_filter_device(dts_handle* h, const dts_filter_params* p,
                      const float* input_dev, float* output_dev, int num_iter);

/* Optimize (src/domain_transform_solver.cu:100-202); the stereo variant is
 * DomainTransformStereo::Optimize (src/domain_transform_stereo.cu:44-158). */
int dts_optimize(dts_handle* h, const dts_optimize_params* p,
                 float overwrite_target_above_conf);
int dts_optimize_stereo(dts_handle* h, const dts_optimize_params* p,
                        float overwrite_target_above_conf);

/* Differentiable Optimize.  No counterpart in the reference, which has no
 * gradients; it differentiates dts_optimize (src/domain_transform_solver.cu:
 * 100-202, non-stereo).  dts_optimize_checkpointed runs the same forward as
 * dts_optimize (identical result in OPTIMIZED_QUANTITY) and keeps the state
 * every ckpt_every iterations (<= 0: ceil(sqrt(K))) plus the states of the
 * last segment.  dts_optimize_adjoint then returns, for a cotangent ybar of
 * the result, tbar = d<ybar, v_K>/dtarget and cbar = d<ybar, v_K>/dconfidence
 * (W*H floats each); gbar (4*W*H floats, may be NULL) receives the guide
 * derivative, zero almost everywhere because the guide only moves the
 * integer window bounds.  device_ptrs != 0: ybar/tbar/cbar/gbar are device
 * pointers.  Any call that changes the frame, dims or bounds in between
 * invalidates the tape (DTS_ERR_STATE). */
int dts_optimize_checkpointed(dts_handle* h, const dts_optimize_params* p,
                              float overwrite_target_above_conf, int ckpt_every);
int dts_optimize_adjoint(dts_handle* h, const float* ybar, float* tbar, float* cbar,
                         float* gbar, int device_ptrs);

/* Recursive-filter domain transform (RF-DT).  No counterpart in the
 * reference, which filters with the box (normalized convolution) domain
 * transform only; it is the recursive variant of Gastal and Oliveira (2011)
 * on the reference's differential (src/domain_transform_common_kernels.cuh:
 * 71-116): per iteration i of num_iter (1..16), every row then every column
 * runs the causal recursion y_k = (1 - w_k) x_k + w_k y_(k-1) and then the
 * anticausal one, with w = exp(-sqrt(2) d / sigma_i), d the transformed-
 * domain step between neighbours (sigma_x / sigma_r along rows, sigma_y /
 * sigma_r along columns) and sigma_i = sigma sqrt(3) 2^(N-1-i) /
 * sqrt(4^N - 1).  The guide is the current frame's (after InitFrame's colour
 * conversion).  keep_tape != 0 keeps the input of every pass (2 num_iter
 * planes) for dts_rf_filter_adjoint, which returns, for a cotangent ybar of
 * the output, xbar = d<ybar, out>/d input (W*H floats) and gbar =
 * d<ybar, out>/d guide (4*W*H floats, RGBA per pixel, the guide taken as
 * real values; may be NULL).  device_ptrs != 0: every plane argument is a
 * device pointer.  A new frame, guide or dims invalidates the tape
 * (DTS_ERR_STATE). */
int dts_rf_filter(dts_handle* h, const dts_filter_params* p, const float* input,
                  float* output, int num_iter, int keep_tape, int device_ptrs);
int dts_rf_filter_adjoint(dts_handle* h, const float* ybar, float* xbar, float* gbar,
                          int device_ptrs);

/* ComputeConfidence (src/domain_transform_solver.cu:206-246). */
int dts_compute_confidence(dts_handle* h, const dts_filter_params* conf_params,
                           float sigma0, int left_side_clear_width);

/* Solve<DtsOptions> (src/edge_aware_solver.h:37-48): differential,
 * integrate, confidence when compute_confidence != 0, optimize. */
int dts_solve(dts_handle* h, const dts_solve_options* o, int compute_confidence);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops): dts_solve on a frame that
 * lives in device memory (guide uchar4 W*H, target and optional confidence
 * W*H floats; NULL confidence means 1), result written to out_dev (W*H
 * floats, not aliasing an input), queued on `stream` (NULL: the default
 * stream) without waiting and without any allocation.  Intermediate planes
 * live in one static device array per GPU, held until the process ends:
 * 48 bytes per pixel for frames of at most 2^20 pixels with sides up to
 * 4096 (DTS_ERR_INVALID_DIMS beyond).  Windows too wide for the LDS segments
 * return DTS_ERR_UNSUPPORTED.  Calls that overlap on one GPU share that array
 * and race; serialise them (one stream, or events). */
int dts_solve_noalloc(int width, int height, int color_space, const void* rgba_dev,
                      const float* target_dev, const float* conf_dev,
                      const dts_solve_options* o, int compute_confidence, float* out_dev,
                      void* stream);

/* Download (src/domain_transform_solver.cu:63-87,
 * src/domain_transform_filter.cu:51-74).  COLOR_IMAGE gives W*H uchar4,
 * every other type W*H floats. */
int dts_download(dts_handle* h, int image_type, void* out);
/* Device pointer of the result plane (OPTIMIZED_QUANTITY), valid until the
 * next call on the handle. */
int dts_result_device(dts_handle* h, const float** out);

/* Window bounds of the current frame for one sigma, for tests: lower/upper
 * absolute indices per pixel in image layout (W*H int32 each). */
int dts_debug_bounds(dts_handle* h, float sigma, int vertical, int* lower,
                     int* upper);
/* 32.32 fixed-point transformed-domain coordinate ct (W*H int64, row-major
 * [y][x]) for tests; vertical selects ct_V. */
int dts_debug_ct(dts_handle* h, int vertical, int64_t* out);

/* Calibration: average time (ms, events on the null stream) of a plain
 * streaming kernel out = a + b + c + d over n floats (n % 4 == 0, 16-byte
 * aligned device pointers), 20 bytes per element, repeated reps times. */
int dts_debug_stream(const float* a, const float* b, const float* c, const float* d,
                     float* out, size_t n, int reps, float* ms);

/* HIP graphs: when enabled, dts_solve and dts_optimize replay a captured
 * graph per (dims, options) instead of launching kernel by kernel. */
int dts_set_graph_mode(dts_handle* h, int enable);

/* Timing helpers: events on the handle's stream. */
int dts_timer_start(dts_handle* h);
int dts_timer_stop(dts_handle* h, float* ms);
/* Per-kernel timing of the line passes: when enabled every horizontal and
 * vertical pass launch is bracketed by events; read sums and counts back. */
int dts_profile_passes(dts_handle* h, int enable);
int dts_profile_read(dts_handle* h, double* x_ms, int* x_count, double* y_ms,
                     int* y_count);
/* Sum and count of the fused-iteration launches (dts_set_fused) timed while
 * pass profiling is on. */
int dts_profile_read_fused(dts_handle* h, double* ms, int* count);

/* Optimize without the target overwrite runs each iteration as one launch
 * (the horizontal half of the rows a vertical tile needs recomputed in the
 * tile) where the windows fit; enable = 0 keeps the two launches per
 * iteration of the reference (src/domain_transform_solver.cu:100-202).  On by
 * default; the checkpointed forward and the adjoint's replays follow the
 * same setting.  The two forms agree to float rounding, not bit for bit: the
 * horizontal prefix sums start at different columns.  No counterpart in the
 * reference. */
int dts_set_fused(dts_handle* h, int enable);

/* ---------------------------------------------------------------- batched */
/* Many frames of uneven size solved together (no counterpart in the
 * reference, whose solvers take one frame per call).  Each frame carries the
 * inputs of one InitFrame + Solve<DtsOptions> with its confidence given
 * (src/edge_aware_solver.h:37-48): its own sigmas, loss, iterations and
 * overwrite threshold. */
typedef struct {
  int width;
  int height;
  int color_space;
  const void* rgba;         /* W*H uchar4 */
  const float* target;      /* W*H */
  const float* confidence;  /* W*H, or NULL for 1 */
  dts_filter_params filter_options;
  dts_optimize_params optim_options;
  float overwrite_target_above_conf;
} dts_batch_frame;

typedef struct dts_batch dts_batch;

/* Packed device planes for up to max_frames frames of max_total_pixels
 * pixels in all (< 2^30). */
int dts_batch_create(int max_frames, size_t max_total_pixels, dts_batch** out);
int dts_batch_destroy(dts_batch* b);
/* Uploads (host pointers, or device pointers when device_ptrs != 0) and
 * lays out the frames; invalidates the captured graph. */
int dts_batch_set_frames(dts_batch* b, const dts_batch_frame* frames, int n, int device_ptrs);
/* How dts_batch_set_frames plans the batched passes (no counterpart in the
 * reference).  enable != 0 (default): it runs one bounds pass over the new
 * frames and sizes every pass class (horizontal segment chunks, vertical
 * register class, fused chunk span) by each frame's measured window extents,
 * which depend only on its guide and sigmas, so the plan holds for every
 * solve until the next dts_batch_set_frames; 0: by the sigma-based caps,
 * without that pass.  Results are the same either way up to float rounding
 * of the fused iterations.  Takes effect at the next dts_batch_set_frames. */
int dts_batch_set_extent_planning(dts_batch* b, int enable);
/* The extents frame's passes are planned for: largest p - l and u - p of its
 * horizontal, then vertical, windows (ext4[0..3]). */
int dts_batch_get_frame_extents(dts_batch* b, int frame, int* ext4);
/* One solve of every frame: a single hipGraph replay (graph mode, default)
 * of the batched setup and iteration launches. */
int dts_batch_solve(dts_batch* b);
int dts_batch_set_graph_mode(dts_batch* b, int enable);
/* Tuning of the batched launches; every choice gives the same results.
 * x_blocks_per_cu: persistent horizontal-pass workgroups per CU (1..16);
 * y_register_tiles: vertical tiles stage their rows in registers (one launch
 * per register class) instead of looping through LDS (one launch). */
int dts_batch_set_tuning(dts_batch* b, int x_blocks_per_cu, int y_register_tiles);
int dts_batch_get_tuning(dts_batch* b, int* x_blocks_per_cu, int* y_register_tiles);
/* One launch per Optimize iteration (the horizontal half of the rows a
 * vertical tile needs recomputed in the tile) for the frames without the
 * target overwrite whose horizontal window span fits max_chunks chunks of 64
 * columns (2, 3 or 4; 0: none); the other frames keep the two launches per
 * iteration in the same replay.  Default 3.  Fused results agree with the
 * two-pass iterations to float rounding, not bit for bit.  get: the setting
 * and (may be null) the number of frames it fuses.  No counterpart in the
 * reference. */
int dts_batch_set_fused(dts_batch* b, int max_chunks);
int dts_batch_get_fused(dts_batch* b, int* max_chunks, int* frames_fused);
/* Wavefronts per fused tile (4 or 8; default 4): with 8 each wavefront
 * recomputes half as many staged rows.  A chunk class whose tile would not
 * fit LDS with 8 keeps 4.  Results agree across the choices to float
 * rounding (the column sums are associated differently).  No counterpart
 * in the reference. */
int dts_batch_set_fused_waves(dts_batch* b, int waves);
int dts_batch_get_fused_waves(dts_batch* b, int* waves);
/* Times x_blocks_per_cu in {1, 2, 4, 8} x y_register_tiles in {0, 1} with
 * two launches per iteration for every frame, then at the best of those
 * each fused chunk limit that some frame reaches, then at the best limit 8
 * wavefronts per fused tile, on the current frames (reps solves each, after
 * one warm-up); keeps the fastest and returns its time per solve. */
int dts_batch_autotune(dts_batch* b, int reps, float* best_ms);
int dts_batch_download(dts_batch* b, int frame, float* out);
int dts_batch_result_device(dts_batch* b, int frame, const float** out);
int dts_batch_synchronize(dts_batch* b);
int dts_batch_timer_start(dts_batch* b);
int dts_batch_timer_stop(dts_batch* b, float* ms);

/* ------------------------------------------------------------------ HFBS */
/* Mirrors HFBSGridParams (src/hfbs.h:18-60); grid_* are derived. */
typedef struct {
  float sigma_x;
  float sigma_y;
  float sigma_l;
  float lambda;
  float alpha;
  float beta;
  int bistoch_iter_max;
  int optim_iter_max;
  float neg_exp_beta;
  int grid_x;
  int grid_y;
  int grid_l;
} hfbs_grid_params;

typedef struct hfbs_handle hfbs_handle;

/* HFBS(max_image_dims, max_grid_x, max_grid_y, max_grid_l) (src/hfbs.cu:578-582). */
int hfbs_create(int max_width, int max_height, int max_grid_x, int max_grid_y,
                int max_grid_l, hfbs_handle** out);
int hfbs_destroy(hfbs_handle* h);
int hfbs_get_stream(hfbs_handle* h, void** stream);
int hfbs_set_stream(hfbs_handle* h, void* stream);
int hfbs_synchronize(hfbs_handle* h);
/* SetImageDim (src/hfbs.cu:527-538). */
int hfbs_set_image_dim(hfbs_handle* h, int width, int height);
/* SetFilterParams (src/hfbs.cu:494-496); defaults are src/hfbs.h:20-52. */
int hfbs_set_params(hfbs_handle* h, const hfbs_grid_params* p);
int hfbs_default_params(hfbs_grid_params* p);
/* InitFrame (src/hfbs.cu:498-553). */
int hfbs_init_frame(hfbs_handle* h, int color_space, const void* rgba,
                    const float* target, const float* confidence);
int hfbs_init_frame_device(hfbs_handle* h, int color_space, const void* rgba_dev,
                           const float* target_dev, const float* conf_dev);
/* ClearAll (src/hfbs.cu:745-759). */
int hfbs_clear_all(hfbs_handle* h);
/* Optimize (src/hfbs.cu:586-743).  Returns DTS_OK also when the grid is too
 * large; the output then equals the target, as in the reference, and
 * *fell_back (optional) is set to 1. */
int hfbs_optimize(hfbs_handle* h, int* fell_back);
/* Download (src/hfbs.cu:555-576): TARGET, CONFIDENCE, OPTIMIZED_QUANTITY. */
int hfbs_download(hfbs_handle* h, int image_type, float* out);
int hfbs_result_device(hfbs_handle* h, const float** out);
/* Grid dims and intermediate grids ([gl][gy][gx] floats) for tests:
 * which = 0 b, 1 blurred splat_c, 2 splat_sum, 3 splat_norm, 4 inv_diag_A,
 * 5 z after the solve. */
int hfbs_debug_grid_dims(hfbs_handle* h, int* gx, int* gy, int* gl);
int hfbs_debug_grid(hfbs_handle* h, int which, float* out);
int hfbs_set_graph_mode(hfbs_handle* h, int enable);
int hfbs_timer_start(hfbs_handle* h);
int hfbs_timer_stop(hfbs_handle* h, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* DTS_CAPI_H_ */
