// The reference's C++ solver classes implemented on the C ABI.
//
// Observable behaviour follows the reference: invalid image sizes print a
// message and keep the previous size (src/domain_transform_filter.cu:114-122,
// src/hfbs.cu:527-538); a failing device call prints the error and exits, as
// GPU_CHECK does (src/error_types_cuda.h:21-35).
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "domain_transform_filter.h"
#include "domain_transform_solver.h"
#include "domain_transform_stereo.h"
#include "dts_capi.h"
#include "hfbs.h"

namespace domain_transform {
namespace {

void Check(int rc, const char* what) {
  if (rc == DTS_OK) return;
  std::cerr << what << ": " << dts_last_error() << std::endl;
  if (rc == DTS_ERR_HIP) std::exit(rc);
}

dts_filter_params ToC(const DomainFilterParams& p) {
  return dts_filter_params{p.sigma_x, p.sigma_y, p.sigma_r};
}

dts_optimize_params ToC(const DomainOptimizeParams& p) {
  return dts_optimize_params{p.num_iterations, static_cast<int>(p.loss), p.lambda,
                             p.sigma_z,        p.step_size,             p.photo_consis_lambda,
                             p.DT_lambda};
}

int ToC(const COLOR_SPACE& c) { return static_cast<int>(c); }

int ToC(const ImageType& t) {
  switch (t) {
    case ImageType::COLOR_IMAGE: return DTS_IMAGE_COLOR_IMAGE;
    case ImageType::DIFFERENTIAL: return DTS_IMAGE_DIFFERENTIAL;
    case ImageType::INTEGRAL: return DTS_IMAGE_INTEGRAL;
    case ImageType::TARGET: return DTS_IMAGE_TARGET;
    case ImageType::CONFIDENCE: return DTS_IMAGE_CONFIDENCE;
    case ImageType::OPTIMIZED_QUANTITY: return DTS_IMAGE_OPTIMIZED_QUANTITY;
    default: return -1;
  }
}

}  // namespace

// ------------------------------------------------------ DomainTransformFilter
DomainTransformFilter::DomainTransformFilter(const ImageDim& max_image_dims)
    : DomainTransformFilter(max_image_dims, 0) {}

DomainTransformFilter::DomainTransformFilter(const ImageDim& max_image_dims, int flags)
    : max_image_dims_(max_image_dims), image_dims_(max_image_dims) {
  dts_handle* h = nullptr;
  const int rc = dts_create(max_image_dims.width, max_image_dims.height, flags, &h);
  if (rc != DTS_OK) {
    std::cerr << "DomainTransformFilter: " << dts_last_error() << std::endl;
    std::exit(rc);
  }
  filter_struct_.handle = h;
  dts_get_stream(h, &filter_struct_.stream);
}

DomainTransformFilter::~DomainTransformFilter() { dts_destroy(filter_struct_.handle); }

void DomainTransformFilter::SetImageDim(const ImageDim& d) {
  if (dts_set_image_dim(filter_struct_.handle, d.width, d.height) == DTS_OK) {
    image_dims_ = d;
  } else {
    std::cerr << "Invalid image dims WxH:" << d.width << " x " << d.height << std::endl;
  }
}

void DomainTransformFilter::InitFrame(const COLOR_SPACE& cs, void* color_image) {
  Check(dts_init_guide(filter_struct_.handle, ToC(cs), color_image), "InitFrame");
}

void DomainTransformFilter::Filter(const DomainFilterParams& p, const float* input,
                                   float* output, const int num_iter) {
  const dts_filter_params fp = ToC(p);
  Check(dts_filter(filter_struct_.handle, &fp, input, output, num_iter), "Filter");
}

void DomainTransformFilter::Filter(const DomainFilterParams& p, const int num_iter,
                                   const float* input_dev, float* output_dev) {
  const dts_filter_params fp = ToC(p);
  Check(dts_filter_device(filter_struct_.handle, &fp, input_dev, output_dev, num_iter),
        "Filter");
}

void DomainTransformFilter::RFFilter(const DomainFilterParams& p, const float* input,
                                     float* output, const int num_iter, const bool keep_tape) {
  const dts_filter_params fp = ToC(p);
  Check(dts_rf_filter(filter_struct_.handle, &fp, input, output, num_iter, keep_tape ? 1 : 0, 0),
        "RFFilter");
}

void DomainTransformFilter::RFFilterAdjoint(const float* ybar, float* xbar, float* gbar) {
  Check(dts_rf_filter_adjoint(filter_struct_.handle, ybar, xbar, gbar, 0), "RFFilterAdjoint");
}

void DomainTransformFilter::Download(const ImageType& t, void* image) const {
  Check(dts_download(filter_struct_.handle, ToC(t), image), "Download");
}

void DomainTransformFilter::ClearAll() { Check(dts_clear_all(filter_struct_.handle), "ClearAll"); }

void DomainTransformFilter::ComputeColorSpaceDifferential(const DomainFilterParams& p) {
  const dts_filter_params fp = ToC(p);
  Check(dts_compute_differential(filter_struct_.handle, &fp), "ComputeColorSpaceDifferential");
}

void DomainTransformFilter::IntegrateColorDifferentials() {
  Check(dts_integrate(filter_struct_.handle), "IntegrateColorDifferentials");
}

DomainTransformFilterStruct* DomainTransformFilter::FilterStructPtr() { return &filter_struct_; }

void DomainTransformFilter::Synchronize() const {
  Check(dts_synchronize(filter_struct_.handle), "Synchronize");
}

// ------------------------------------------------------ DomainTransformSolver
DomainTransformSolver::DomainTransformSolver(const ImageDim& max_image_dims)
    : DomainTransformFilter(max_image_dims, 0) {}

DomainTransformSolver::DomainTransformSolver(const ImageDim& max_image_dims, int flags)
    : DomainTransformFilter(max_image_dims, flags) {}

DomainTransformSolver::~DomainTransformSolver() {}

void DomainTransformSolver::InitFrame(const COLOR_SPACE& cs, void* color_image, float* target,
                                      float* confidence) {
  Check(dts_init_frame(filter_struct_.handle, ToC(cs), color_image, target, confidence),
        "InitFrame");
}

void DomainTransformSolver::ComputeColorSpaceDifferential(const DomainFilterParams& p) {
  filter_params_ = p;
  DomainTransformFilter::ComputeColorSpaceDifferential(p);
}

void DomainTransformSolver::IntegrateColorDifferentials() {
  DomainTransformFilter::IntegrateColorDifferentials();
}

void DomainTransformSolver::Optimize(const DomainOptimizeParams& p, const float overwrite) {
  const dts_optimize_params op = ToC(p);
  Check(dts_optimize(filter_struct_.handle, &op, overwrite), "Optimize");
}

void DomainTransformSolver::Download(const ImageType& t, void* image) const {
  DomainTransformFilter::Download(t, image);
}

void DomainTransformSolver::ComputeConfidence(const DomainFilterParamsVec<1>& cp,
                                              const int left_side_clear_width) {
  const dts_filter_params fp{cp.sigma_x, cp.sigma_y, cp.sigma_r};
  Check(dts_compute_confidence(filter_struct_.handle, &fp, cp.sigmas[0], left_side_clear_width),
        "ComputeConfidence");
}

void DomainTransformSolver::ClearAll() { DomainTransformFilter::ClearAll(); }

// ------------------------------------------------------ DomainTransformStereo
DomainTransformStereo::DomainTransformStereo(const ImageDim& max_image_dims)
    : DomainTransformSolver(max_image_dims, DTS_FLAG_STEREO) {}

DomainTransformStereo::~DomainTransformStereo() {}

void DomainTransformStereo::InitRightFrame(const float center_offset, const COLOR_SPACE& cs,
                                           void* color_image) {
  center_offset_ = center_offset;
  Check(dts_init_right_frame(filter_struct_.handle, center_offset, ToC(cs), color_image),
        "InitRightFrame");
}

void DomainTransformStereo::ProcessRightFrame() {
  Check(dts_process_right_frame(filter_struct_.handle), "ProcessRightFrame");
}

void DomainTransformStereo::Optimize(const DomainOptimizeParams& p, const float overwrite) {
  const dts_optimize_params op = ToC(p);
  Check(dts_optimize_stereo(filter_struct_.handle, &op, overwrite), "Optimize");
}

void DomainTransformStereo::SetImageDim(const ImageDim& d) { DomainTransformFilter::SetImageDim(d); }

// ----------------------------------------------------------------------- HFBS
HFBS::HFBS(const ImageDim& max_image_dims, const int max_grid_x, const int max_grid_y,
           const int max_grid_l)
    : image_dim_(max_image_dims) {
  const int rc = hfbs_create(max_image_dims.width, max_image_dims.height, max_grid_x, max_grid_y,
                             max_grid_l, &handle_);
  if (rc != DTS_OK) {
    std::cerr << "HFBS: " << dts_last_error() << std::endl;
    std::exit(rc);
  }
}

HFBS::~HFBS() { hfbs_destroy(handle_); }

void HFBS::SetImageDim(const ImageDim& d) {
  if (hfbs_set_image_dim(handle_, d.width, d.height) == DTS_OK)
    image_dim_ = d;
  else
    std::cerr << "Image size is more than the reserved Cuda memory in constructor." << std::endl;
}

void HFBS::InitFrame(const COLOR_SPACE& cs, void* color_image, float* target, float* confidence) {
  Check(hfbs_init_frame(handle_, ToC(cs), color_image, target, confidence), "InitFrame");
}

void HFBS::SetFilterParams(const HFBSGridParams& gp) {
  grid_params_ = gp;
  const hfbs_grid_params p{gp.sigma_x, gp.sigma_y, gp.sigma_l, gp.lambda, gp.alpha, gp.beta,
                           gp.bistoch_iter_max, gp.optim_iter_max, gp.neg_exp_beta,
                           gp.grid_x, gp.grid_y, gp.grid_l};
  Check(hfbs_set_params(handle_, &p), "SetFilterParams");
}

void HFBS::ClearAll() { Check(hfbs_clear_all(handle_), "ClearAll"); }

void HFBS::Optimize() {
  int fell_back = 0;
  Check(hfbs_optimize(handle_, &fell_back), "Optimize");
  hfbs_debug_grid_dims(handle_, &grid_params_.grid_x, &grid_params_.grid_y, &grid_params_.grid_l);
}

void HFBS::Download(const ImageType& t, float* image) const {
  Check(hfbs_download(handle_, ToC(t), image), "Download");
}

void HFBS::Synchronize() const { Check(hfbs_synchronize(handle_), "Synchronize"); }

}  // namespace domain_transform
