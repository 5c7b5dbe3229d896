// Solver-agnostic entry points: option bundles, Solve<> and
// EdgeAwareOptimize<>.  Interface of the reference's src/edge_aware_solver.h:13-118.
// EdgeAwareOptimize takes any image type with a `data` pointer to tightly
// packed rows.  DefaultImage is cv::Mat when OpenCV's headers are available
// and domain_transform::PlaneImage otherwise; ProblemImageData and
// EdgeAwareSolver<S, O> use it, so code written against the reference's
// header (`EdgeAwareSolver<DomainTransformSolver, DtsOptions>`,
// `ProblemImageData data;`) compiles unchanged.
#ifndef EDGE_AWARE_SOLVER_H_
#define EDGE_AWARE_SOLVER_H_

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#if defined(__has_include)
#if __has_include(<opencv2/core.hpp>)
#include <opencv2/core.hpp>
#define DTS_AMD_HAVE_OPENCV 1
#endif
#endif

#include "dts_capi.h"
#include "domain_transform_solver.h"
#include "hfbs.h"

namespace domain_transform {

struct DtsOptions {
  DomainFilterParams filter_options;
  DomainOptimizeParams optim_options;
  DomainFilterParamsVec<1> conf_params;
  COLOR_SPACE color_space;
  ImageDim image_dims;
  float overwrite_target_above_conf = -1;
};

struct HfbsOptions {
  HFBSGridParams grid_params;
  domain_transform::COLOR_SPACE color_space;
  ImageDim image_dims;
  float overwrite_target_above_conf = -1;
};

template <typename SolverOptions, typename SolverType>
void Solve(const SolverOptions& solver_options, const bool compute_confidence,
           SolverType* solver);

// DTS: differential, integration, optional confidence, optimisation.
template <>
inline void Solve(const DtsOptions& o, const bool compute_confidence,
                  DomainTransformSolver* solver) {
  solver->ComputeColorSpaceDifferential(o.filter_options);
  solver->IntegrateColorDifferentials();
  if (compute_confidence) solver->ComputeConfidence(o.conf_params, 0);
  solver->Optimize(o.optim_options, o.overwrite_target_above_conf);
}

// HFBS: the reference runs Optimize() with the solver's current parameters
// (SetFilterParams is the caller's business).
template <>
inline void Solve(const HfbsOptions&, const bool, HFBS* solver) {
  solver->Optimize();
}

namespace detail {
inline void TimerStart(DomainTransformFilter* s) {
  dts_timer_start(s->FilterStructPtr()->handle);
}
inline float TimerStop(DomainTransformFilter* s) {
  float ms = 0;
  dts_timer_stop(s->FilterStructPtr()->handle, &ms);
  return ms;
}
inline void TimerStart(HFBS* s) { hfbs_timer_start(s->Handle()); }
inline float TimerStop(HFBS* s) {
  float ms = 0;
  hfbs_timer_stop(s->Handle(), &ms);
  return ms;
}
}  // namespace detail

// Uploads the frame, times Solve<> with GPU events on the solver's stream
// (upload and download excluded, as in the reference) and downloads the
// result.  Returns the solve time in milliseconds.
template <typename SolverOptions, typename SolverType, typename ImageT>
float EdgeAwareOptimize(const ImageT& rgba_image, const ImageT& target_image,
                        const ImageT& confidence, const SolverOptions& solver_options,
                        SolverType* solver, ImageT* optim_image) {
  solver->SetImageDim(solver_options.image_dims);
  solver->ClearAll();
  solver->InitFrame(solver_options.color_space, (void*)rgba_image.data,
                    reinterpret_cast<float*>(target_image.data),
                    reinterpret_cast<float*>(confidence.data));
  detail::TimerStart(solver);
  Solve(solver_options, confidence.data == nullptr, solver);
  const float ms = detail::TimerStop(solver);
  solver->Download(ImageType::OPTIMIZED_QUANTITY, reinterpret_cast<float*>(optim_image->data));
  return ms;
}

// Minimal host image usable with the templates above when OpenCV is not
// around: rows x cols pixels of elem_bytes each, tightly packed, like a
// continuous cv::Mat.
struct PlaneImage {
  int rows = 0, cols = 0, elem_bytes = 0;
  std::vector<unsigned char> storage;
  unsigned char* data = nullptr;
  PlaneImage() = default;
  PlaneImage(int r, int c, int bytes)
      : rows(r), cols(c), elem_bytes(bytes), storage((size_t)r * c * bytes) {
    data = storage.data();
  }
  // Copies own their pixels: `data` always points into this object's storage.
  PlaneImage(const PlaneImage& o)
      : rows(o.rows), cols(o.cols), elem_bytes(o.elem_bytes), storage(o.storage) {
    data = storage.empty() ? nullptr : storage.data();
  }
  PlaneImage(PlaneImage&& o) noexcept
      : rows(o.rows), cols(o.cols), elem_bytes(o.elem_bytes), storage(std::move(o.storage)) {
    data = storage.empty() ? nullptr : storage.data();
    o.data = nullptr;
  }
  PlaneImage& operator=(PlaneImage o) noexcept {
    rows = o.rows;
    cols = o.cols;
    elem_bytes = o.elem_bytes;
    storage.swap(o.storage);
    data = storage.empty() ? nullptr : storage.data();
    return *this;
  }
  bool empty() const { return storage.empty(); }
  template <typename T>
  T* ptr() {
    return reinterpret_cast<T*>(data);
  }
};

#ifdef DTS_AMD_HAVE_OPENCV
using DefaultImage = cv::Mat;
#else
using DefaultImage = PlaneImage;
#endif

// Image data of one edge-aware problem (src/edge_aware_solver.h:88-94), for
// any image type.
template <typename ImageT>
struct ProblemImageDataT {
  ImageT rgba_image;
  ImageT target;
  ImageT confidence;
  ImageT gt;
};

// The reference's struct: the four images of one problem as DefaultImage.
struct ProblemImageData : ProblemImageDataT<DefaultImage> {};

// Per-application solver wrapper with pre/post-processing hooks
// (src/edge_aware_solver.h:96-118); ImageT defaults to DefaultImage, so the
// reference's two-argument form names the same class.
template <typename SolverType, typename SolverOptionsType, typename ImageT = DefaultImage>
class EdgeAwareSolver {
 public:
  EdgeAwareSolver() {}
  virtual ~EdgeAwareSolver() {}
  virtual void PreProcess(ImageT* rgba_image, ImageT* target_image, ImageT* confidence) {}
  virtual void PostProcessResult(ImageT* optimized_result) {}
  SolverType& solver() { return *solver_; }
  const SolverOptionsType& options() { return solver_options_; }

 protected:
  std::unique_ptr<SolverType> solver_;
  SolverOptionsType solver_options_;
};

}  // namespace domain_transform

#endif  // EDGE_AWARE_SOLVER_H_
