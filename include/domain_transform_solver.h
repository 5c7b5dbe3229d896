// Domain Transform Solver (gradient descent on the DT-smoothness problem).
// Public interface of the reference's src/domain_transform_solver.h:34-85.
//
//   DomainTransformSolver solver(max_dims);
//   solver.SetImageDim(dims);
//   solver.InitFrame(COLOR_SPACE::YCbCr, rgba, target, confidence);
//   solver.ComputeColorSpaceDifferential(filter_params);
//   solver.IntegrateColorDifferentials();
//   solver.Optimize(optimize_params);
//   solver.Download(ImageType::OPTIMIZED_QUANTITY, result);
#ifndef DOMAIN_TRANSFORM_SOLVER_H_
#define DOMAIN_TRANSFORM_SOLVER_H_

#include "domain_transform_common.h"
#include "domain_transform_filter.h"
#include "domain_transform_filter_params.h"
#include "domain_transform_optimize.h"

namespace domain_transform {

class DomainTransformSolver : public DomainTransformFilter {
 public:
  explicit DomainTransformSolver(const ImageDim& max_image_dims);
  ~DomainTransformSolver() override;

  // confidence == nullptr fills the confidence with 1.
  void InitFrame(const COLOR_SPACE& color_space, void* color_image, float* target,
                 float* confidence = nullptr);

  void ComputeColorSpaceDifferential(const DomainFilterParams& filter_params);
  void IntegrateColorDifferentials();

  // overwrite_target_above_conf in (0,1) resets pixels whose confidence is
  // above it to the target before each half-step.
  virtual void Optimize(const DomainOptimizeParams& optimize_params,
                        const float overwrite_target_above_conf = -1);

  void Download(const ImageType& image_type, void* image) const override;

  // Edge-aware confidence from the DT variance of the target; columns
  // x <= left_side_clear_width get 0.
  void ComputeConfidence(const DomainFilterParamsVec<1>& conf_params,
                         const int left_side_clear_width);

  void ClearAll();

 protected:
  DomainTransformSolver(const ImageDim& max_image_dims, int flags);

  DomainFilterParams filter_params_{};
};

}  // namespace domain_transform

#endif  // DOMAIN_TRANSFORM_SOLVER_H_
