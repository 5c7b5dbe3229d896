// Stereo refinement: the solver plus a photometric term against a right
// view.  Public interface of the reference's src/domain_transform_stereo.h:36-57.
#ifndef DOMAIN_TRANSFORM_STEREO_H_
#define DOMAIN_TRANSFORM_STEREO_H_

#include "domain_transform_filter.h"
#include "domain_transform_solver.h"

namespace domain_transform {

class DomainTransformStereo : public DomainTransformSolver {
 public:
  explicit DomainTransformStereo(const ImageDim& max_image_dims);
  ~DomainTransformStereo() override;

  // center_offset: offset of the centre pixels of the two views (Middlebury doff).
  void InitRightFrame(const float center_offset, const COLOR_SPACE& color_space,
                      void* color_image);

  // Call after ComputeColorSpaceDifferential().
  void ProcessRightFrame();

  void Optimize(const DomainOptimizeParams& optimize_params,
                const float overwrite_target_above_conf = -1) override;

  void SetImageDim(const ImageDim& image_dim) override;

 protected:
  float center_offset_ = 0;
};

}  // namespace domain_transform

#endif  // DOMAIN_TRANSFORM_STEREO_H_
