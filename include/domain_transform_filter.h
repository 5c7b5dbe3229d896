// Domain transform filter (normalised-convolution box variant) on the GPU.
// Public interface of the reference's src/domain_transform_filter.h:19-70;
// the device buffers live behind a dts_handle of the C ABI (dts_capi.h).
#ifndef DOMAIN_TRANSFORM_FILTER_H_
#define DOMAIN_TRANSFORM_FILTER_H_

#include "domain_transform_common.h"
#include "domain_transform_filter_params.h"

struct dts_handle;

namespace domain_transform {

// Device state of a filter: the C-ABI handle and its stream.  Stands where
// the reference keeps its libcua buffers (src/domain_transform_filter_struct.cuh).
struct DomainTransformFilterStruct {
  dts_handle* handle = nullptr;
  void* stream = nullptr;
};

class DomainTransformFilter {
 public:
  // Allocates device planes for images up to max_image_dims.
  explicit DomainTransformFilter(const ImageDim& max_image_dims);
  virtual ~DomainTransformFilter();
  DomainTransformFilter(const DomainTransformFilter&) = delete;
  DomainTransformFilter& operator=(const DomainTransformFilter&) = delete;

  // Invalid sizes print a message and keep the previous size.
  virtual void SetImageDim(const ImageDim& image_dim);

  // Uploads a W*H RGBA uint8 guide and converts it to color_space.
  void InitFrame(const COLOR_SPACE& convert_to_color_space, void* color_image);

  // Filters a W*H float host image (num_iter horizontal+vertical passes).
  void Filter(const DomainFilterParams& filter_params, const float* input, float* output,
              const int num_iter = 3);

  // Recursive-filter domain transform of a W*H float host image (not in the
  // reference; see dts_rf_filter in dts_capi.h).  keep_tape keeps what
  // RFFilterAdjoint needs.
  void RFFilter(const DomainFilterParams& filter_params, const float* input, float* output,
                const int num_iter = 3, const bool keep_tape = false);
  // Derivatives of <ybar, RFFilter output> with respect to the input (W*H)
  // and the guide (4*W*H, RGBA; may be null), for the last RFFilter with
  // keep_tape (see dts_rf_filter_adjoint).
  void RFFilterAdjoint(const float* ybar, float* xbar, float* gbar);

  virtual void Download(const ImageType& image_type, void* image) const;

  void ClearAll();

  void ComputeColorSpaceDifferential(const DomainFilterParams& domain_filter_params);
  void IntegrateColorDifferentials();

  DomainTransformFilterStruct* FilterStructPtr();

  // Blocks until the handle's stream is idle.
  void Synchronize() const;

 protected:
  DomainTransformFilter(const ImageDim& max_image_dims, int flags);

  // Filter on device pointers of W*H floats (the reference's overload takes
  // its libcua arrays here).
  void Filter(const DomainFilterParams& filter_params, const int num_iter,
              const float* input_dev, float* output_dev);

  DomainTransformFilterStruct filter_struct_;
  ImageDim max_image_dims_;
  ImageDim image_dims_;
};

}  // namespace domain_transform

#endif  // DOMAIN_TRANSFORM_FILTER_H_
