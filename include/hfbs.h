// HFBS bilateral-grid solver.  Public interface and parameter defaults of the
// reference's src/hfbs.h:18-150.
#ifndef HFBS_H_
#define HFBS_H_

#include "domain_transform_common.h"

struct hfbs_handle;

namespace domain_transform {

struct HFBSGridParams {
  // Grid sigmas in pixels (x, y) and luma units (l); use values >= 1.
  float sigma_x = 8.0f;
  float sigma_y = 8.0f;
  float sigma_l = 4.0f;
  // Smoothness weight.
  float lambda = 8.0f;
  // Heavy-ball step size and momentum.
  float alpha = 1.0f;
  float beta = 0.9f;
  // Bistochastization and solver iterations.
  int bistoch_iter_max = 20;
  int optim_iter_max = 140;
  // Decay of the exponential blur used for the initial guess.
  float neg_exp_beta = 0.9;
  // Derived from the image size and sigmas by Optimize().
  int grid_x = 1;
  int grid_y = 1;
  int grid_l = 1;
};

class HFBS {
 public:
  HFBS(const ImageDim& max_image_dims, const int max_grid_x, const int max_grid_y,
       const int max_grid_l);
  ~HFBS();
  HFBS(const HFBS&) = delete;
  HFBS& operator=(const HFBS&) = delete;

  void SetImageDim(const ImageDim& image_dim);
  // confidence == nullptr means 1 everywhere.
  void InitFrame(const COLOR_SPACE& color_space, void* color_image, float* target,
                 float* confidence = nullptr);
  void SetFilterParams(const HFBSGridParams& grid_params);
  void ClearAll();
  // A grid above the constructor's limits prints a message and leaves the
  // output equal to the target.
  void Optimize();
  void Download(const ImageType& image_type, float* image) const;
  void Synchronize() const;
  hfbs_handle* Handle() const { return handle_; }

 protected:
  hfbs_handle* handle_ = nullptr;
  ImageDim image_dim_;
  HFBSGridParams grid_params_;
};

}  // namespace domain_transform

#endif  // HFBS_H_
