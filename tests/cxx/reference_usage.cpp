// Compile-and-link check of the drop-in headers against the way the
// reference's programs use them: subclasses of the two-argument
// EdgeAwareSolver<S, O> for DTS and HFBS with PreProcess / PostProcessResult
// overrides (src/benchmark_depth_super_res_cli.cpp:25-122), a
// ProblemImageData filled by value (src/depth_super_res_dataset.h:38-85) and
// the EdgeAwareOptimize call on its members.  Nothing here touches a GPU
// unless the program is run.
#include <cstdio>

#include "edge_aware_solver.h"

namespace benchmark {

using domain_transform::DefaultImage;

class DtsStyleSolver : public domain_transform::EdgeAwareSolver<
                           domain_transform::DomainTransformSolver, domain_transform::DtsOptions> {
 public:
  explicit DtsStyleSolver(const domain_transform::ImageDim& max_dims) {
    solver_.reset(new domain_transform::DomainTransformSolver(max_dims));
    solver_options_.filter_options.sigma_x = 8;
    solver_options_.filter_options.sigma_y = 8;
    solver_options_.filter_options.sigma_r = 0.1f;
    solver_options_.optim_options.loss = domain_transform::RobustLoss::L2;
    solver_options_.optim_options.num_iterations = 10;
    solver_options_.color_space = domain_transform::COLOR_SPACE::YCbCr;
  }
  void PreProcess(DefaultImage* rgba, DefaultImage* target, DefaultImage* conf) override {}
  void PostProcessResult(DefaultImage* result) override {}
};

class HfbsStyleSolver : public domain_transform::EdgeAwareSolver<domain_transform::HFBS,
                                                                 domain_transform::HfbsOptions> {
 public:
  explicit HfbsStyleSolver(const domain_transform::ImageDim& max_dims) {
    solver_.reset(new domain_transform::HFBS(max_dims, max_dims.width, max_dims.height, 64));
    solver_options_.grid_params.sigma_x = 8;
    solver_options_.grid_params.sigma_y = 8;
    solver_options_.grid_params.sigma_l = 16;
    solver_options_.color_space = domain_transform::COLOR_SPACE::YCbCr;
  }
  void PreProcess(DefaultImage* rgba, DefaultImage* target, DefaultImage* conf) override {}
  void PostProcessResult(DefaultImage* result) override {}
};

domain_transform::ProblemImageData MakeProblem(int w, int h) {
  domain_transform::ProblemImageData data;
  data.rgba_image = DefaultImage(h, w, 4);
  data.target = DefaultImage(h, w, 4);
  data.confidence = DefaultImage(h, w, 4);
  data.gt = DefaultImage(h, w, 4);
  return data;
}

template <typename EdgeAwareSolverType>
float RunOnce(EdgeAwareSolverType* s, int w, int h) {
  domain_transform::ProblemImageData data = MakeProblem(w, h);
  s->PreProcess(&data.rgba_image, &data.target, &data.confidence);
  auto opt = s->options();
  opt.image_dims.width = w;
  opt.image_dims.height = h;
  DefaultImage result(h, w, 4);
  const float ms = domain_transform::EdgeAwareOptimize(data.rgba_image, data.target,
                                                      data.confidence, opt, &s->solver(), &result);
  s->PostProcessResult(&result);
  return ms;
}

}  // namespace benchmark

int main(int argc, char** argv) {
  domain_transform::ImageDim dims;
  dims.width = 64;
  dims.height = 48;
  benchmark::DtsStyleSolver dts(dims);
  benchmark::HfbsStyleSolver hfbs(dims);
  const float a = benchmark::RunOnce(&dts, 64, 48);
  const float b = benchmark::RunOnce(&hfbs, 64, 48);
  std::printf("dts %.3f ms, hfbs %.3f ms\n", a, b);
  return 0;
}
