// Drop-in check of the C++ class API: the same calls a reference CLI makes
// (src/benchmark_depth_super_res_cli.cpp:130-175 via EdgeAwareOptimize,
// src/edge_aware_solver.h:60-90), compiled against include/ only.
//
// usage: dropin_solve <dts|hfbs|rf> W H in_rgba.u8 in_target.f32 in_conf.f32 out.f32
//        [sigma_x sigma_y sigma_r sigma_z iters color_space]
// Prints "ms=<solve time>" and writes the W*H float result.  rf: the
// recursive-filter transform of the target (DomainTransformFilter::RFFilter,
// iters iterations), then RFFilterAdjoint with the confidence plane as the
// cotangent; out.f32 receives the result, out.f32.xbar and out.f32.gbar the
// input and guide cotangents.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "edge_aware_solver.h"

using namespace domain_transform;

static bool ReadAll(const char* path, PlaneImage* img) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read(reinterpret_cast<char*>(img->data), img->storage.size());
  return static_cast<size_t>(f.gcount()) == img->storage.size();
}

int main(int argc, char** argv) {
  if (argc < 8) {
    std::cerr << "usage: dropin_solve <dts|hfbs|rf> W H rgba target conf out [params]\n";
    return 2;
  }
  const std::string algo = argv[1];
  const int W = std::atoi(argv[2]), H = std::atoi(argv[3]);
  PlaneImage rgba(H, W, 4), target(H, W, 4), conf(H, W, 4), out(H, W, 4);
  if (!ReadAll(argv[4], &rgba) || !ReadAll(argv[5], &target) || !ReadAll(argv[6], &conf)) {
    std::cerr << "cannot read inputs\n";
    return 2;
  }
  auto farg = [&](int i, float d) { return argc > i ? (float)std::atof(argv[i]) : d; };
  ImageDim dims;
  dims.width = W;
  dims.height = H;
  float ms = 0;
  if (algo == "dts") {
    DtsOptions o;
    o.filter_options.sigma_x = farg(8, 64);
    o.filter_options.sigma_y = farg(9, 64);
    o.filter_options.sigma_r = farg(10, 0.1f);
    o.optim_options.loss = RobustLoss::L2;
    o.optim_options.lambda = 0.99f;
    o.optim_options.step_size = 0.99f;
    o.optim_options.sigma_z = farg(11, 35.555557f);
    o.optim_options.num_iterations = (int)farg(12, 10);
    o.conf_params.sigma_x = 8;
    o.conf_params.sigma_y = 8;
    o.conf_params.sigma_r = 0.2f;
    o.conf_params.sigmas[0] = 16;
    o.color_space = static_cast<COLOR_SPACE>((int)farg(13, 1));
    o.image_dims = dims;
    DomainTransformSolver solver(dims);
    ms = EdgeAwareOptimize(rgba, target, conf, o, &solver, &out);
  } else if (algo == "rf") {
    DomainFilterParams fp;
    fp.sigma_x = farg(8, 16);
    fp.sigma_y = farg(9, 16);
    fp.sigma_r = farg(10, 0.3f);
    const int n = (int)farg(12, 3);
    DomainTransformFilter filter(dims);
    filter.SetImageDim(dims);
    filter.InitFrame(static_cast<COLOR_SPACE>((int)farg(13, 0)), rgba.data);
    filter.RFFilter(fp, target.ptr<float>(), out.ptr<float>(), n, true);
    std::vector<float> xbar((size_t)W * H), gbar((size_t)4 * W * H);
    filter.RFFilterAdjoint(conf.ptr<float>(), xbar.data(), gbar.data());
    std::ofstream fx(std::string(argv[7]) + ".xbar", std::ios::binary);
    fx.write(reinterpret_cast<const char*>(xbar.data()), xbar.size() * 4);
    std::ofstream fg(std::string(argv[7]) + ".gbar", std::ios::binary);
    fg.write(reinterpret_cast<const char*>(gbar.data()), gbar.size() * 4);
  } else {
    HfbsOptions o;
    o.grid_params.sigma_x = farg(8, 8);
    o.grid_params.sigma_y = farg(9, 8);
    o.grid_params.sigma_l = farg(10, 8);
    o.grid_params.optim_iter_max = (int)farg(12, 20);
    o.color_space = static_cast<COLOR_SPACE>((int)farg(13, 0));
    o.image_dims = dims;
    HFBS solver(dims, W / 2 + 2, H / 2 + 2, 256);
    solver.SetImageDim(dims);
    solver.SetFilterParams(o.grid_params);
    ms = EdgeAwareOptimize(rgba, target, conf, o, &solver, &out);
  }
  std::ofstream f(argv[7], std::ios::binary);
  f.write(reinterpret_cast<const char*>(out.data), out.storage.size());
  std::printf("ms=%.4f\n", ms);
  return 0;
}
