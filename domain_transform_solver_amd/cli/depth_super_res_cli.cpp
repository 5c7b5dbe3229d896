// Depth super-resolution benchmark on the MI355X library: the program of
// src/benchmark_depth_super_res_cli.cpp with the same flags, solver presets,
// dataset walk (src/depth_super_res_dataset.h:15-131), post-processing and
// printed result ("RMSE <geometric mean>, time <mean ms> ms, num_pixels
// <mean>").  Images go through the zlib PNG codec of image_io.h instead of
// OpenCV; results are written as 8-bit PNGs to
// <path_to_data_top_dir>/<algorithm_name>/<scene>/output_<k>_n.png.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "edge_aware_solver.h"
#include "flags.h"
#include "image_io.h"

using domain_transform::PlaneImage;

namespace {

constexpr int kNumImages = 12;
const char* const kFolders[3] = {"art_big", "books_big", "moebius_big"};

// src/benchmark_depth_super_res_cli.cpp:24-85
struct DtsSuperRes {
  std::unique_ptr<domain_transform::DomainTransformSolver> solver;
  domain_transform::DtsOptions opts;
  explicit DtsSuperRes(const domain_transform::ImageDim& max_dims)
      : solver(new domain_transform::DomainTransformSolver(max_dims)) {
    opts.filter_options.sigma_x = 8;
    opts.filter_options.sigma_y = 8;
    opts.filter_options.sigma_r = 0.1f;
    opts.optim_options.loss = domain_transform::RobustLoss::L2;
    opts.optim_options.num_iterations = 10;
    opts.optim_options.lambda = 0.99f;
    opts.optim_options.step_size = 0.99f;
    opts.optim_options.photo_consis_lambda = 0.02f;
    opts.conf_params.sigma_x = 8;
    opts.conf_params.sigma_y = 8;
    opts.conf_params.sigma_r = 0.2f;
    opts.conf_params.sigmas[0] = 16;
    opts.color_space = domain_transform::COLOR_SPACE::YCbCr;
  }
  void PreProcess(std::vector<float>*) {}
  float PostValue(float v) const { return v; }
  float Run(const PlaneImage& rgba, const PlaneImage& t, const PlaneImage& c, int upsample,
            PlaneImage* out) {
    domain_transform::DtsOptions o = opts;
    o.filter_options.sigma_x = 8 * upsample;
    o.filter_options.sigma_y = 8 * upsample;
    o.optim_options.sigma_z = 8 * upsample / (1 + 0.2 * 4);
    o.image_dims.height = rgba.rows;
    o.image_dims.width = rgba.cols;
    return domain_transform::EdgeAwareOptimize(rgba, t, c, o, solver.get(), out);
  }
};

// src/benchmark_depth_super_res_cli.cpp:87-126.  The grid parameters of the
// options are set but, as in the reference, Solve<HfbsOptions> runs
// Optimize() with the solver's own (default) parameters.
struct HfbsSuperRes {
  std::unique_ptr<domain_transform::HFBS> solver;
  domain_transform::HfbsOptions opts;
  explicit HfbsSuperRes(const domain_transform::ImageDim& max_dims)
      : solver(new domain_transform::HFBS(max_dims, max_dims.width, max_dims.height, 64)) {
    opts.grid_params.sigma_x = 8;
    opts.grid_params.sigma_y = 8;
    opts.grid_params.sigma_l = 16;
    opts.color_space = domain_transform::COLOR_SPACE::YCbCr;
  }
  void PreProcess(std::vector<float>* t) {
    for (float& v : *t) v = (v - 0.5f) * 128;
  }
  float PostValue(float v) const { return v / 128.0f + 0.5f; }
  float Run(const PlaneImage& rgba, const PlaneImage& t, const PlaneImage& c, int upsample,
            PlaneImage* out) {
    domain_transform::HfbsOptions o = opts;
    o.grid_params.sigma_x = 8 * upsample;
    o.grid_params.sigma_y = 8 * upsample;
    o.image_dims.height = rgba.rows;
    o.image_dims.width = rgba.cols;
    return domain_transform::EdgeAwareOptimize(rgba, t, c, o, solver.get(), out);
  }
};

PlaneImage FloatPlane(int h, int w, const std::vector<float>& v) {
  PlaneImage p(h, w, 4);
  std::memcpy(p.data, v.data(), v.size() * 4);
  return p;
}

template <typename Solver>
int Benchmark(const std::string& top, const std::string& algo, const std::string& input) {
  domain_transform::ImageDim max_dims;
  max_dims.width = 1536;
  max_dims.height = 1536;
  Solver s(max_dims);
  double log_rmse = 0, time = 0, pixels = 0;
  for (int im = 0; im < kNumImages; im++) {
    const std::string f = kFolders[im / 4];
    const int k = im % 4 + 1;
    const int upsample = 1 << k;
    std::string err;
    int w, h, w2, h2, w3, h3, w4, h4;
    std::vector<uint8_t> rgba, gt;
    std::vector<float> target, conf;
    if (!dts_cli::ReadRgba(top + "/_input/" + f + "/" + f + "_color.png", &w, &h, &rgba, &err) ||
        !dts_cli::ReadFloat(top + "/" + input + "/" + f + "/output_" + std::to_string(k) +
                                "_n.png",
                            1.0f / (256.0f * 256.0f - 1), &w2, &h2, &target, &err) ||
        !dts_cli::ReadGray8(top + "/_gth/" + f + ".png", &w3, &h3, &gt, &err) ||
        !dts_cli::ReadFloat(top + "/weights/" + f + "/weight_" + std::to_string(k) + "_n.png",
                            1.0f / (256.0f * 256.0f - 1), &w4, &h4, &conf, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    if (w2 != w || h2 != h || w3 != w || h3 != h || w4 != w || h4 != h) {
      std::fprintf(stderr, "image sizes differ in %s\n", f.c_str());
      return 1;
    }
    // src/depth_super_res_dataset.h:66-70: confidence normalised by its max.
    float cmax = 0;
    for (float v : conf) cmax = std::max(cmax, v);
    for (float& v : conf) v /= cmax;
    s.PreProcess(&target);
    PlaneImage prgba(h, w, 4);
    std::memcpy(prgba.data, rgba.data(), rgba.size());
    PlaneImage pt = FloatPlane(h, w, target), pc = FloatPlane(h, w, conf), out(h, w, 4);
    const float ms = s.Run(prgba, pt, pc, upsample, &out);
    const float* o = out.ptr<float>();
    std::vector<uint16_t> res((size_t)w * h);
    double se = 0;
    for (size_t i = 0; i < res.size(); i++) {
      // cv::Mat::convertTo(CV_8UC1, 255)
      res[i] = dts_cli::SaturateU8(s.PostValue(o[i]) * 255.0f);
      const double d = (double)gt[i] - (double)res[i];
      se += d * d;
    }
    // src/benchmark_util.h:62-68 and :19-25 (log of the per-image RMSE).
    log_rmse += std::log(std::sqrt(se / (double)res.size()));
    time += ms;
    pixels += (double)w * h;
    const std::string out_path =
        top + "/" + algo + "/" + f + "/output_" + std::to_string(k) + "_n.png";
    if (!dts_cli::WritePngGray(out_path, w, h, 8, res, &err))
      std::fprintf(stderr, "%s\n", err.c_str());
  }
  std::fprintf(stderr, "RMSE %g, time %g ms, num_pixels %g\n", std::exp(log_rmse / kNumImages),
               time / kNumImages, pixels / kNumImages);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  dts_cli::Flags flags;
  flags.Define("path_to_data_top_dir", "",
               "This is the path to the data as provided by Jon Barron(from his site).");
  flags.Define("algorithm_name", "dts",
               "This is the algorithm name used to process the data. Options: dts | hfbs.");
  flags.Define("algorithm_input_name", "bicubic", "This is the algorithm name used as the input.");
  bool help = false;
  if (!flags.Parse(argc, argv, &help)) return -1;
  if (help) {
    flags.PrintHelp(argv[0]);
    return 0;
  }
  const std::string top = flags.Get("path_to_data_top_dir");
  const std::string algo = flags.Get("algorithm_name");
  const std::string input = flags.Get("algorithm_input_name");
  if (top.empty() || algo.empty() || input.empty()) {
    std::fprintf(stderr, "Invalid command line inputs.\n");
    return -1;
  }
  if (algo == "dts") {
    std::printf("Using %s algorithm.\n", algo.c_str());
    return Benchmark<DtsSuperRes>(top, algo, input);
  }
  if (algo == "hfbs") {
    std::printf("Using %s algorithm.\n", algo.c_str());
    return Benchmark<HfbsSuperRes>(top, algo, input);
  }
  std::fprintf(stderr, "Unsupported algorithm.\n");
  return -1;
}
