// Stereo benchmark on the MI355X library: the program of
// src/benchmark_stereo_cli.cpp with the same flag, dataset layout
// (<top>/Art/im0.png, im1.png, im0.png_disp.bin, calib.txt, disp0GT.pfm),
// solver preset and printed results: the minimum and maximum of the
// estimate and of the ground truth, then "RMSE <px>, time <ms> ms,
// num_pixels <n>" on stderr.  The timed region covers ComputeConfidence,
// the differential, the integration, ProcessRightFrame and Optimize, as in
// the reference; the estimate is written as <top>/dts_Art.png (8 bits,
// rounded and clamped, as cv::imwrite stores a float image).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "domain_transform_stereo.h"
#include "edge_aware_solver.h"
#include "flags.h"
#include "image_io.h"

namespace {

// Third line of calib.txt, "doffs=<int>" (src/benchmark_stereo_cli.cpp:120-144):
// the first whitespace-separated token, split at '=', integer part of the
// value; 0 when absent.
int ReadDoff(const std::string& path) {
  std::ifstream f(path);
  if (!f.is_open()) {
    std::fprintf(stderr, "Could not open calibration file at %s\n", path.c_str());
    return 0;
  }
  std::string line, tok;
  std::getline(f, line);
  std::getline(f, line);
  f >> tok;
  std::vector<std::string> parts;
  size_t p = 0;
  while (p <= tok.size()) {
    const size_t q = tok.find('=', p);
    const std::string s = tok.substr(p, q == std::string::npos ? std::string::npos : q - p);
    if (!s.empty()) parts.push_back(s);
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return parts.size() == 2 ? std::stoi(parts[1]) : 0;
}

// Raw float32 disparities of the initialisation, row-major, top row first
// (src/benchmark_stereo_cli.cpp:167-200); zeros when the file is missing.
std::vector<float> ReadTarget(const std::string& path, size_t n) {
  std::vector<float> t(n, 0.0f);
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) {
    std::fprintf(stderr, "Couldnt read the binary initialization at %s\n", path.c_str());
    return t;
  }
  f.read(reinterpret_cast<char*>(t.data()), (std::streamsize)(n * sizeof(float)));
  return t;
}

void MinMax(const char* name, const std::vector<float>& v) {
  float lo = std::numeric_limits<float>::infinity(), hi = -lo;
  for (float x : v) {
    lo = std::min(lo, x);
    hi = std::max(hi, x);
  }
  std::fprintf(stderr, "%s min:%g max:%g\n", name, (double)lo, (double)hi);
}

}  // namespace

int main(int argc, char** argv) {
  dts_cli::Flags flags;
  flags.Define("path_to_data_top_dir", "", "This is the path to the stereo data.");
  bool help = false;
  if (!flags.Parse(argc, argv, &help)) return -1;
  if (help) {
    flags.PrintHelp(argv[0]);
    return 0;
  }
  const std::string top = flags.Get("path_to_data_top_dir");
  if (top.empty()) {
    std::fprintf(stderr, "Invalid command line inputs.\n");
    return -1;
  }
  std::printf("Using DTS algorithm.\n");
  std::fflush(stdout);

  domain_transform::ImageDim max_dims;
  max_dims.width = 768;
  max_dims.height = 768;
  domain_transform::DomainTransformStereo solver(max_dims);

  // src/benchmark_stereo_cli.cpp:203-256
  domain_transform::DtsOptions opts;
  opts.filter_options.sigma_x = 8;
  opts.filter_options.sigma_y = 8;
  opts.filter_options.sigma_r = 0.1f;
  opts.optim_options.loss = domain_transform::RobustLoss::CHARBONNIER;
  opts.optim_options.num_iterations = 1;
  opts.optim_options.lambda = 0.99f;
  opts.optim_options.step_size = 0.99f;
  opts.optim_options.sigma_z = 30;
  opts.optim_options.photo_consis_lambda = 0.01f;
  opts.optim_options.DT_lambda = 0.0f;
  opts.conf_params.sigma_x = 8;
  opts.conf_params.sigma_y = 8;
  opts.conf_params.sigma_r = 0.2f;
  opts.conf_params.sigmas[0] = 16;
  opts.color_space = domain_transform::COLOR_SPACE::RGB;

  const std::string dir = top + "/Art/";
  std::string err;
  int w, h, w1, h1;
  std::vector<uint8_t> left, right;
  if (!dts_cli::ReadRgba(dir + "im0.png", &w, &h, &left, &err) ||
      !dts_cli::ReadRgba(dir + "im1.png", &w1, &h1, &right, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  if (w1 != w || h1 != h || w > max_dims.width || h > max_dims.height) {
    std::fprintf(stderr, "the views must have one size of at most %dx%d\n", max_dims.width,
                 max_dims.height);
    return 1;
  }
  const size_t n = (size_t)w * h;
  std::vector<float> target = ReadTarget(dir + "im0.png_disp.bin", n);
  const int doff = ReadDoff(dir + "calib.txt");
  // Ground truth: Middlebury stores the rows bottom first; infinities
  // (unknown) become 0 (src/benchmark_stereo_cli.cpp:146-165).
  std::vector<float> gt(n, 0.0f);
  {
    int gw = 0, gh = 0;
    std::vector<float> g;
    if (dts_cli::ReadPfm(dir + "disp0GT.pfm", &gw, &gh, &g, &err) && gw == w && gh == h) {
      for (size_t i = 0; i < n; i++) gt[i] = std::isinf(g[i]) ? 0.0f : g[i];
    } else {
      std::fprintf(stderr, "Could not read the groundtruth disparoty pfm file.\n");
    }
  }

  domain_transform::ImageDim dims;
  dims.width = w;
  dims.height = h;
  // src/benchmark_stereo_cli.cpp:258-303 (EdgeAwareStereo)
  solver.SetImageDim(dims);
  solver.ClearAll();
  solver.InitFrame(opts.color_space, left.data(), target.data(), nullptr);
  solver.InitRightFrame((float)doff, opts.color_space, right.data());
  constexpr int kLeftSideClearWidth = 20;
  domain_transform::detail::TimerStart(&solver);
  solver.ComputeConfidence(opts.conf_params, kLeftSideClearWidth);
  solver.ComputeColorSpaceDifferential(opts.filter_options);
  solver.IntegrateColorDifferentials();
  solver.ProcessRightFrame();
  solver.Optimize(opts.optim_options, opts.overwrite_target_above_conf);
  const float ms = domain_transform::detail::TimerStop(&solver);
  std::vector<float> est(n);
  solver.Download(domain_transform::ImageType::OPTIMIZED_QUANTITY, est.data());

  MinMax("disparity_estimate", est);
  MinMax("image_data.disparity_gt", gt);
  double se = 0;
  for (size_t i = 0; i < n; i++) {
    const double d = (double)gt[i] - (double)est[i];
    se += d * d;
  }
  std::vector<uint16_t> out8(n);
  for (size_t i = 0; i < n; i++) out8[i] = dts_cli::SaturateU8(est[i]);
  if (!dts_cli::WritePngGray(top + "/dts_Art.png", w, h, 8, out8, &err))
    std::fprintf(stderr, "%s\n", err.c_str());
  // One image: the geometric mean of the RMSE is the RMSE itself.
  std::fprintf(stderr, "RMSE %g, time %g ms, num_pixels %g\n", std::sqrt(se / (double)n),
               (double)ms, (double)n);
  return 0;
}
