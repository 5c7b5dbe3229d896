// Colorization benchmark on the MI355X library: the program of
// src/benchmark_colorization_cli.cpp with the same flags, solver presets,
// dataset (<top>/<name>.bmp, <name>_strokes.bmp, <name>_gt.bmp for IMG_3743
// and IMG_3748), stroke-to-confidence rule, colour conversions and printed
// results: per image "<name>, SSIM: [..], PSNR: [..]" on stdout, then
// "RMSE 0, time <mean ms> ms, num_pixels <mean>" on stderr (the reference
// records no RMSE for this task).  The Cb and Cr planes are solved one after
// the other; the result is written as <top>/<algorithm>_<name>.bmp.
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

const char* const kNames[2] = {"IMG_3743", "IMG_3748"};

struct Bgr {
  int w = 0, h = 0;
  std::vector<uint8_t> px;  // h * w * 3, B G R
};

// JPEG YCbCr of an 8-bit BGR image as float planes
// (src/benchmark_colorization_cli.cpp:100-122).
void BgrToYCbCr(const Bgr& im, std::vector<float>* y, std::vector<float>* cb,
                std::vector<float>* cr) {
  const size_t n = (size_t)im.w * im.h;
  y->resize(n);
  cb->resize(n);
  cr->resize(n);
  for (size_t i = 0; i < n; i++) {
    const float b = im.px[3 * i], g = im.px[3 * i + 1], r = im.px[3 * i + 2];
    (*y)[i] = 0.299f * r + 0.587f * g + 0.114f * b;
    (*cb)[i] = 128 - 0.168736f * r - 0.331264f * g + 0.5f * b;
    (*cr)[i] = 128 + 0.5f * r - 0.418688f * g - 0.081312f * b;
  }
}

// The same conversion stored into 8 bits by truncation, as the reference's
// uchar assignment does (src/benchmark_colorization_cli.cpp:124-146).
Bgr BgrToYCbCr8(const Bgr& im) {
  std::vector<float> y, cb, cr;
  BgrToYCbCr(im, &y, &cb, &cr);
  Bgr o = im;
  auto tr = [](float v) { return (uint8_t)(v <= 0 ? 0 : v >= 255 ? 255 : (int)v); };
  for (size_t i = 0; i < y.size(); i++) {
    o.px[3 * i] = tr(y[i]);
    o.px[3 * i + 1] = tr(cb[i]);
    o.px[3 * i + 2] = tr(cr[i]);
  }
  return o;
}

// src/benchmark_colorization_cli.cpp:148-172
Bgr YCbCrToBgr(int w, int h, const std::vector<float>& y, const std::vector<float>& cb,
               const std::vector<float>& cr) {
  Bgr o;
  o.w = w;
  o.h = h;
  o.px.resize((size_t)w * h * 3);
  auto cv8 = [](float v) { return (uint8_t)(v > 255 ? 255 : v < 0 ? 0 : std::round(v)); };
  for (size_t i = 0; i < y.size(); i++) {
    const float r = y[i] + 1.402f * (cr[i] - 128);
    const float g = y[i] - 0.344136f * (cb[i] - 128) - 0.714136 * (cr[i] - 128);
    const float b = y[i] + 1.772f * (cb[i] - 128);
    o.px[3 * i] = cv8(b);
    o.px[3 * i + 1] = cv8(g);
    o.px[3 * i + 2] = cv8(r);
  }
  return o;
}

// Separable 11x11 Gaussian (sigma 1.5) of a float plane, borders reflected
// without repeating the edge pixel (OpenCV's BORDER_REFLECT_101).
std::vector<float> Gauss11(const std::vector<float>& in, int w, int h) {
  constexpr int kR = 5;
  float k[2 * kR + 1];
  double ks = 0;
  for (int i = -kR; i <= kR; i++) ks += std::exp(-(i * i) / (2 * 1.5 * 1.5));
  for (int i = -kR; i <= kR; i++) k[i + kR] = (float)(std::exp(-(i * i) / (2 * 1.5 * 1.5)) / ks);
  auto refl = [](int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i : 2 * n - 2 - i;
    return i;
  };
  std::vector<float> tmp(in.size()), out(in.size());
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float s = 0;
      for (int i = -kR; i <= kR; i++) s += k[i + kR] * in[(size_t)y * w + refl(x + i, w)];
      tmp[(size_t)y * w + x] = s;
    }
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      float s = 0;
      for (int i = -kR; i <= kR; i++) s += k[i + kR] * tmp[(size_t)refl(y + i, h) * w + x];
      out[(size_t)y * w + x] = s;
    }
  return out;
}

// Mean SSIM per channel of two 8-bit 3-channel images
// (src/benchmark_colorization_cli.cpp:45-98); the fourth entry stays 0.
void Ssim(const Bgr& a, const Bgr& b, double out[4]) {
  const float c1 = (0.01f * 255) * (0.01f * 255), c2 = (0.03f * 255) * (0.03f * 255);
  const size_t n = (size_t)a.w * a.h;
  out[3] = 0;
  for (int c = 0; c < 3; c++) {
    std::vector<float> x(n), y(n), xx(n), yy(n), xy(n);
    for (size_t i = 0; i < n; i++) {
      x[i] = a.px[3 * i + c];
      y[i] = b.px[3 * i + c];
      xx[i] = x[i] * x[i];
      yy[i] = y[i] * y[i];
      xy[i] = x[i] * y[i];
    }
    const std::vector<float> m1 = Gauss11(x, a.w, a.h), m2 = Gauss11(y, a.w, a.h),
                             s11 = Gauss11(xx, a.w, a.h), s22 = Gauss11(yy, a.w, a.h),
                             s12 = Gauss11(xy, a.w, a.h);
    double sum = 0;
    for (size_t i = 0; i < n; i++) {
      const float m12 = m1[i] * m2[i], m1s = m1[i] * m1[i], m2s = m2[i] * m2[i];
      const float v1 = s11[i] - m1s, v2 = s22[i] - m2s, v12 = s12[i] - m12;
      sum += ((2 * m12 + c1) * (2 * v12 + c2)) / ((m1s + m2s + c1) * (v1 + v2 + c2));
    }
    out[c] = sum / (double)n;
  }
}

// PSNR of the Cb and Cr channels as the reference defines it,
// 20 log10(255) - 10 log10(RMSE) (src/benchmark_colorization_cli.cpp:22-43).
void PsnrCbCr(const Bgr& gt, const Bgr& test, double out[4]) {
  out[0] = out[3] = 0;
  const size_t n = (size_t)gt.w * gt.h;
  for (int c = 1; c <= 2; c++) {
    double se = 0;
    for (size_t i = 0; i < n; i++) {
      const double d = (double)gt.px[3 * i + c] - test.px[3 * i + c];
      se += d * d;
    }
    const double rmse = std::sqrt(se / (double)n);
    out[c] = (float)(20 * std::log10(255) - 10 * std::log10(rmse));
  }
}

std::string Fmt4(const double v[4]) {
  char b[160];
  std::snprintf(b, sizeof(b), "[%g, %g, %g, %g]", v[0], v[1], v[2], v[3]);
  return b;
}

PlaneImage FloatPlane(int h, int w, const std::vector<float>& v) {
  PlaneImage p(h, w, 4);
  std::memcpy(p.data, v.data(), v.size() * 4);
  return p;
}

// src/benchmark_colorization_cli.cpp:284-345
struct DtsColorization {
  std::unique_ptr<domain_transform::DomainTransformSolver> solver;
  domain_transform::DtsOptions opts;
  explicit DtsColorization(const domain_transform::ImageDim& max_dims)
      : solver(new domain_transform::DomainTransformSolver(max_dims)) {
    opts.filter_options.sigma_x = 64;
    opts.filter_options.sigma_y = 64;
    opts.filter_options.sigma_r = 0.25f;
    opts.optim_options.loss = domain_transform::RobustLoss::CHARBONNIER;
    opts.optim_options.num_iterations = 10;
    opts.optim_options.lambda = 0.99f;
    opts.optim_options.step_size = 0.99f;
    opts.optim_options.sigma_z = 30;
    opts.optim_options.photo_consis_lambda = 0.0f;
    opts.optim_options.DT_lambda = 0.0f;
    opts.conf_params.sigma_x = 8;
    opts.conf_params.sigma_y = 8;
    opts.conf_params.sigma_r = 0.2f;
    opts.conf_params.sigmas[0] = 16;
    opts.color_space = domain_transform::COLOR_SPACE::RGB;
    opts.overwrite_target_above_conf = 0.85f;
  }
  void PreProcess(std::vector<float>*, std::vector<float>*) {}
  float PostValue(float v) const { return v; }
  float Run(const PlaneImage& rgba, const PlaneImage& t, const PlaneImage& c, PlaneImage* out) {
    domain_transform::DtsOptions o = opts;
    o.image_dims.height = rgba.rows;
    o.image_dims.width = rgba.cols;
    return domain_transform::EdgeAwareOptimize(rgba, t, c, o, solver.get(), out);
  }
};

// src/benchmark_colorization_cli.cpp:347-387.  As for every HFBS preset of
// the reference, Solve<HfbsOptions> runs Optimize() with the solver's own
// parameters; the grid parameters set here equal its defaults.
struct HfbsColorization {
  std::unique_ptr<domain_transform::HFBS> solver;
  domain_transform::HfbsOptions opts;
  explicit HfbsColorization(const domain_transform::ImageDim& max_dims)
      : solver(new domain_transform::HFBS(max_dims, (int)std::ceil(3280.0 / 8),
                                          (int)std::ceil(3280.0 / 8), 64)) {
    opts.grid_params.sigma_x = 8;
    opts.grid_params.sigma_y = 8;
    opts.grid_params.sigma_l = 4;
    opts.grid_params.optim_iter_max = 140;
    opts.grid_params.bistoch_iter_max = 20;
    opts.grid_params.lambda = 8.0f;
    opts.color_space = domain_transform::COLOR_SPACE::RGB;
  }
  void PreProcess(std::vector<float>* t, std::vector<float>* c) {
    for (float& v : *t) v = v - 128;
    for (float& v : *c) v = (float)(v * 110000000164.0);
  }
  float PostValue(float v) const { return v + 128; }
  float Run(const PlaneImage& rgba, const PlaneImage& t, const PlaneImage& c, PlaneImage* out) {
    domain_transform::HfbsOptions o = opts;
    o.image_dims.height = rgba.rows;
    o.image_dims.width = rgba.cols;
    return domain_transform::EdgeAwareOptimize(rgba, t, c, o, solver.get(), out);
  }
};

bool ReadBgrImage(const std::string& path, Bgr* im) {
  std::string err;
  if (!dts_cli::ReadBgr(path, &im->w, &im->h, &im->px, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return false;
  }
  return true;
}

template <typename Solver>
int Benchmark(const std::string& top, const std::string& algo) {
  domain_transform::ImageDim max_dims;
  max_dims.width = 3280;
  max_dims.height = 3280;
  Solver s(max_dims);
  double time = 0, pixels = 0;
  for (const char* name : kNames) {
    Bgr gray, marked, gt;
    if (!ReadBgrImage(top + "/" + name + ".bmp", &gray) ||
        !ReadBgrImage(top + "/" + name + "_strokes.bmp", &marked) ||
        !ReadBgrImage(top + "/" + name + "_gt.bmp", &gt))
      return 1;
    if (marked.w != gray.w || marked.h != gray.h || gt.w != gray.w || gt.h != gray.h) {
      std::fprintf(stderr, "image sizes differ for %s\n", name);
      return 1;
    }
    if (gray.w > max_dims.width || gray.h > max_dims.height) {
      std::fprintf(stderr, "%s is larger than %dx%d\n", name, max_dims.width, max_dims.height);
      return 1;
    }
    const int w = gray.w, h = gray.h;
    const size_t n = (size_t)w * h;
    // Guide: the first channel of the gray image as (y, y, y, 0)
    // (src/benchmark_colorization_cli.cpp:206-216).
    PlaneImage yyya(h, w, 4);
    std::vector<float> yplane(n);
    for (size_t i = 0; i < n; i++) {
      const uint8_t v = gray.px[3 * i];
      yyya.data[4 * i] = yyya.data[4 * i + 1] = yyya.data[4 * i + 2] = v;
      yyya.data[4 * i + 3] = 0;
      yplane[i] = v;
    }
    // Strokes: Cb/Cr targets, confidence 1 where either 8-bit chroma leaves
    // [127, 129] (src/benchmark_colorization_cli.cpp:218-246).
    std::vector<float> ym, cb_t, cr_t, cb_c(n), cr_c(n);
    BgrToYCbCr(marked, &ym, &cb_t, &cr_t);
    for (size_t i = 0; i < n; i++) {
      const int cb8 = dts_cli::SaturateU8(cb_t[i]), cr8 = dts_cli::SaturateU8(cr_t[i]);
      const float m = (cb8 > 129 || cb8 < 127 || cr8 > 129 || cr8 < 127) ? 1.0f : 0.0f;
      cb_c[i] = m;
      cr_c[i] = m;
    }
    const Bgr gt_ycbcr = BgrToYCbCr8(gt);

    std::vector<float> est[2];
    float ms = 0;
    for (int ch = 0; ch < 2; ch++) {
      std::vector<float>& t = ch == 0 ? cb_t : cr_t;
      std::vector<float>& c = ch == 0 ? cb_c : cr_c;
      s.PreProcess(&t, &c);
      PlaneImage pt = FloatPlane(h, w, t), pc = FloatPlane(h, w, c), out(h, w, 4);
      ms += s.Run(yyya, pt, pc, &out);
      const float* o = out.ptr<float>();
      est[ch].resize(n);
      for (size_t i = 0; i < n; i++) est[ch][i] = s.PostValue(o[i]);
    }
    time += ms;
    pixels += (double)n;

    const Bgr result = YCbCrToBgr(w, h, yplane, est[0], est[1]);
    std::string err;
    if (!dts_cli::WriteBmp(top + "/" + algo + "_" + name + ".bmp", w, h, result.px, &err))
      std::fprintf(stderr, "%s\n", err.c_str());
    const Bgr res_ycbcr = BgrToYCbCr8(result);
    double ssim[4], psnr[4];
    Ssim(gt_ycbcr, res_ycbcr, ssim);
    PsnrCbCr(gt_ycbcr, res_ycbcr, psnr);
    std::printf("%s, SSIM: %s, PSNR: %s\n", name, Fmt4(ssim).c_str(), Fmt4(psnr).c_str());
    std::fflush(stdout);
  }
  // The reference records an RMSE of 0 for every image (log 0 averaged and
  // exponentiated gives 0).
  std::fprintf(stderr, "RMSE 0, time %g ms, num_pixels %g\n", time / 2, pixels / 2);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  dts_cli::Flags flags;
  flags.Define("path_to_data_top_dir", "",
               "This is the path to the colorization data. The folder should hav the grayscale "
               "image, image with color stroes and gt.");
  flags.Define("algorithm_name", "dts",
               "This is the algorithm name used to process the data. Options: dts | hfbs.");
  bool help = false;
  if (!flags.Parse(argc, argv, &help)) return -1;
  if (help) {
    flags.PrintHelp(argv[0]);
    return 0;
  }
  const std::string top = flags.Get("path_to_data_top_dir");
  const std::string algo = flags.Get("algorithm_name");
  if (top.empty() || algo.empty()) {
    std::fprintf(stderr, "Invalid command line inputs.\n");
    return -1;
  }
  if (algo == "dts") {
    std::printf("Using %s algorithm.\n", algo.c_str());
    return Benchmark<DtsColorization>(top, algo);
  }
  if (algo == "hfbs") {
    std::printf("Using %s algorithm.\n", algo.c_str());
    return Benchmark<HfbsColorization>(top, algo);
  }
  std::fprintf(stderr, "Unsupported algorithm.\n");
  return -1;
}
