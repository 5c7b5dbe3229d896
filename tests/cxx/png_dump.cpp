// Test harness of the command-line programs' image codec
// (domain_transform_solver_amd/cli/image_io.h):
//   png_dump read <in.png> <out.raw>        header line "W H C D" on stdout and
//                                           the 16-bit samples to out.raw
//   png_dump write16 <W> <H> <in.raw> <out.png>   16-bit gray PNG from raw u16
//   png_dump write8 <W> <H> <in.raw> <out.png>    8-bit gray PNG from raw u16
//   png_dump pfm <in.pfm> <out.raw>         "W H" and float32 rows, top first
//   png_dump bgr <in.png|in.bmp> <out.raw>  "W H" and 8-bit BGR rows, top first
//   png_dump wbmp <W> <H> <in.raw> <out.bmp>      24-bit BMP from raw BGR
//   png_dump check <file>...                decodes every file with each reader
//                                           that applies, printing "<file> <ok>"
//                                           before moving on (fuzz and
//                                           sanitizer runs: a fault names the
//                                           last file printed)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "image_io.h"

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "check") {
    for (int i = 2; i < argc; i++) {
      std::string err;
      std::printf("%s ", argv[i]);
      std::fflush(stdout);
      int w = 0, h = 0;
      std::vector<uint8_t> bgr;
      std::vector<float> f;
      dts_cli::PngImage img;
      const bool a = dts_cli::ReadBgr(argv[i], &w, &h, &bgr, &err);
      const bool b = dts_cli::ReadPng(argv[i], &img, &err);
      const bool c = dts_cli::ReadPfm(argv[i], &w, &h, &f, &err);
      std::printf("%d%d%d\n", a, b, c);
      std::fflush(stdout);
    }
    return 0;
  }
  if (argc < 4) return 2;
  const std::string mode = argv[1];
  std::string err;
  if (mode == "read") {
    dts_cli::PngImage img;
    if (!dts_cli::ReadPng(argv[2], &img, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    std::printf("%d %d %d %d\n", img.width, img.height, img.channels, img.bit_depth);
    std::ofstream f(argv[3], std::ios::binary);
    f.write(reinterpret_cast<const char*>(img.samples.data()),
            (std::streamsize)(img.samples.size() * 2));
    return 0;
  }
  if ((mode == "write16" || mode == "write8") && argc >= 6) {
    const int w = std::atoi(argv[2]), h = std::atoi(argv[3]);
    std::vector<uint16_t> s((size_t)w * h);
    std::ifstream f(argv[4], std::ios::binary);
    f.read(reinterpret_cast<char*>(s.data()), (std::streamsize)(s.size() * 2));
    if (!dts_cli::WritePngGray(argv[5], w, h, mode == "write16" ? 16 : 8, s, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    return 0;
  }
  if (mode == "pfm") {
    int w, h;
    std::vector<float> d;
    if (!dts_cli::ReadPfm(argv[2], &w, &h, &d, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    std::printf("%d %d\n", w, h);
    std::ofstream f(argv[3], std::ios::binary);
    f.write(reinterpret_cast<const char*>(d.data()), (std::streamsize)(d.size() * 4));
    return 0;
  }
  if (mode == "bgr") {
    int w, h;
    std::vector<uint8_t> d;
    if (!dts_cli::ReadBgr(argv[2], &w, &h, &d, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    std::printf("%d %d\n", w, h);
    std::ofstream f(argv[3], std::ios::binary);
    f.write(reinterpret_cast<const char*>(d.data()), (std::streamsize)d.size());
    return 0;
  }
  if (mode == "wbmp" && argc >= 6) {
    const int w = std::atoi(argv[2]), h = std::atoi(argv[3]);
    std::vector<uint8_t> d((size_t)w * h * 3);
    std::ifstream f(argv[4], std::ios::binary);
    f.read(reinterpret_cast<char*>(d.data()), (std::streamsize)d.size());
    if (!dts_cli::WriteBmp(argv[5], w, h, d, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    return 0;
  }
  return 2;
}
