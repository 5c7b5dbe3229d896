// Test harness of the command-line programs' image codec
// (domain_transform_solver_amd/cli/image_io.h):
//   png_dump read <in.png> <out.raw>        header line "W H C D" on stdout and
//                                           the 16-bit samples to out.raw
//   png_dump write16 <W> <H> <in.raw> <out.png>   16-bit gray PNG from raw u16
//   png_dump write8 <W> <H> <in.raw> <out.png>    8-bit gray PNG from raw u16
//   png_dump pfm <in.pfm> <out.raw>         "W H" and float32 rows, top first
//   png_dump bgr <in.png|in.bmp> <out.raw>  "W H" and 8-bit BGR rows, top first
//   png_dump wbmp <W> <H> <in.raw> <out.bmp>      24-bit BMP from raw BGR
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "image_io.h"

int main(int argc, char** argv) {
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
