// Image files for the command-line programs: PNG (1-16-bit gray, gray +
// alpha, RGB, RGBA and palette; non-interlaced) read and written with zlib,
// BMP and PFM.  Replaces the OpenCV calls of the reference's CLIs
// (src/domain_transform_image_io.h, src/pfm_image.cpp) with the same pixel
// conventions.
#ifndef DTS_CLI_IMAGE_IO_H_
#define DTS_CLI_IMAGE_IO_H_

#include <cstdint>
#include <string>
#include <vector>

namespace dts_cli {

// Decoded PNG: channels 1-4 (gray, gray+alpha, RGB, RGBA; palettes are
// expanded to RGB/RGBA), samples of 8 or 16 bits kept in 16-bit storage.
struct PngImage {
  int width = 0, height = 0, channels = 0, bit_depth = 0;
  std::vector<uint16_t> samples;  // height * width * channels
  uint16_t at(int y, int x, int c) const {
    return samples[((size_t)y * width + x) * channels + c];
  }
};

// Returns false (with a message in *err) on unreadable or unsupported files.
bool ReadPng(const std::string& path, PngImage* img, std::string* err);
// Gray PNG of 8 or 16 bits from row-major samples.
bool WritePngGray(const std::string& path, int width, int height, int bit_depth,
                  const std::vector<uint16_t>& samples, std::string* err);

// BMP: uncompressed 1/4/8-bit palette, 16-, 24- and 32-bit pixels (with
// bit-field masks), bottom-up or top-down; returned as 8-bit BGR, top row
// first.  Written as 24-bit bottom-up files.
bool ReadBmp(const std::string& path, int* w, int* h, std::vector<uint8_t>* bgr,
             std::string* err);
bool WriteBmp(const std::string& path, int w, int h, const std::vector<uint8_t>& bgr,
              std::string* err);
// Colour image of a PNG or BMP file (told apart by its signature) as 8-bit
// BGR, what cv::imread(path, IMREAD_COLOR) returns: gray is replicated,
// alpha dropped, 16-bit samples shifted down to 8.
bool ReadBgr(const std::string& path, int* w, int* h, std::vector<uint8_t>* bgr,
             std::string* err);

// Colour image as uchar4 RGBA with alpha 0, the layout
// ReadRGBAImageFrom (src/domain_transform_image_io.h:17-23) produces.
bool ReadRgba(const std::string& path, int* w, int* h, std::vector<uint8_t>* rgba,
              std::string* err);
// Gray value of channel 0 (8 or 16 bit, as stored) times scale.
bool ReadFloat(const std::string& path, float scale, int* w, int* h, std::vector<float>* out,
               std::string* err);
// 8-bit gray image (cv::IMREAD_GRAYSCALE of a gray file).
bool ReadGray8(const std::string& path, int* w, int* h, std::vector<uint8_t>* out,
               std::string* err);

// PFM: little- or big-endian single-channel float maps, rows stored bottom
// first (src/pfm_image.cpp:19-72); returned top row first.
bool ReadPfm(const std::string& path, int* w, int* h, std::vector<float>* out,
             std::string* err);
bool WritePfm(const std::string& path, int w, int h, const std::vector<float>& data,
              std::string* err);

// cv::saturate_cast<uchar>(float): round half to even, clamp to [0, 255].
inline uint8_t SaturateU8(float v) {
  const float r = __builtin_rintf(v);
  return (uint8_t)(r < 0 ? 0 : r > 255 ? 255 : r);
}

}  // namespace dts_cli

#endif  // DTS_CLI_IMAGE_IO_H_
