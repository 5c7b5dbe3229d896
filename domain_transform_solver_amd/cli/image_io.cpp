// PNG (on zlib), BMP and PFM files; see image_io.h.
#include "image_io.h"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace dts_cli {
namespace {

// Largest width or height accepted from a file: the library's limit, which
// also keeps every size product below in 64-bit range.
constexpr int64_t kMaxDim = 32767;
// zlib's largest compression ratio (deflate emits at least one bit per 258
// output bytes); inflated sizes beyond it mean a corrupt header.
constexpr uint64_t kMaxInflateRatio = 1032;

uint32_t be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

void put_be32(std::vector<uint8_t>* v, uint32_t x) {
  v->push_back((uint8_t)(x >> 24));
  v->push_back((uint8_t)(x >> 16));
  v->push_back((uint8_t)(x >> 8));
  v->push_back((uint8_t)x);
}

bool fail(std::string* err, const std::string& msg) {
  if (err) *err = msg;
  return false;
}

bool read_file(const std::string& path, std::vector<uint8_t>* buf) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  buf->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

uint32_t le32(const uint8_t* p) {
  return ((uint32_t)p[3] << 24) | ((uint32_t)p[2] << 16) | ((uint32_t)p[1] << 8) | p[0];
}
uint16_t le16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }

void put_le32(std::vector<uint8_t>* v, uint32_t x) {
  for (int i = 0; i < 4; i++) v->push_back((uint8_t)(x >> (8 * i)));
}
void put_le16(std::vector<uint8_t>* v, uint16_t x) {
  v->push_back((uint8_t)x);
  v->push_back((uint8_t)(x >> 8));
}

// Value of the bits of a BMP channel mask, scaled to 8 bits.
uint8_t mask_channel(uint32_t px, uint32_t mask) {
  if (!mask) return 0;
  int shift = 0, bits = 0;
  while (!((mask >> shift) & 1)) shift++;
  while (shift + bits < 32 && ((mask >> (shift + bits)) & 1)) bits++;
  const uint32_t v = (px & mask) >> shift;
  const uint32_t mx = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1);
  return (uint8_t)(bits >= 8 ? v >> (bits - 8) : (v * 255 + mx / 2) / mx);
}

uint8_t paeth(uint8_t a, uint8_t b, uint8_t c) {
  const int p = (int)a + b - c;
  const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

}  // namespace

bool ReadPng(const std::string& path, PngImage* img, std::string* err) {
  std::vector<uint8_t> buf;
  if (!read_file(path, &buf)) return fail(err, "cannot read " + path);
  static const uint8_t kSig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (buf.size() < 8 || std::memcmp(buf.data(), kSig, 8) != 0)
    return fail(err, path + ": not a PNG file");
  int w = 0, h = 0, depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  size_t pos = 8;
  while (pos + 12 <= buf.size()) {
    const uint32_t len = be32(&buf[pos]);
    if (pos + 12 + (size_t)len > buf.size()) return fail(err, path + ": truncated chunk");
    const char* type = reinterpret_cast<const char*>(&buf[pos + 4]);
    const uint8_t* d = &buf[pos + 8];
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len < 13) return fail(err, path + ": bad IHDR");
      const uint32_t w32 = be32(d), h32 = be32(d + 4);
      if (w32 > (uint32_t)kMaxDim || h32 > (uint32_t)kMaxDim)
        return fail(err, path + ": image larger than 32767 pixels on a side");
      w = (int)w32;
      h = (int)h32;
      depth = d[8];
      ctype = d[9];
      interlace = d[12];
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(d, d + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(d, d + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), d, d + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    pos += 12 + len;
  }
  if (w <= 0 || h <= 0) return fail(err, path + ": no image header");
  if (interlace != 0) return fail(err, path + ": interlaced PNG is not supported");
  if (depth != 1 && depth != 2 && depth != 4 && depth != 8 && depth != 16)
    return fail(err, path + ": bad bit depth");
  int src_ch;
  switch (ctype) {
    case 0: src_ch = 1; break;
    case 2: src_ch = 3; break;
    case 3: src_ch = 1; break;
    case 4: src_ch = 2; break;
    case 6: src_ch = 4; break;
    default: return fail(err, path + ": unknown colour type");
  }
  // Depths below 8 are allowed for gray and palette images only.
  if (depth < 8 && ctype != 0 && ctype != 3) return fail(err, path + ": bad bit depth");
  if (ctype == 3 && (depth == 16 || plte.size() < 3)) return fail(err, path + ": bad palette");
  // Filters work on whole bytes: the byte distance to the previous pixel is
  // at least 1 for packed depths.
  const size_t bpp = std::max<size_t>(1, (size_t)src_ch * depth / 8);
  const size_t stride = ((size_t)w * src_ch * depth + 7) / 8;
  // w, h <= 32767: the products stay below 2^36.  A size the compressed
  // data cannot inflate to is rejected before it is allocated.
  const uint64_t raw_bytes = (uint64_t)(stride + 1) * (uint64_t)h;
  if (raw_bytes > kMaxInflateRatio * (uint64_t)idat.size() + 1024)
    return fail(err, path + ": corrupt image data");
  std::vector<uint8_t> raw((size_t)raw_bytes);
  uLongf raw_len = (uLongf)raw.size();
  if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK ||
      raw_len != raw.size())
    return fail(err, path + ": corrupt image data");
  // Undo the per-row filters in place.
  std::vector<uint8_t> prev(stride, 0);
  for (int y = 0; y < h; y++) {
    uint8_t* row = &raw[(stride + 1) * y];
    const int ft = row[0];
    uint8_t* r = row + 1;
    for (size_t i = 0; i < stride; i++) {
      const uint8_t a = i >= bpp ? r[i - bpp] : 0;
      const uint8_t b = prev[i];
      const uint8_t c = i >= bpp ? prev[i - bpp] : 0;
      switch (ft) {
        case 0: break;
        case 1: r[i] = (uint8_t)(r[i] + a); break;
        case 2: r[i] = (uint8_t)(r[i] + b); break;
        case 3: r[i] = (uint8_t)(r[i] + (uint8_t)(((int)a + b) >> 1)); break;
        case 4: r[i] = (uint8_t)(r[i] + paeth(a, b, c)); break;
        default: return fail(err, path + ": bad row filter");
      }
    }
    std::memcpy(prev.data(), r, stride);
  }
  const bool alpha_from_trns = ctype == 3 && !trns.empty();
  img->width = w;
  img->height = h;
  img->bit_depth = depth == 16 ? 16 : 8;
  img->channels = ctype == 3 ? (alpha_from_trns ? 4 : 3) : src_ch;
  img->samples.assign((size_t)w * h * img->channels, 0);
  for (int y = 0; y < h; y++) {
    const uint8_t* r = &raw[(stride + 1) * y + 1];
    for (int x = 0; x < w; x++) {
      uint16_t* out = &img->samples[((size_t)y * w + x) * img->channels];
      if (depth < 8) {
        // Packed samples, most significant bits first.
        const int bit = x * depth;
        const int v = (r[bit >> 3] >> (8 - depth - (bit & 7))) & ((1 << depth) - 1);
        if (ctype == 0) {
          out[0] = (uint16_t)(v * 255 / ((1 << depth) - 1));  // gray scaled to 8 bits
          continue;
        }
        if ((size_t)(3 * v + 2) >= plte.size()) return fail(err, path + ": palette index");
        out[0] = plte[3 * v];
        out[1] = plte[3 * v + 1];
        out[2] = plte[3 * v + 2];
        if (alpha_from_trns) out[3] = (size_t)v < trns.size() ? trns[v] : 255;
        continue;
      }
      if (ctype == 3) {
        const int idx = r[x];
        if ((size_t)(3 * idx + 2) >= plte.size()) return fail(err, path + ": palette index");
        out[0] = plte[3 * idx];
        out[1] = plte[3 * idx + 1];
        out[2] = plte[3 * idx + 2];
        if (alpha_from_trns) out[3] = (size_t)idx < trns.size() ? trns[idx] : 255;
        continue;
      }
      for (int c = 0; c < src_ch; c++) {
        const size_t o = ((size_t)x * src_ch + c) * (depth / 8);
        out[c] = depth == 8 ? r[o] : (uint16_t)((r[o] << 8) | r[o + 1]);
      }
    }
  }
  return true;
}

bool WritePngGray(const std::string& path, int width, int height, int bit_depth,
                  const std::vector<uint16_t>& samples, std::string* err) {
  if (bit_depth != 8 && bit_depth != 16) return fail(err, "bit depth must be 8 or 16");
  if (samples.size() != (size_t)width * height) return fail(err, "sample count mismatch");
  const size_t bps = bit_depth / 8;
  std::vector<uint8_t> raw;
  raw.reserve(((size_t)width * bps + 1) * height);
  for (int y = 0; y < height; y++) {
    raw.push_back(0);
    for (int x = 0; x < width; x++) {
      const uint16_t v = samples[(size_t)y * width + x];
      if (bps == 2) raw.push_back((uint8_t)(v >> 8));
      raw.push_back((uint8_t)(bps == 2 ? v & 0xff : v));
    }
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
    return fail(err, "compression failed");
  z.resize(zlen);
  std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
  auto chunk = [&](const char* type, const std::vector<uint8_t>& data) {
    put_be32(&out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc =
        (uint32_t)crc32(0L, &out[start], (uInt)(out.size() - start));
    put_be32(&out, crc);
  };
  std::vector<uint8_t> ihdr;
  put_be32(&ihdr, (uint32_t)width);
  put_be32(&ihdr, (uint32_t)height);
  ihdr.push_back((uint8_t)bit_depth);
  ihdr.push_back(0);  // gray
  ihdr.push_back(0);
  ihdr.push_back(0);
  ihdr.push_back(0);
  chunk("IHDR", ihdr);
  chunk("IDAT", z);
  chunk("IEND", {});
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(err, "cannot write " + path);
  f.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)out.size());
  return (bool)f || fail(err, "write failed: " + path);
}

bool ReadBmp(const std::string& path, int* w, int* h, std::vector<uint8_t>* bgr,
             std::string* err) {
  std::vector<uint8_t> buf;
  if (!read_file(path, &buf)) return fail(err, "cannot read " + path);
  if (buf.size() < 54 || buf[0] != 'B' || buf[1] != 'M') return fail(err, path + ": not a BMP file");
  const uint32_t off = le32(&buf[10]);
  const uint32_t hsize = le32(&buf[14]);
  if (hsize < 40 || 14 + (size_t)hsize > buf.size())
    return fail(err, path + ": unsupported BMP header");
  const int32_t width = (int32_t)le32(&buf[18]);
  const int32_t height_signed = (int32_t)le32(&buf[22]);
  const int bpp = le16(&buf[28]);
  const uint32_t comp = le32(&buf[30]);
  uint32_t ncol = le32(&buf[46]);
  const bool top_down = height_signed < 0;
  // -INT32_MIN is not an int: such a height is rejected as too large.
  const int64_t height = top_down ? -(int64_t)height_signed : (int64_t)height_signed;
  if (width <= 0 || height <= 0 || width > kMaxDim || height > kMaxDim)
    return fail(err, path + ": bad size");
  if (comp != 0 && !(comp == 3 && (bpp == 16 || bpp == 32)))
    return fail(err, path + ": compressed BMP is not supported");
  if (bpp != 1 && bpp != 4 && bpp != 8 && bpp != 16 && bpp != 24 && bpp != 32)
    return fail(err, path + ": unsupported bits per pixel");
  // Channel masks of 16- and 32-bit pixels: the defaults of uncompressed
  // files, or the bit fields after the 40-byte header.
  uint32_t rm = bpp == 16 ? 0x7c00u : 0x00ff0000u, gm = bpp == 16 ? 0x03e0u : 0x0000ff00u,
           bm = bpp == 16 ? 0x001fu : 0x000000ffu;
  if (comp == 3) {
    if (14 + 40 + 12 > buf.size()) return fail(err, path + ": truncated bit fields");
    rm = le32(&buf[54]);
    gm = le32(&buf[58]);
    bm = le32(&buf[62]);
  }
  std::vector<uint8_t> pal;
  if (bpp <= 8) {
    if (ncol == 0) ncol = 1u << bpp;
    const size_t p0 = 14 + hsize;
    if (ncol > 256 || p0 + 4 * (size_t)ncol > buf.size()) return fail(err, path + ": bad palette");
    pal.assign(buf.begin() + p0, buf.begin() + p0 + 4 * ncol);
  }
  const size_t stride = (((size_t)width * bpp + 31) / 32) * 4;
  if ((size_t)off + stride * (size_t)height > buf.size()) return fail(err, path + ": truncated");
  *w = width;
  *h = (int)height;
  bgr->assign((size_t)width * height * 3, 0);
  for (int y = 0; y < height; y++) {
    const uint8_t* r = &buf[off + stride * (size_t)(top_down ? y : height - 1 - y)];
    uint8_t* o = &(*bgr)[(size_t)y * width * 3];
    for (int x = 0; x < width; x++, o += 3) {
      if (bpp <= 8) {
        const int bit = x * bpp;
        const uint32_t idx = (r[bit >> 3] >> (8 - bpp - (bit & 7))) & ((1u << bpp) - 1);
        if (idx >= ncol) return fail(err, path + ": palette index");
        std::memcpy(o, &pal[4 * idx], 3);
      } else if (bpp == 24) {
        std::memcpy(o, r + 3 * x, 3);
      } else {
        const uint32_t px = bpp == 16 ? le16(r + 2 * x) : le32(r + 4 * x);
        o[0] = mask_channel(px, bm);
        o[1] = mask_channel(px, gm);
        o[2] = mask_channel(px, rm);
      }
    }
  }
  return true;
}

bool WriteBmp(const std::string& path, int w, int h, const std::vector<uint8_t>& bgr,
              std::string* err) {
  if (w <= 0 || h <= 0 || bgr.size() != (size_t)w * h * 3) return fail(err, "bad image size");
  const uint32_t stride = (uint32_t)(((size_t)w * 24 + 31) / 32) * 4;
  std::vector<uint8_t> out;
  out.push_back('B');
  out.push_back('M');
  put_le32(&out, 54 + stride * (uint32_t)h);
  put_le32(&out, 0);
  put_le32(&out, 54);
  put_le32(&out, 40);
  put_le32(&out, (uint32_t)w);
  put_le32(&out, (uint32_t)h);  // bottom row first
  put_le16(&out, 1);
  put_le16(&out, 24);
  put_le32(&out, 0);
  put_le32(&out, stride * (uint32_t)h);
  for (int i = 0; i < 4; i++) put_le32(&out, 0);
  for (int y = h - 1; y >= 0; y--) {
    const uint8_t* r = &bgr[(size_t)y * w * 3];
    out.insert(out.end(), r, r + (size_t)w * 3);
    out.insert(out.end(), stride - (size_t)w * 3, 0);
  }
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(err, "cannot write " + path);
  f.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)out.size());
  return (bool)f || fail(err, "write failed: " + path);
}

bool ReadBgr(const std::string& path, int* w, int* h, std::vector<uint8_t>* bgr,
             std::string* err) {
  std::ifstream f(path, std::ios::binary);
  char sig[2] = {0, 0};
  if (!f || !f.read(sig, 2)) return fail(err, "cannot read " + path);
  f.close();
  if (sig[0] == 'B' && sig[1] == 'M') return ReadBmp(path, w, h, bgr, err);
  PngImage img;
  if (!ReadPng(path, &img, err)) return false;
  *w = img.width;
  *h = img.height;
  bgr->assign((size_t)img.width * img.height * 3, 0);
  const int sh = img.bit_depth == 16 ? 8 : 0;
  for (int y = 0; y < img.height; y++)
    for (int x = 0; x < img.width; x++) {
      uint8_t* o = &(*bgr)[((size_t)y * img.width + x) * 3];
      if (img.channels >= 3) {
        for (int c = 0; c < 3; c++) o[2 - c] = (uint8_t)(img.at(y, x, c) >> sh);
      } else {
        o[0] = o[1] = o[2] = (uint8_t)(img.at(y, x, 0) >> sh);
      }
    }
  return true;
}

bool ReadRgba(const std::string& path, int* w, int* h, std::vector<uint8_t>* rgba,
              std::string* err) {
  std::vector<uint8_t> bgr;
  if (!ReadBgr(path, w, h, &bgr, err)) return false;
  const size_t n = (size_t)*w * *h;
  rgba->assign(n * 4, 0);
  for (size_t i = 0; i < n; i++) {
    (*rgba)[4 * i + 0] = bgr[3 * i + 2];
    (*rgba)[4 * i + 1] = bgr[3 * i + 1];
    (*rgba)[4 * i + 2] = bgr[3 * i + 0];
  }
  return true;
}

bool ReadFloat(const std::string& path, float scale, int* w, int* h, std::vector<float>* out,
               std::string* err) {
  PngImage img;
  if (!ReadPng(path, &img, err)) return false;
  if (img.channels != 1) return fail(err, path + ": expected a gray image");
  *w = img.width;
  *h = img.height;
  out->resize((size_t)img.width * img.height);
  for (size_t i = 0; i < out->size(); i++) (*out)[i] = (float)img.samples[i] * scale;
  return true;
}

bool ReadGray8(const std::string& path, int* w, int* h, std::vector<uint8_t>* out,
               std::string* err) {
  PngImage img;
  if (!ReadPng(path, &img, err)) return false;
  if (img.channels != 1) return fail(err, path + ": expected a gray image");
  *w = img.width;
  *h = img.height;
  out->resize((size_t)img.width * img.height);
  const int sh = img.bit_depth == 16 ? 8 : 0;
  for (size_t i = 0; i < out->size(); i++) (*out)[i] = (uint8_t)(img.samples[i] >> sh);
  return true;
}

bool ReadPfm(const std::string& path, int* w, int* h, std::vector<float>* out,
             std::string* err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(err, "cannot read " + path);
  std::string magic;
  double scale = 0;
  f >> magic >> *w >> *h >> scale;
  if (!f) return fail(err, path + ": bad PFM header");
  f.get();  // the single whitespace after the header
  if (magic != "Pf") return fail(err, path + ": only single-channel PFM (Pf) is supported");
  if (*w <= 0 || *h <= 0 || *w > kMaxDim || *h > kMaxDim) return fail(err, path + ": bad size");
  {
    // The float rows must be in the file before they are allocated.
    const std::streamoff here = f.tellg();
    f.seekg(0, std::ios::end);
    const std::streamoff end = f.tellg();
    f.seekg(here);
    if (here < 0 || end - here < (std::streamoff)((int64_t)*w * *h * 4))
      return fail(err, path + ": truncated");
  }
  const bool little = scale < 0;
  std::vector<uint8_t> raw((size_t)*w * *h * 4);
  f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)raw.size());
  if (!f) return fail(err, path + ": truncated");
  out->resize((size_t)*w * *h);
  for (int y = 0; y < *h; y++)
    for (int x = 0; x < *w; x++) {
      const uint8_t* p = &raw[((size_t)(*h - 1 - y) * *w + x) * 4];
      uint32_t u = little ? ((uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0])
                          : ((uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]);
      float v;
      std::memcpy(&v, &u, 4);
      (*out)[(size_t)y * *w + x] = v;
    }
  return true;
}

bool WritePfm(const std::string& path, int w, int h, const std::vector<float>& data,
              std::string* err) {
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(err, "cannot write " + path);
  f << "Pf\n" << w << " " << h << "\n-1\n";
  for (int y = h - 1; y >= 0; y--)
    f.write(reinterpret_cast<const char*>(&data[(size_t)y * w]), (std::streamsize)w * 4);
  return (bool)f || fail(err, "write failed: " + path);
}

}  // namespace dts_cli
