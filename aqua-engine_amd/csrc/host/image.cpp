// Image files: reading textures (JPEG) and writing frames -- 8-bit PNG for
// display (the role of the reference's `image` dependency, Cargo.toml:13)
// and float PFM for radiance without loss.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "internal.h"

namespace aqua {

int read_image_file(const std::string& path, AquaImage* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(AQUA_ERR_IO, "cannot open image '%s'", path.c_str());
  std::vector<uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (data.size() >= 2 && data[0] == 0xFF && data[1] == 0xD8) {
    if (int e = decode_jpeg(data, &out->w, &out->h, &out->rgba))
      return fail(e, "image '%s': %s", path.c_str(), last_error_cstr());
    return AQUA_OK;
  }
  return fail(AQUA_ERR_UNSUPPORTED, "image '%s' is not a JPEG file", path.c_str());
}

namespace {

void put32be(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back(uint8_t(x >> 24));
  v.push_back(uint8_t(x >> 16));
  v.push_back(uint8_t(x >> 8));
  v.push_back(uint8_t(x));
}

// One PNG chunk: length, type, data, CRC-32 of type + data (ISO 3309, the
// same CRC zlib computes).
void chunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data) {
  put32be(out, uint32_t(data.size()));
  size_t at = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  uLong crc = crc32(0L, Z_NULL, 0);
  crc = crc32(crc, out.data() + at, uInt(out.size() - at));
  put32be(out, uint32_t(crc));
}

}  // namespace

int write_png(const char* path, uint32_t w, uint32_t h, uint32_t channels, const uint8_t* px) {
  // raw scanlines, each led by filter type 0 (none)
  const size_t row = size_t(w) * channels;
  std::vector<uint8_t> raw(size_t(h) * (row + 1));
  for (uint32_t y = 0; y < h; ++y) {
    raw[size_t(y) * (row + 1)] = 0;
    memcpy(&raw[size_t(y) * (row + 1) + 1], px + size_t(y) * row, row);
  }
  uLongf zlen = compressBound(uLong(raw.size()));
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK)
    return fail(AQUA_ERR_OOM, "png: deflate failed");
  z.resize(zlen);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  std::vector<uint8_t> ihdr;
  put32be(ihdr, w);
  put32be(ihdr, h);
  ihdr.push_back(8);                        // bit depth
  ihdr.push_back(channels == 4 ? 6 : 2);    // colour type: RGBA / RGB
  ihdr.push_back(0);                        // deflate
  ihdr.push_back(0);                        // adaptive filtering
  ihdr.push_back(0);                        // no interlace
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
  FILE* f = fopen(path, "wb");
  if (!f) return fail(AQUA_ERR_IO, "cannot write '%s'", path);
  size_t n = fwrite(out.data(), 1, out.size(), f);
  int rc = fclose(f);
  if (n != out.size() || rc != 0) return fail(AQUA_ERR_IO, "short write to '%s'", path);
  return AQUA_OK;
}

int write_pfm(const char* path, uint32_t w, uint32_t h, const float* rgb) {
  // "PF" (colour), size, scale -1 (little endian); rows stored bottom to top
  FILE* f = fopen(path, "wb");
  if (!f) return fail(AQUA_ERR_IO, "cannot write '%s'", path);
  bool ok = fprintf(f, "PF\n%u %u\n-1.0\n", w, h) > 0;
  for (uint32_t y = h; ok && y-- > 0;) {
    const float* r = rgb + size_t(y) * w * 3;
    for (size_t i = 0; ok && i < size_t(w) * 3; ++i) {
      uint32_t b;
      memcpy(&b, &r[i], 4);
      uint8_t le[4] = {uint8_t(b), uint8_t(b >> 8), uint8_t(b >> 16), uint8_t(b >> 24)};
      ok = fwrite(le, 1, 4, f) == 4;
    }
  }
  if (fclose(f) != 0) ok = false;
  if (!ok) return fail(AQUA_ERR_IO, "short write to '%s'", path);
  return AQUA_OK;
}

}  // namespace aqua

extern "C" {

int aqua_image_load(const char* path, uint32_t* width, uint32_t* height, uint8_t* rgba) {
  if (!path || !width || !height) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  try {
    AquaImage img;
    if (int e = aqua::read_image_file(path, &img)) return e;
    *width = img.w;
    *height = img.h;
    if (rgba) memcpy(rgba, img.rgba.data(), img.rgba.size());
    return AQUA_OK;
  } catch (...) {
    return aqua::fail(AQUA_ERR_OOM, "image '%s': host allocation failed", path);
  }
}

int aqua_image_write_png(const char* path, uint32_t width, uint32_t height, uint32_t channels,
                         const uint8_t* pixels) {
  if (!path || !pixels) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (width == 0 || height == 0 || (channels != 3 && channels != 4))
    return aqua::fail(AQUA_ERR_INVALID_ARG, "png: %ux%u x %u channels", width, height, channels);
  try {
    return aqua::write_png(path, width, height, channels, pixels);
  } catch (...) {
    return aqua::fail(AQUA_ERR_OOM, "png: host allocation failed");
  }
}

int aqua_image_write_pfm(const char* path, uint32_t width, uint32_t height, const float* rgb) {
  if (!path || !rgb) return aqua::fail(AQUA_ERR_INVALID_ARG, "null argument");
  if (width == 0 || height == 0) return aqua::fail(AQUA_ERR_INVALID_ARG, "pfm: empty image");
  return aqua::write_pfm(path, width, height, rgb);
}

}  // extern "C"
