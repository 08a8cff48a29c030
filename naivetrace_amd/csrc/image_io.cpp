// Linear HDR image files: PFM and OpenEXR (uncompressed scanlines, 32-bit float R, G, B), the
// floating-point counterpart of the reference's 8-bit sRGB output (src/renderbuffer.hxx:99-125).
//
// PFM: "PF\n<w> <h>\n-1.0\n", little-endian float RGB, bottom row first.
// OpenEXR 2.0 single-part scanline file: magic 20000630, version 2, header attributes
// channels (B, G, R as FLOAT), compression NONE, dataWindow / displayWindow, lineOrder
// INCREASING_Y, pixelAspectRatio, screenWindowCenter, screenWindowWidth; then one offset per
// scanline and the scanline chunks (y, byte count, the B, G and R rows).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ttgpu.h"

namespace {

void Put(std::vector<uint8_t> &b, const void *p, size_t n) {
  const uint8_t *c = (const uint8_t *)p;
  b.insert(b.end(), c, c + n);
}
void PutI32(std::vector<uint8_t> &b, int32_t v) { Put(b, &v, 4); }  // the host is little-endian
void PutU64(std::vector<uint8_t> &b, uint64_t v) { Put(b, &v, 8); }
void PutF32(std::vector<uint8_t> &b, float v) { Put(b, &v, 4); }
void PutStr(std::vector<uint8_t> &b, const char *s) { Put(b, s, std::strlen(s) + 1); }
void Attr(std::vector<uint8_t> &b, const char *name, const char *type, const std::vector<uint8_t> &value) {
  PutStr(b, name);
  PutStr(b, type);
  PutI32(b, (int32_t)value.size());
  Put(b, value.data(), value.size());
}

std::vector<uint8_t> Exr(const float *rgb, int w, int h) {
  std::vector<uint8_t> b;
  const uint8_t magic[4] = {0x76, 0x2f, 0x31, 0x01};
  Put(b, magic, 4);
  PutI32(b, 2);  // version 2, single-part scanline
  {
    std::vector<uint8_t> ch;
    for (const char *c : {"B", "G", "R"}) {  // channels in alphabetical order
      PutStr(ch, c);
      PutI32(ch, 2);  // FLOAT
      const uint8_t linear_reserved[4] = {0, 0, 0, 0};
      Put(ch, linear_reserved, 4);
      PutI32(ch, 1);
      PutI32(ch, 1);
    }
    ch.push_back(0);
    Attr(b, "channels", "chlist", ch);
  }
  Attr(b, "compression", "compression", {0});
  std::vector<uint8_t> box;
  PutI32(box, 0); PutI32(box, 0); PutI32(box, w - 1); PutI32(box, h - 1);
  Attr(b, "dataWindow", "box2i", box);
  Attr(b, "displayWindow", "box2i", box);
  Attr(b, "lineOrder", "lineOrder", {0});
  std::vector<uint8_t> f;
  PutF32(f, 1.f);
  Attr(b, "pixelAspectRatio", "float", f);
  std::vector<uint8_t> c;
  PutF32(c, 0.f); PutF32(c, 0.f);
  Attr(b, "screenWindowCenter", "v2f", c);
  Attr(b, "screenWindowWidth", "float", f);
  b.push_back(0);  // end of header
  const size_t table = b.size();
  const int32_t row_bytes = w * 3 * 4;
  const uint64_t first = table + 8ull * h;
  for (int y = 0; y < h; ++y) PutU64(b, first + (uint64_t)y * (8 + row_bytes));
  for (int y = 0; y < h; ++y) {
    PutI32(b, y);
    PutI32(b, row_bytes);
    for (int c3 : {2, 1, 0})
      for (int x = 0; x < w; ++x) PutF32(b, rgb[3 * ((size_t)y * w + x) + c3]);
  }
  return b;
}

std::vector<uint8_t> Pfm(const float *rgb, int w, int h) {
  std::vector<uint8_t> b;
  char head[64];
  const int n = std::snprintf(head, sizeof(head), "PF\n%d %d\n-1.0\n", w, h);
  Put(b, head, (size_t)n);
  for (int y = h - 1; y >= 0; --y) Put(b, rgb + 3 * (size_t)y * w, (size_t)w * 3 * 4);
  return b;
}

}  // namespace

extern "C" int ttgpu_write_rgb_image(const float *rgb, int32_t width, int32_t height, const char *path, int32_t format) {
  if (!rgb || !path || width <= 0 || height <= 0) return TTGPU_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t> data;
  if (format == TTGPU_IMAGE_PFM) data = Pfm(rgb, width, height);
  else if (format == TTGPU_IMAGE_EXR) data = Exr(rgb, width, height);
  else return TTGPU_ERR_INVALID_ARGUMENT;
  FILE *f = std::fopen(path, "wb");
  if (!f) return TTGPU_ERR_RUNTIME;
  const size_t wrote = std::fwrite(data.data(), 1, data.size(), f);
  const int closed = std::fclose(f);
  return wrote == data.size() && closed == 0 ? TTGPU_OK : TTGPU_ERR_RUNTIME;
}
