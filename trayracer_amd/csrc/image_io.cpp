// Image files (host): portable float map and PNG. The PNG encoder is self-contained:
// zlib stream of stored (uncompressed) deflate blocks, CRC-32 per chunk (ISO 3309
// polynomial 0xEDB88320, table built at first use), Adler-32 of the raw data
// (RFC 1950/1951, PNG specification 2nd edition).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/trt.h"

namespace {

uint32_t crc32_update(uint32_t crc, const uint8_t* d, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  for (size_t i = 0; i < n; ++i) crc = table[(crc ^ d[i]) & 0xFFu] ^ (crc >> 8);
  return crc;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back(uint8_t(x >> 24));
  v.push_back(uint8_t(x >> 16));
  v.push_back(uint8_t(x >> 8));
  v.push_back(uint8_t(x));
}

void chunk(std::vector<uint8_t>& out, const char type[4], const std::vector<uint8_t>& data) {
  put_be32(out, uint32_t(data.size()));
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  const uint32_t crc = crc32_update(0xFFFFFFFFu, out.data() + start, out.size() - start) ^ 0xFFFFFFFFu;
  put_be32(out, crc);
}

bool write_all(const char* path, const void* data, size_t n) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = std::fwrite(data, 1, n, f) == n;
  return std::fclose(f) == 0 && ok;
}

}  // namespace

extern "C" {

trt_status trt_write_pfm(const char* path, uint32_t width, uint32_t height, const float* rgba) {
  if (!path || !rgba || width == 0 || height == 0) return TRT_INVALID_ARG;
  const std::string head = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1.0\n";
  std::vector<uint8_t> out(head.begin(), head.end());
  const size_t n = size_t(width) * height;
  out.reserve(out.size() + n * 12);
  for (size_t i = 0; i < n; ++i) {  // PFM rows run bottom to top, as stored; little endian
    uint8_t b[12];
    std::memcpy(b, rgba + i * 4, 12);
    out.insert(out.end(), b, b + 12);
  }
  return write_all(path, out.data(), out.size()) ? TRT_OK : TRT_INVALID_ARG;
}

trt_status trt_write_png(const char* path, uint32_t width, uint32_t height, const uint8_t* rgba8) {
  if (!path || !rgba8 || width == 0 || height == 0 || width > (1u << 24) || height > (1u << 24))
    return TRT_INVALID_ARG;
  // raw scanlines, top row first (filter type 0 per row)
  const size_t row = size_t(width) * 4;
  std::vector<uint8_t> raw;
  raw.reserve((row + 1) * height);
  for (uint32_t y = 0; y < height; ++y) {
    raw.push_back(0);
    const uint8_t* src = rgba8 + size_t(height - 1 - y) * row;
    raw.insert(raw.end(), src, src + row);
  }
  // zlib: CMF/FLG for deflate with a 32K window, stored blocks of at most 65535 bytes
  std::vector<uint8_t> z = {0x78, 0x01};
  size_t pos = 0;
  do {
    const size_t len = raw.size() - pos < 65535 ? raw.size() - pos : 65535;
    const bool last = pos + len == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back(uint8_t(len & 0xFF));
    z.push_back(uint8_t(len >> 8));
    z.push_back(uint8_t(~len & 0xFF));
    z.push_back(uint8_t((~len >> 8) & 0xFF));
    z.insert(z.end(), raw.begin() + long(pos), raw.begin() + long(pos + len));
    pos += len;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  put_be32(z, (b << 16) | a);
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, width);
  put_be32(ihdr, height);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, adaptive filters, no interlace
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
  return write_all(path, out.data(), out.size()) ? TRT_OK : TRT_INVALID_ARG;
}

}  // extern "C"
