// KTX 1.1 container of encoded surfaces (host code).
//
// One level follows ImageWriterKTX::WriteImage (IO/src/ImageWriterKTX.cpp:69-160)
// byte for byte for the formats it writes (BPTC, DXT1, DXT5): identifier,
// endianness word, glType 0, glTypeSize 1, glFormat 0, the compressed internal
// format and its base format, the size, one face, the "KTXorientation" =
// "S=r,T=d" key/value pair padded to 4 bytes with the key's first bytes, then
// imageSize and the data.  Beyond the reference: several mip levels (level i
// is max(1, w >> i) x max(1, h >> i) texels, stored as its grid of whole
// blocks, imageSize the grid's bytes) and ETC1 (GL_ETC1_RGB8_OES).  For one
// level whose sides are multiples of 4 imageSize equals the reference's
// width * height (* 1/2 for DXT1).
#include <stdint.h>
#include <string.h>

#include "../../include/fasthip.h"
#include "internal.h"

namespace fh {
namespace {

constexpr uint32_t kGlRgb = 0x1907, kGlRgba = 0x1908;
constexpr uint32_t kGlBptc = 0x8E8C;        // GL_COMPRESSED_RGBA_BPTC_UNORM
constexpr uint32_t kGlDxt1 = 0x83F0;        // GL_COMPRESSED_RGB_S3TC_DXT1_EXT
constexpr uint32_t kGlDxt5 = 0x83F3;        // GL_COMPRESSED_RGBA_S3TC_DXT5_EXT
constexpr uint32_t kGlEtc1 = 0x8D64;        // GL_ETC1_RGB8_OES

struct Writer {
  uint8_t *out;
  size_t cap, n = 0;
  void bytes(const void *p, size_t len) {
    if (out && n + len <= cap) memcpy(out + n, p, len);
    n += len;
  }
  void u32(uint32_t v) { bytes(&v, 4); }  // little-endian host, as the reference writes
};

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_ktx_write(uint32_t format, uint32_t width, uint32_t height,
                                  uint32_t levels, const uint8_t *const *level_data, uint8_t *out,
                                  size_t out_capacity, size_t *out_size) {
  if (!out_size || width == 0 || height == 0 || levels == 0)
    return set_error("null size, empty image or no level"), FH_ERR_INVALID_ARGUMENT;
  uint32_t internal, base, block_bytes;
  switch (format) {
    case FH_FORMAT_BPTC: internal = kGlBptc; base = kGlRgba; block_bytes = 16; break;
    case FH_FORMAT_DXT1: internal = kGlDxt1; base = kGlRgb; block_bytes = 8; break;
    case FH_FORMAT_DXT5: internal = kGlDxt5; base = kGlRgba; block_bytes = 16; break;
    case FH_FORMAT_ETC1: internal = kGlEtc1; base = kGlRgb; block_bytes = 8; break;
    default: return set_error("Unknown compression format"), FH_ERR_UNKNOWN_FORMAT;
  }
  uint32_t max_levels = 1;
  for (uint32_t w = width, h = height; w > 1 || h > 1; max_levels++) {
    w = w > 1 ? w / 2 : 1;
    h = h > 1 ? h / 2 : 1;
  }
  if (levels > max_levels) return set_error("more levels than the mip chain has"), FH_ERR_INVALID_ARGUMENT;
  if (out && !level_data) return set_error("null level data"), FH_ERR_INVALID_ARGUMENT;
  Writer wr{out, out_capacity};
  static const uint8_t kIdentifier[12] = {0xAB, 0x4B, 0x54, 0x58, 0x20, 0x31,
                                          0x31, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A};
  wr.bytes(kIdentifier, 12);
  wr.u32(0x04030201);
  static const char kKey[] = "KTXorientation", kValue[] = "S=r,T=d";
  const uint32_t kv = (uint32_t)(sizeof(kKey) + sizeof(kValue));  // both with their NUL
  const uint32_t total_kv = (kv + 4 + 3) & ~3u;
  wr.u32(0);         // glType
  wr.u32(1);         // glTypeSize
  wr.u32(0);         // glFormat
  wr.u32(internal);  // glInternalFormat
  wr.u32(base);      // glBaseInternalFormat
  wr.u32(width);
  wr.u32(height);
  wr.u32(0);         // pixelDepth
  wr.u32(0);         // numberOfArrayElements
  wr.u32(1);         // numberOfFaces
  wr.u32(levels);    // numberOfMipmapLevels
  wr.u32(total_kv);  // bytesOfKeyValueData
  wr.u32(kv);        // keyAndValueByteSize
  wr.bytes(kKey, sizeof(kKey));
  wr.bytes(kValue, sizeof(kValue));
  wr.bytes(kKey, total_kv - kv - 4);  // padding, taken from the key as the reference does
  uint32_t w = width, h = height;
  for (uint32_t l = 0; l < levels; l++) {
    const uint64_t size = (uint64_t)((w + 3) / 4) * ((h + 3) / 4) * block_bytes;
    if (size > 0xFFFFFFFFull) return set_error("level too large for KTX"), FH_ERR_INVALID_ARGUMENT;
    wr.u32((uint32_t)size);  // imageSize (a multiple of 8: no mip padding follows)
    if (out) {
      if (!level_data[l]) return set_error("null level data"), FH_ERR_INVALID_ARGUMENT;
      wr.bytes(level_data[l], (size_t)size);
    } else {
      wr.n += (size_t)size;
    }
    w = w > 1 ? w / 2 : 1;
    h = h > 1 ? h / 2 : 1;
  }
  *out_size = wr.n;
  if (out && wr.n > out_capacity)
    return set_error("KTX output buffer too small"), FH_ERR_BUFFER_TOO_SMALL;
  return FH_OK;
}
