// gst/encoder.hpp -- C++ host API mirroring GammaUNC/GST codec/encoder.h (namespace GenTC)
// over the C ABI of gst_encode.h.  Header-only, CPU only; link with libgst_encode.so.
//
// CompressDXT(width, height, rgb, dxt) is the reference's overload of the same name
// (codec/encoder.cpp:182-186): the .gtc bytes gentenc writes, byte for byte.  An empty
// `dxt` compresses the blocks with stb_dxt v1.04 first, as DXTImage(width, height, rgb)
// does (codec/dxt_image.cpp:444-452).  Where gentenc stops on an assertion or takes a
// size it cannot code, this throws GenTC::EncodeError with the GSTE_* status.  The
// file-name overload (PNG / .crn loading) and the DXTImage overload are not provided:
// callers pass pixels and, optionally, blocks.
#ifndef GST_ENCODER_HPP_
#define GST_ENCODER_HPP_

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gst_encode.h"

namespace GenTC {

class EncodeError : public std::runtime_error {
 public:
  EncodeError(const std::string &where, int code)
      : std::runtime_error(where + " failed with status " + std::to_string(code)), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// The file and, in *blocks (may be null), the DXT1 blocks after gentenc's palette reuse:
// the reference's DXTImage::PhysicalBlocks(), what the file decodes to.
inline std::vector<uint8_t> CompressDXT(int width, int height, const std::vector<uint8_t> &rgb_data,
                                        const std::vector<uint8_t> &dxt_data, std::vector<uint8_t> *blocks) {
  if (width <= 0 || height <= 0 || rgb_data.size() != static_cast<size_t>(width) * height * 3 ||
      (!dxt_data.empty() && dxt_data.size() != static_cast<size_t>(width) * height / 2))
    throw EncodeError("CompressDXT", GSTE_ERR_ARGUMENT);
  const uint32_t w = static_cast<uint32_t>(width), h = static_cast<uint32_t>(height);
  std::vector<uint8_t> out(gste_max_encoded_size(w, h));
  std::vector<uint8_t> blk(static_cast<size_t>(w) * h / 2);
  size_t n = 0;
  const int rc = gste_compress_dxt(rgb_data.data(), w, h, dxt_data.empty() ? nullptr : dxt_data.data(), out.data(),
                                   out.size(), &n, blk.data());
  if (rc != GSTE_OK) throw EncodeError("CompressDXT", rc);
  out.resize(n);
  if (blocks) *blocks = std::move(blk);
  return out;
}

// codec/encoder.h:14-16.
inline std::vector<uint8_t> CompressDXT(int width, int height, const std::vector<uint8_t> &rgb_data,
                                        const std::vector<uint8_t> &dxt_data) {
  return CompressDXT(width, height, rgb_data, dxt_data, nullptr);
}

}  // namespace GenTC

#endif  // GST_ENCODER_HPP_
