// gst/ans.hpp -- C++ counterpart of the reference's ans::ocl::OpenCLDecoder
// (ans/ans_ocl.h:26-72) on top of the C ABI in gst_ans.h.  Header-only; link with
// libgst_amd.so.  Same constructor arguments and methods; the OpenCL types become
// plain ones (cl_uchar -> uint8_t, cl_ushort -> uint16_t, cl_uint -> uint32_t), and
// failures throw GenTC::DecodeError instead of asserting.
#ifndef GST_ANS_HPP_
#define GST_ANS_HPP_

#include <cstdint>
#include <memory>
#include <vector>

#include "gst/decoder.hpp"
#include "gst_ans.h"

namespace ans {
namespace hip {

static const size_t kANSTableSize = GST_ANS_TABLE_SIZE;            // ans/ans.h:73
static const size_t kNumEncodedSymbols = GST_ANS_SYMBOLS_PER_STREAM; // ans/ans.h:74

// ans/ans_ocl_encode.cpp:6-8 (GenerateHistogram(F, 2048), ans/histogram.cpp:40-121).
inline std::vector<uint32_t> NormalizeFrequencies(const std::vector<uint32_t> &F) {
  std::vector<uint32_t> out(F.size());
  const int rc = gst_ans_normalize_frequencies(F.data(), static_cast<uint32_t>(F.size()),
                                               static_cast<int32_t>(kANSTableSize), out.data());
  if (rc != GST_OK) throw GenTC::DecodeError("NormalizeFrequencies", rc);
  return out;
}

class HIPDecoder {
 public:
  // ans/ans_ocl.h:28-32.  The context only selects the device here: the decoder
  // keeps its own stream.
  HIPDecoder(const std::unique_ptr<gpu::GPUContext> &ctx, const std::vector<uint32_t> &F, int num_interleaved) {
    if (ctx) (void)hipSetDevice(ctx->Device());
    const int rc = gst_ans_decoder_create(F.data(), static_cast<uint32_t>(F.size()),
                                          static_cast<uint32_t>(num_interleaved), &d_);
    if (rc != GST_OK) throw GenTC::DecodeError("HIPDecoder", rc);
  }
  ~HIPDecoder() { (void)gst_ans_decoder_destroy(d_); }
  HIPDecoder(const HIPDecoder &) = delete;
  HIPDecoder &operator=(const HIPDecoder &) = delete;

  // ans/ans_ocl.h:35-37: one stream of 256 symbols.
  std::vector<uint8_t> Decode(uint32_t state, const std::vector<uint8_t> &data) const {
    std::vector<uint8_t> out(kNumEncodedSymbols);
    const uint8_t *p = data.data();
    const size_t len = data.size();
    Check(gst_ans_decode(d_, &state, 1, &p, &len, 1, out.data()), "Decode");
    return out;
  }

  // ans/ans_ocl.h:39-41: states.size() <= num_interleaved streams interleaved in one buffer.
  std::vector<std::vector<uint8_t>> Decode(const std::vector<uint32_t> &states,
                                           const std::vector<uint8_t> &data) const {
    std::vector<uint8_t> all(states.size() * kNumEncodedSymbols);
    const uint8_t *p = data.data();
    const size_t len = data.size();
    Check(gst_ans_decode(d_, states.data(), static_cast<uint32_t>(states.size()), &p, &len, 1, all.data()),
          "Decode");
    return Split(all, states.size());
  }

  // ans/ans_ocl.h:43-45: data.size() independent buffers of num_interleaved streams each.
  std::vector<std::vector<uint8_t>> Decode(const std::vector<uint32_t> &states,
                                           const std::vector<std::vector<uint8_t>> &data) const {
    std::vector<const uint8_t *> ptrs;
    std::vector<size_t> lens;
    for (const auto &v : data) {
      ptrs.push_back(v.data());
      lens.push_back(v.size());
    }
    std::vector<uint8_t> all(states.size() * kNumEncodedSymbols);
    Check(gst_ans_decode(d_, states.data(), static_cast<uint32_t>(states.size()), ptrs.data(), lens.data(),
                         static_cast<uint32_t>(data.size()), all.data()),
          "Decode");
    return Split(all, states.size());
  }

  // ans/ans_ocl.h:47.
  void RebuildTable(const std::vector<uint32_t> &F) {
    Check(gst_ans_rebuild_table(d_, F.data(), static_cast<uint32_t>(F.size())), "RebuildTable");
  }

  // ans/ans_ocl.h:49-51.
  std::vector<uint8_t> GetSymbols() const {
    std::vector<uint8_t> v(kANSTableSize);
    Check(gst_ans_get_table(d_, v.data(), nullptr, nullptr), "GetSymbols");
    return v;
  }
  std::vector<uint16_t> GetFrequencies() const {
    std::vector<uint16_t> v(kANSTableSize);
    Check(gst_ans_get_table(d_, nullptr, v.data(), nullptr), "GetFrequencies");
    return v;
  }
  std::vector<uint16_t> GetCumulativeFrequencies() const {
    std::vector<uint16_t> v(kANSTableSize);
    Check(gst_ans_get_table(d_, nullptr, nullptr, v.data()), "GetCumulativeFrequencies");
    return v;
  }

 private:
  static void Check(int rc, const char *what) {
    if (rc != GST_OK) throw GenTC::DecodeError(what, rc);
  }
  static std::vector<std::vector<uint8_t>> Split(const std::vector<uint8_t> &all, size_t n) {
    std::vector<std::vector<uint8_t>> out;
    out.reserve(n);
    for (size_t i = 0; i < n; ++i)
      out.emplace_back(all.begin() + i * kNumEncodedSymbols, all.begin() + (i + 1) * kNumEncodedSymbols);
    return out;
  }

  gst_ans_decoder *d_ = nullptr;
};

}  // namespace hip
}  // namespace ans

#endif  // GST_ANS_HPP_
