// gst/decoder.hpp -- C++ host API mirroring GammaUNC/GST codec/decoder.h (namespace GenTC)
// on top of the C ABI in gst_decoder.h.  Header-only; link with libgst_amd.so.
//
// The reference's OpenCL handles become HIP ones: cl_command_queue -> hipStream_t,
// cl_mem -> device pointer, cl_event -> hipEvent_t (returned events are owned by the
// caller, as in the reference, demo/photos_sf.cpp:872-874).  gpu::GPUContext keeps the
// reference's queue round robin (gpu/gpu.h:56-59) over HIP streams.
#ifndef GST_DECODER_HPP_
#define GST_DECODER_HPP_

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gst_decoder.h"

namespace gpu {

// Minimal stand-in for gpu::GPUContext (gpu/gpu.h:50-174): one device and a few
// created streams handed out round robin.
class GPUContext {
 public:
  static std::unique_ptr<GPUContext> InitializeHIP(int device = 0, int num_queues = 4) {
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    std::unique_ptr<GPUContext> ctx(new GPUContext(device));
    for (int i = 0; i < num_queues; ++i) {
      hipStream_t s;
      if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
      ctx->queues_.push_back(s);
    }
    return ctx;
  }
  ~GPUContext() {
    for (hipStream_t s : queues_) (void)hipStreamDestroy(s);
  }
  int Device() const { return device_; }
  hipStream_t GetDefaultCommandQueue() const { return queues_[0]; }
  hipStream_t GetNextQueue() { return queues_[next_.fetch_add(1) % queues_.size()]; }

 private:
  explicit GPUContext(int device) : device_(device), next_(0) {}
  int device_;
  std::vector<hipStream_t> queues_;
  std::atomic<size_t> next_;
};

}  // namespace gpu

namespace GenTC {

class DecodeError : public std::runtime_error {
 public:
  DecodeError(const std::string &where, int code)
      : std::runtime_error(where + " failed with status " + std::to_string(code)), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// codec/codec_base.h:9-20
struct GenTCHeader {
  uint32_t width;
  uint32_t height;
  uint32_t palette_bytes;
  uint32_t y_cmp_sz;
  uint32_t chroma_cmp_sz;
  uint32_t palette_sz;
  uint32_t indices_sz;

  void LoadFrom(const uint8_t *buf) { memcpy(this, buf, sizeof(*this)); }
  const gst_header *c() const { return reinterpret_cast<const gst_header *>(this); }
};
static_assert(sizeof(GenTCHeader) == sizeof(gst_header), "header layout");

static const size_t kWaveletBlockDim = 32;

// codec/dxt_image.h:14-21
union PhysicalDXTBlock {
  struct {
    uint16_t ep1;
    uint16_t ep2;
    uint32_t interpolation;
  };
  uint64_t dxt_block;
};

// The part of DXTImage (codec/dxt_image.h:39-119) the decoder returns:
// DXTImage(w, h, dxt) and PhysicalBlocks().
class DXTImage {
 public:
  DXTImage(int width, int height, const std::vector<uint8_t> &dxt_data)
      : width_(width), height_(height),
        blocks_(reinterpret_cast<const PhysicalDXTBlock *>(dxt_data.data()),
                reinterpret_cast<const PhysicalDXTBlock *>(dxt_data.data()) +
                    static_cast<size_t>(width / 4) * (height / 4)) {}
  int Width() const { return width_; }
  int Height() const { return height_; }
  int BlocksWide() const { return width_ / 4; }
  int BlocksHigh() const { return height_ / 4; }
  const std::vector<PhysicalDXTBlock> &PhysicalBlocks() const { return blocks_; }

 private:
  int width_, height_;
  std::vector<PhysicalDXTBlock> blocks_;
};

inline void Check(int rc, const char *where) {
  if (rc != GST_OK) throw DecodeError(where, rc);
}

// codec/decoder.h:15
inline bool InitializeDecoder(const std::unique_ptr<gpu::GPUContext> &ctx) {
  return ctx && gst_initialize_decoder(ctx->Device()) == GST_OK;
}

// codec/decoder.h:16-17
inline DXTImage DecompressDXT(const std::unique_ptr<gpu::GPUContext> &ctx,
                              const std::vector<uint8_t> &cmp_data) {
  (void)ctx;
  GenTCHeader hdr;
  Check(gst_parse_header(cmp_data.data(), cmp_data.size(), reinterpret_cast<gst_header *>(&hdr)),
        "DecompressDXT");
  std::vector<uint8_t> out(static_cast<size_t>(hdr.width) * hdr.height / 2);
  Check(gst_decompress_dxt(cmp_data.data(), cmp_data.size(), out.data(), out.size()), "DecompressDXT");
  return DXTImage(static_cast<int>(hdr.width), static_cast<int>(hdr.height), out);
}

// codec/decoder.h:19-21
inline hipEvent_t LoadCompressedDXT(const std::unique_ptr<gpu::GPUContext> &, const GenTCHeader &hdr,
                                    hipStream_t queue, const void *cmp_data, void *output,
                                    uint32_t num_init, const hipEvent_t *init) {
  gst_event ev = nullptr;
  Check(gst_load_compressed_dxt(hdr.c(), queue, cmp_data, output, num_init,
                                reinterpret_cast<const gst_event *>(init), &ev),
        "LoadCompressedDXT");
  return static_cast<hipEvent_t>(ev);
}

// codec/decoder.h:23-25 (images may differ in size here)
inline hipEvent_t LoadCompressedDXTs(const std::unique_ptr<gpu::GPUContext> &,
                                     const std::vector<GenTCHeader> &hdrs, hipStream_t queue,
                                     const void *cmp_data, void *output, uint32_t num_init,
                                     const hipEvent_t *init) {
  gst_event ev = nullptr;
  Check(gst_load_compressed_dxts(reinterpret_cast<const gst_header *>(hdrs.data()),
                                 static_cast<uint32_t>(hdrs.size()), queue, cmp_data, output, num_init,
                                 reinterpret_cast<const gst_event *>(init), &ev),
        "LoadCompressedDXTs");
  return static_cast<hipEvent_t>(ev);
}

// codec/decoder.h:27-33
inline hipEvent_t LoadRGB(const std::unique_ptr<gpu::GPUContext> &, const GenTCHeader &hdr,
                          hipStream_t queue, const void *cmp_data, void *output, uint32_t num_init,
                          const hipEvent_t *init) {
  gst_event ev = nullptr;
  Check(gst_load_rgb(hdr.c(), queue, cmp_data, output, num_init, reinterpret_cast<const gst_event *>(init),
                     &ev),
        "LoadRGB");
  return static_cast<hipEvent_t>(ev);
}

inline hipEvent_t LoadRGBs(const std::unique_ptr<gpu::GPUContext> &, const std::vector<GenTCHeader> &hdrs,
                           hipStream_t queue, const void *cmp_data, void *output, uint32_t num_init,
                           const hipEvent_t *init) {
  gst_event ev = nullptr;
  Check(gst_load_rgbs(reinterpret_cast<const gst_header *>(hdrs.data()), static_cast<uint32_t>(hdrs.size()),
                      queue, cmp_data, output, num_init, reinterpret_cast<const gst_event *>(init), &ev),
        "LoadRGBs");
  return static_cast<hipEvent_t>(ev);
}

// codec/decoder.h:35-37
inline size_t RequiredScratchMem(const GenTCHeader &hdr) { return gst_required_scratch_mem(hdr.c()); }

inline void PreallocateDecompressor(const std::unique_ptr<gpu::GPUContext> &, size_t req_sz) {
  Check(gst_preallocate_decompressor(req_sz), "PreallocateDecompressor");
}

inline void FreeDecompressor() { Check(gst_free_decompressor(), "FreeDecompressor"); }

}  // namespace GenTC

#endif  // GST_DECODER_HPP_
