// Stand-alone interleaved rANS decoder (C ABI of include/gst_ans.h): the counterpart
// of the reference's ans::ocl::OpenCLDecoder (ans/ans_ocl.h:26-72, ans/ans_ocl.cpp).
// The table is built by the texture path's BuildTableKernel (one of its four slots)
// and the groups are decoded by AnsGroupsKernel; this file holds the host side:
// frequency normalisation, the data packing and the blocking copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "gst_ans.h"
#include "gst_histogram.h"
#include "gst_internal.h"

struct gst_ans_decoder {
  uint32_t num_interleaved = 1;
  uint8_t *d_freqs = nullptr;   // 4 x 256 u16: the build stage makes four tables, all alike
  uint32_t *d_tables = nullptr; // 4 x 2048 packed entries + 4 status words
  uint32_t *d_err = nullptr;
  hipStream_t stream = nullptr;
  mutable std::mutex mu;        // one decode or rebuild at a time per decoder
};

namespace {

int HipFail(hipError_t e) { return e == hipErrorOutOfMemory ? GST_ERR_OUT_OF_MEMORY : GST_ERR_HIP; }

#define ANS_HIP(call)                         \
  do {                                        \
    hipError_t _e = (call);                   \
    if (_e != hipSuccess) return HipFail(_e); \
  } while (0)

using gst::hist::GenerateHistogram;

int BuildTable(gst_ans_decoder *d, const uint32_t *F, uint32_t n) {
  if (!F || n == 0 || n > 256) return GST_ERR_INVALID_ARGUMENT;
  std::vector<uint32_t> norm;
  if (!GenerateHistogram(std::vector<uint32_t>(F, F + n), static_cast<int>(GST_ANS_TABLE_SIZE), &norm))
    return GST_ERR_INVALID_ARGUMENT;
  uint16_t freqs[4][256] = {};
  for (uint32_t k = 0; k < 4; ++k)
    for (uint32_t i = 0; i < n; ++i) freqs[k][i] = static_cast<uint16_t>(norm[i]);
  ANS_HIP(hipMemcpyAsync(d->d_freqs, freqs, sizeof(freqs), hipMemcpyHostToDevice, d->stream));
  ANS_HIP(hipMemsetAsync(d->d_err, 0, sizeof(uint32_t), d->stream));
  int rc = gst_stage_build_table(d->d_freqs, d->d_tables, d->d_err, d->stream);
  if (rc != GST_OK) return rc;
  uint32_t err = 0;
  ANS_HIP(hipMemcpyAsync(&err, d->d_err, sizeof(err), hipMemcpyDeviceToHost, d->stream));
  ANS_HIP(hipStreamSynchronize(d->stream));
  return err ? GST_ERR_INVALID_ARGUMENT : GST_OK;
}

size_t RoundUp4(size_t x) { return (x + 3) & ~static_cast<size_t>(3); }

}  // namespace

extern "C" {

int gst_ans_normalize_frequencies(const uint32_t *counts, uint32_t n, int32_t M, uint32_t *out) {
  if (!counts || !out || n == 0) return GST_ERR_INVALID_ARGUMENT;
  std::vector<uint32_t> h;
  if (!GenerateHistogram(std::vector<uint32_t>(counts, counts + n), M, &h)) return GST_ERR_INVALID_ARGUMENT;
  std::memcpy(out, h.data(), n * sizeof(uint32_t));
  return GST_OK;
}

int gst_ans_decoder_create(const uint32_t *F, uint32_t n, uint32_t num_interleaved, gst_ans_decoder **out) {
  if (!out || num_interleaved == 0 || num_interleaved > GST_ANS_MAX_INTERLEAVED) return GST_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int rc = gst::EnsureDecoderInit();
  if (rc != GST_OK) return rc;
  gst_ans_decoder *d = new gst_ans_decoder;
  d->num_interleaved = num_interleaved;
  hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&d->d_freqs, 4 * 256 * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc(&d->d_tables, (4 * GST_ANS_TABLE_SIZE + 4) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&d->d_err, sizeof(uint32_t));
  rc = e == hipSuccess ? BuildTable(d, F, n) : HipFail(e);
  if (rc != GST_OK) {
    gst_ans_decoder_destroy(d);
    return rc;
  }
  *out = d;
  return GST_OK;
}

int gst_ans_rebuild_table(gst_ans_decoder *d, const uint32_t *F, uint32_t n) {
  if (!d) return GST_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(d->mu);
  return BuildTable(d, F, n);
}

int gst_ans_get_table(const gst_ans_decoder *d, uint8_t *symbols, uint16_t *frequencies,
                      uint16_t *cumulative_frequencies) {
  if (!d) return GST_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(d->mu);
  std::vector<uint32_t> tab(GST_ANS_TABLE_SIZE);
  ANS_HIP(hipMemcpyAsync(tab.data(), d->d_tables, GST_ANS_TABLE_SIZE * sizeof(uint32_t), hipMemcpyDeviceToHost,
                         d->stream));
  ANS_HIP(hipStreamSynchronize(d->stream));
  // Packed entry: bits 0-11 f, bits 12-22 slot - cum, bits 24-31 symbol.
  for (uint32_t s = 0; s < GST_ANS_TABLE_SIZE; ++s) {
    const uint32_t e = tab[s];
    if (symbols) symbols[s] = static_cast<uint8_t>(e >> 24);
    if (frequencies) frequencies[s] = static_cast<uint16_t>(e & 0xFFFu);
    if (cumulative_frequencies) cumulative_frequencies[s] = static_cast<uint16_t>(s - ((e >> 12) & 0x7FFu));
  }
  return GST_OK;
}

int gst_ans_decode(const gst_ans_decoder *d, const uint32_t *states, uint32_t num_streams,
                   const uint8_t *const *data, const size_t *data_lens, uint32_t n_data, uint8_t *out) {
  if (!d || !states || !data || !data_lens || !out || n_data == 0 || num_streams == 0 ||
      num_streams % n_data != 0)
    return GST_ERR_INVALID_ARGUMENT;
  const uint32_t lanes = num_streams / n_data;
  if (lanes > GST_ANS_MAX_INTERLEAVED || (n_data > 1 && lanes != d->num_interleaved) ||
      (n_data == 1 && lanes > d->num_interleaved && lanes > 1))
    return GST_ERR_INVALID_ARGUMENT;
  // ans/ans_ocl.cpp:330-356: u32 end offsets, then per data stream its words right-
  // aligned to a 4-byte boundary (zero bytes before them) and its states.
  size_t total = 4 * static_cast<size_t>(n_data);
  for (uint32_t i = 0; i < n_data; ++i) {
    if (!data[i] && data_lens[i]) return GST_ERR_INVALID_ARGUMENT;
    if (data_lens[i] % 2) return GST_ERR_INVALID_ARGUMENT;  // whole 16-bit words
    total += RoundUp4(data_lens[i]) + 4 * static_cast<size_t>(lanes);
  }
  if (total > 0xFFFFFFF0u) return GST_ERR_INVALID_ARGUMENT;
  std::vector<uint8_t> packed(total, 0);
  size_t at = 4 * static_cast<size_t>(n_data);
  for (uint32_t i = 0; i < n_data; ++i) {
    const size_t end = at + RoundUp4(data_lens[i]) + 4 * static_cast<size_t>(lanes);
    if (data_lens[i]) std::memcpy(&packed[end - 4 * lanes - data_lens[i]], data[i], data_lens[i]);
    std::memcpy(&packed[end - 4 * lanes], states + static_cast<size_t>(i) * lanes, 4 * lanes);
    const uint32_t end32 = static_cast<uint32_t>(end);
    std::memcpy(&packed[4 * i], &end32, 4);
    at = end;
  }
  const size_t out_bytes = static_cast<size_t>(num_streams) * GST_ANS_SYMBOLS_PER_STREAM;
  std::lock_guard<std::mutex> lock(d->mu);
  uint8_t *d_data = nullptr, *d_out = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&d_data), total, d->stream);
  if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void **>(&d_out), out_bytes, d->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_data, packed.data(), total, hipMemcpyHostToDevice, d->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_out, 0, out_bytes, d->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d->d_err, 0, sizeof(uint32_t), d->stream);
  if (e == hipSuccess)
    e = gst::LaunchAnsGroups(d->d_tables, d_data, static_cast<uint32_t>(total), n_data, lanes, d_out, d->d_err,
                             d->stream);
  uint32_t err = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, d->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&err, d->d_err, sizeof(err), hipMemcpyDeviceToHost, d->stream);
  if (d_data) (void)hipFreeAsync(d_data, d->stream);
  if (d_out) (void)hipFreeAsync(d_out, d->stream);
  const hipError_t s = hipStreamSynchronize(d->stream);
  if (e == hipSuccess) e = s;
  if (e != hipSuccess) return HipFail(e);
  return err ? GST_ERR_CORRUPT_STREAM : GST_OK;
}

int gst_ans_decoder_destroy(gst_ans_decoder *d) {
  if (!d) return GST_ERR_INVALID_ARGUMENT;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  if (d->d_freqs) (void)hipFree(d->d_freqs);
  if (d->d_tables) (void)hipFree(d->d_tables);
  if (d->d_err) (void)hipFree(d->d_err);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return GST_OK;
}

}  // extern "C"
