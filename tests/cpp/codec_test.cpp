// codec_test.cpp -- the reference's end-to-end decoder test (codec/test/codec_test.cpp:36-48)
// restated against the C++ host API (include/gst/decoder.hpp).
//
// usage: codec_test <file.gtc> <expected.dxt>
// Decodes through DecompressDXT and through LoadCompressedDXT on a device staging
// buffer, and requires every 64-bit block to equal the expected PhysicalBlocks().
#include <cstdio>
#include <fstream>
#include <iterator>
#include <vector>

#include "gst/decoder.hpp"

static std::vector<uint8_t> ReadFile(const char *fn) {
  std::ifstream is(fn, std::ios::binary);
  return std::vector<uint8_t>((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
}

static int CompareBlocks(const GenTC::DXTImage &img, const std::vector<uint8_t> &expected, const char *what) {
  const auto &blks = img.PhysicalBlocks();
  const auto *exp = reinterpret_cast<const GenTC::PhysicalDXTBlock *>(expected.data());
  if (blks.size() * 8 != expected.size()) {
    printf("%s: %zu blocks, expected %zu\n", what, blks.size(), expected.size() / 8);
    return 1;
  }
  for (size_t i = 0; i < blks.size(); ++i) {
    if (blks[i].dxt_block != exp[i].dxt_block) {
      printf("%s: block %zu differs\n", what, i);
      return 1;
    }
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: codec_test <file.gtc> <expected.dxt>\n");
    return 2;
  }
  std::vector<uint8_t> cmp = ReadFile(argv[1]), expected = ReadFile(argv[2]);
  auto ctx = gpu::GPUContext::InitializeHIP(0);
  if (!ctx || !GenTC::InitializeDecoder(ctx)) {
    printf("InitializeDecoder failed\n");
    return 1;
  }
  try {
    GenTC::DXTImage img = GenTC::DecompressDXT(ctx, cmp);
    if (CompareBlocks(img, expected, "DecompressDXT")) return 1;

    // Device-resident path: staging buffer (offsets block + file minus header at 512).
    GenTC::GenTCHeader hdr;
    hdr.LoadFrom(cmp.data());
    size_t staging_sz = 0, streams_off = 0;
    GenTC::Check(gst_staging_size(hdr.c(), 1, &staging_sz, &streams_off), "gst_staging_size");
    std::vector<uint8_t> staging(staging_sz);
    const uint8_t *files[1] = {cmp.data()};
    const size_t lens[1] = {cmp.size()};
    GenTC::Check(gst_pack_staging(files, lens, 1, staging.data(), staging.size()), "gst_pack_staging");
    void *d_cmp = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_cmp, staging.size()) != hipSuccess || hipMalloc(&d_out, expected.size()) != hipSuccess)
      return 1;
    hipStream_t q = ctx->GetNextQueue();
    (void)hipMemcpyAsync(d_cmp, staging.data(), staging.size(), hipMemcpyHostToDevice, q);
    hipEvent_t upload;
    (void)hipEventCreate(&upload);
    (void)hipEventRecord(upload, q);
    hipStream_t q2 = ctx->GetNextQueue();
    hipEvent_t done = GenTC::LoadCompressedDXT(ctx, hdr, q2, d_cmp, d_out, 1, &upload);
    (void)hipEventSynchronize(done);
    std::vector<uint8_t> out(expected.size());
    (void)hipMemcpy(out.data(), d_out, out.size(), hipMemcpyDeviceToHost);
    (void)hipEventDestroy(done);
    (void)hipEventDestroy(upload);
    (void)hipFree(d_cmp);
    (void)hipFree(d_out);
    GenTC::DXTImage img2(static_cast<int>(hdr.width), static_cast<int>(hdr.height), out);
    if (CompareBlocks(img2, expected, "LoadCompressedDXT")) return 1;
  } catch (const GenTC::DecodeError &e) {
    printf("%s\n", e.what());
    return 1;
  }
  printf("codec_test: %s decoded bit-exact\n", argv[1]);
  return 0;
}
