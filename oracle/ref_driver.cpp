// ref_driver.cpp -- command-line driver over the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against the
// reference sources where they lie (encoder, CPU rANS coder, CPU wavelet); the
// binary goes to oracle/_ref/ and is used to write the golden vectors under
// tests/golden/ and to cross-check the oracle.  It contains no reference code:
// it only calls the reference's public functions.
//
// Commands (all binary files are raw little-endian):
//   encode_rgb <rgb8.raw> <w> <h> <out.gtc> <out.dxt>
//       GenTC::DXTImage(w, h, rgb) (stb_dxt + Reencode, codec/dxt_image.cpp:444-452)
//       then GenTC::CompressDXT (codec/encoder.cpp:193-195).  out.dxt holds
//       PhysicalBlocks(), the expected decoder output (codec/test/codec_test.cpp:44-47).
//   encode_png <image.png> <out.gtc> <out.dxt>
//       Same, image loaded by DXTImage(orig_fn, NULL) (codec/dxt_image.cpp:387-442).
//   encode_rgb_dxt <rgb8.raw> <w> <h> <in.dxt> <out.gtc> <out.dxt>
//       GenTC::DXTImage(w, h, rgb, dxt) (given blocks + Reencode,
//       codec/dxt_image.cpp:454-468) then GenTC::CompressDXT.
//   png_rgb <image.png> <out.rgb>
//       The RGB8 pixels stbi_load(fn, .., 3) gives DXTImage(orig_fn, NULL).
//   decompress_rgb <in.dxt> <w> <h> <out.rgb>
//       GenTC::DXTImage(w, h, dxt) (codec/dxt_image.cpp:470-480, no re-encoding), then
//       DecompressedImage() (codec/dxt_image.cpp:665-687) with the alpha dropped.
//   stages <in.gtc> <out.sym> <out.planes>
//       Per-group ans::DecodeInterleaved (ans/decode.cpp:240-292) of every stream
//       into the reference symbol layout, then InverseWavelet2D
//       (codec/wavelet.cpp:133-155) over dims 2,4,8,16,32 of every 32x32 tile
//       into six int8 raster planes.
//   hist <M> <count>...
//       ans::GenerateHistogram (ans/histogram.cpp:40-121); prints the result.
//   ans_group <seed> <out.bin> <count>...
//       8192 symbols drawn from the counts by a fixed LCG, normalised with
//       ans::ocl::NormalizeFrequencies and encoded with ans::EncodeInterleaved
//       (ans/encode.cpp:224-259) over 32 streams.  out.bin =
//       [u16 freq[256]][u32 group_bytes][group bytes][8192 symbols].
#include <cassert>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "ans.h"
#include "codec_base.h"
#include "dxt_image.h"
#include "encoder.h"
#include "histogram.h"
#include "wavelet.h"

// Declarations only; the implementation is compiled into codec/dxt_image.cpp.
#include "stb_image.h"

static std::vector<uint8_t> ReadFile(const char *fn) {
  std::ifstream is(fn, std::ios::binary);
  return std::vector<uint8_t>((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
}

static void WriteFile(const char *fn, const void *data, size_t n) {
  FILE *f = fopen(fn, "wb");
  if (!f || fwrite(data, 1, n, f) != n) {
    fprintf(stderr, "write failed: %s\n", fn);
    exit(2);
  }
  fclose(f);
}

// Encoder output is chatty on stdout; keep our own results on stderr-free stdout
// by silencing std::cout during reference calls.
struct QuietCout {
  std::streambuf *old;
  std::ostringstream sink;
  QuietCout() : old(std::cout.rdbuf(sink.rdbuf())) {}
  ~QuietCout() { std::cout.rdbuf(old); }
};

static void WriteEncoded(const GenTC::DXTImage &img, const char *gtc, const char *dxt) {
  std::vector<uint8_t> cmp;
  {
    QuietCout q;
    cmp = GenTC::CompressDXT(img);
  }
  WriteFile(gtc, cmp.data(), cmp.size());
  const auto &blks = img.PhysicalBlocks();
  WriteFile(dxt, blks.data(), blks.size() * sizeof(blks[0]));
}

static int CmdEncodeRGB(int argc, char **argv) {
  if (argc != 7) return 1;
  std::vector<uint8_t> rgb = ReadFile(argv[2]);
  const int w = atoi(argv[3]), h = atoi(argv[4]);
  if (rgb.size() != static_cast<size_t>(w) * h * 3) {
    fprintf(stderr, "rgb size mismatch\n");
    return 2;
  }
  QuietCout *q = new QuietCout;
  GenTC::DXTImage img(w, h, rgb.data());
  delete q;
  WriteEncoded(img, argv[5], argv[6]);
  return 0;
}

static int CmdEncodePNG(int argc, char **argv) {
  if (argc != 5) return 1;
  QuietCout *q = new QuietCout;
  GenTC::DXTImage img(argv[2], NULL);
  delete q;
  WriteEncoded(img, argv[3], argv[4]);
  return 0;
}

static int CmdEncodeRGBDXT(int argc, char **argv) {
  if (argc != 8) return 1;
  std::vector<uint8_t> rgb = ReadFile(argv[2]);
  const int w = atoi(argv[3]), h = atoi(argv[4]);
  std::vector<uint8_t> dxt = ReadFile(argv[5]);
  if (rgb.size() != static_cast<size_t>(w) * h * 3 || dxt.size() != static_cast<size_t>(w) * h / 2) {
    fprintf(stderr, "rgb or dxt size mismatch\n");
    return 2;
  }
  QuietCout *q = new QuietCout;
  GenTC::DXTImage img(w, h, rgb, dxt);
  delete q;
  WriteEncoded(img, argv[6], argv[7]);
  return 0;
}

static int CmdPngRGB(int argc, char **argv) {
  if (argc != 4) return 1;
  int w = 0, h = 0;
  stbi_uc *data = stbi_load(argv[2], &w, &h, NULL, 3);
  if (!data) return 2;
  WriteFile(argv[3], data, static_cast<size_t>(w) * h * 3);
  stbi_image_free(data);
  printf("%d %d\n", w, h);
  return 0;
}

static int CmdDecompressRGB(int argc, char **argv) {
  if (argc != 6) return 1;
  std::vector<uint8_t> dxt = ReadFile(argv[2]);
  const int w = atoi(argv[3]), h = atoi(argv[4]);
  if (dxt.size() != static_cast<size_t>(w) * h / 2) {
    fprintf(stderr, "dxt size mismatch\n");
    return 2;
  }
  GenTC::DXTImage img(w, h, dxt);
  std::unique_ptr<GenTC::RGBAImage> rgba = img.DecompressedImage();
  std::vector<uint8_t> rgb;
  rgb.reserve(static_cast<size_t>(w) * h * 3);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const GenTC::RGBA p = rgba->GetAt(x, y);
      rgb.push_back(std::get<0>(p));
      rgb.push_back(std::get<1>(p));
      rgb.push_back(std::get<2>(p));
    }
  WriteFile(argv[5], rgb.data(), rgb.size());
  return 0;
}

static uint32_t Rd32(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

static int CmdStages(int argc, char **argv) {
  if (argc != 5) return 1;
  std::vector<uint8_t> f = ReadFile(argv[2]);
  GenTC::GenTCHeader hdr;
  memcpy(&hdr, f.data(), sizeof(hdr));
  const size_t nv = static_cast<size_t>(hdr.width) * hdr.height / 16;
  const size_t P = hdr.palette_bytes;
  const size_t syms[4] = {2 * nv, 4 * nv, P, nv};
  const size_t sizes[4] = {hdr.y_cmp_sz, hdr.chroma_cmp_sz, hdr.palette_sz, hdr.indices_sz};
  std::vector<uint8_t> symbols(7 * nv + P, 0);
  size_t in_off = 28 + 2048, out_off = 0;
  for (int k = 0; k < 4; ++k) {
    std::vector<uint32_t> F(256);
    for (int i = 0; i < 256; ++i) F[i] = f[28 + 512 * k + 2 * i] | (f[28 + 512 * k + 2 * i + 1] << 8);
    while (!F.empty() && F.back() == 0) F.pop_back();
    const ans::Options opts = ans::ocl::GetOpenCLOptions(F);
    const uint8_t *stream = f.data() + in_off;
    const size_t G = syms[k] / 8192;
    for (size_t g = 0; g < G; ++g) {
      const uint32_t lo = g == 0 ? static_cast<uint32_t>(4 * G) : Rd32(stream + 4 * (g - 1));
      const uint32_t hi = Rd32(stream + 4 * g);
      std::vector<uint8_t> group(stream + lo, stream + hi);
      std::vector<uint8_t> dec = ans::DecodeInterleaved(group, 8192, opts, 32);
      memcpy(symbols.data() + out_off + g * 8192, dec.data(), 8192);
    }
    in_off += sizes[k];
    out_off += syms[k];
  }
  WriteFile(argv[3], symbols.data(), symbols.size());

  const size_t bx = hdr.width / 4, tiles_x = bx / 32, tiles = nv / 1024;
  std::vector<int8_t> planes(6 * nv);
  for (size_t p = 0; p < 6; ++p) {
    for (size_t t = 0; t < tiles; ++t) {
      int16_t blk[32 * 32];
      const uint8_t *in = symbols.data() + p * nv + t * 1024;
      for (int i = 0; i < 1024; ++i) blk[i] = static_cast<int16_t>(static_cast<int>(in[i]) - 128);
      for (size_t dim = 2; dim <= 32; dim *= 2)
        GenTC::InverseWavelet2D(blk, 64, blk, 64, dim);
      const size_t tx = t % tiles_x, ty = t / tiles_x;
      for (int y = 0; y < 32; ++y)
        for (int x = 0; x < 32; ++x)
          planes[p * nv + (ty * 32 + y) * bx + tx * 32 + x] = static_cast<int8_t>(blk[y * 32 + x]);
    }
  }
  WriteFile(argv[4], planes.data(), planes.size());
  return 0;
}

static int CmdHist(int argc, char **argv) {
  if (argc < 4) return 1;
  const int M = atoi(argv[2]);
  std::vector<uint32_t> counts;
  for (int i = 3; i < argc; ++i) counts.push_back(static_cast<uint32_t>(strtoul(argv[i], NULL, 10)));
  std::vector<uint32_t> h = ans::GenerateHistogram(counts, M);
  for (size_t i = 0; i < h.size(); ++i) printf("%s%u", i ? " " : "", h[i]);
  printf("\n");
  return 0;
}

static int CmdAnsGroup(int argc, char **argv) {
  if (argc < 5) return 1;
  uint64_t s = strtoull(argv[2], NULL, 10) * 6364136223846793005ull + 1442695040888963407ull;
  std::vector<uint32_t> counts;
  for (int i = 4; i < argc; ++i) counts.push_back(static_cast<uint32_t>(strtoul(argv[i], NULL, 10)));
  uint64_t total = 0;
  for (auto c : counts) total += c;
  std::vector<uint8_t> symbols(8192);
  for (auto &sym : symbols) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    uint64_t r = (s >> 33) % total;
    uint8_t k = 0;
    while (r >= counts[k]) r -= counts[k++];
    sym = k;
  }
  std::vector<uint32_t> F = ans::ocl::NormalizeFrequencies(counts);
  std::vector<uint8_t> enc = ans::EncodeInterleaved(symbols, ans::ocl::GetOpenCLOptions(F), 32);
  std::vector<uint8_t> out(512, 0);
  for (size_t i = 0; i < F.size(); ++i) {
    out[2 * i] = static_cast<uint8_t>(F[i]);
    out[2 * i + 1] = static_cast<uint8_t>(F[i] >> 8);
  }
  const uint32_t n = static_cast<uint32_t>(enc.size());
  out.insert(out.end(), reinterpret_cast<const uint8_t *>(&n), reinterpret_cast<const uint8_t *>(&n) + 4);
  out.insert(out.end(), enc.begin(), enc.end());
  out.insert(out.end(), symbols.begin(), symbols.end());
  WriteFile(argv[3], out.data(), out.size());
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: ref_driver encode_rgb|encode_png|encode_rgb_dxt|png_rgb|decompress_rgb|stages|hist|ans_group ...\n");
    return 1;
  }
  const std::string cmd = argv[1];
  int rc = 1;
  if (cmd == "encode_rgb") rc = CmdEncodeRGB(argc, argv);
  else if (cmd == "encode_png") rc = CmdEncodePNG(argc, argv);
  else if (cmd == "encode_rgb_dxt") rc = CmdEncodeRGBDXT(argc, argv);
  else if (cmd == "png_rgb") rc = CmdPngRGB(argc, argv);
  else if (cmd == "decompress_rgb") rc = CmdDecompressRGB(argc, argv);
  else if (cmd == "stages") rc = CmdStages(argc, argv);
  else if (cmd == "hist") rc = CmdHist(argc, argv);
  else if (cmd == "ans_group") rc = CmdAnsGroup(argc, argv);
  if (rc == 1) fprintf(stderr, "bad arguments for %s\n", cmd.c_str());
  return rc;
}
