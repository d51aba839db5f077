// encoder_test.cpp -- GenTC::CompressDXT through the C++ API (include/gst/encoder.hpp).
//
// Encodes the 512x512 seed-0 synthetic texture (noise 0.3) from its pixels, as the
// reference's DXTImage(w, h, rgb) + CompressDXT do, and writes the file and the blocks;
// the caller compares them with the reference's.  Also checks that the given-blocks
// overload takes the blocks (a second pass of palette reuse over them still writes a
// file that decodes to its own blocks, which the caller checks too) and that an
// unsupported size throws.
//
// usage: encoder_test <out.gtc> <out.dxt> <out2.gtc> <out2.dxt>
#include <cstdio>
#include <vector>

#include "gst/encoder.hpp"

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: encoder_test <out.gtc> <out.dxt> <out2.gtc> <out2.dxt>\n");
    return 2;
  }
  const int w = 512, h = 512;
  std::vector<uint8_t> rgb(static_cast<size_t>(w) * h * 3);
  gste_generate_rgb(0, w, h, 300, rgb.data());
  std::vector<uint8_t> blocks;
  const std::vector<uint8_t> gtc = GenTC::CompressDXT(w, h, rgb, std::vector<uint8_t>(), &blocks);
  std::vector<uint8_t> blocks2;
  const std::vector<uint8_t> gtc2 = GenTC::CompressDXT(w, h, rgb, blocks, &blocks2);
  try {
    GenTC::CompressDXT(128, 128, std::vector<uint8_t>(128 * 128 * 3), std::vector<uint8_t>());
    printf("128x128 did not throw\n");
    return 1;
  } catch (const GenTC::EncodeError &e) {
    if (e.code() != GSTE_ERR_ARGUMENT) return 1;
  }
  const std::vector<uint8_t> *outs[4] = {&gtc, &blocks, &gtc2, &blocks2};
  for (int k = 0; k < 4; ++k) {
    FILE *f = fopen(argv[1 + k], "wb");
    if (!f) return 2;
    fwrite(outs[k]->data(), 1, outs[k]->size(), f);
    fclose(f);
  }
  printf("ok %zu bytes\n", gtc.size());
  return 0;
}
