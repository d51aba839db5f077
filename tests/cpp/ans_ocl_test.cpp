// ans_ocl_test.cpp -- the reference's ANS decoder tests (ans/ans_ocl_test.cpp:64-343)
// restated against the C++ API (include/gst/ans.hpp).
//
// usage: ans_ocl_test
// Runs Initialization, TableRebuilding, DecodeSingleStream, DecodeInterleavedStreams
// and DecodeMultipleInterleavedStreams; prints each result and exits non-zero on the
// first failure.  The encoder is the reference's rANS step (ans/encode.cpp:49-69) over
// the frequencies NormalizeFrequencies returns, symbols from a seeded generator.
#include <cstdio>
#include <random>
#include <vector>

#include "gst/ans.hpp"

using ans::hip::HIPDecoder;
using ans::hip::kANSTableSize;
using ans::hip::kNumEncodedSymbols;

namespace {

// ans/encode.cpp:49-69 with b = 2^16, k = 2^4, M = 2048.
class Encoder {
 public:
  explicit Encoder(const std::vector<uint32_t> &norm) : f_(norm), c_(norm.size(), 0) {
    for (size_t i = 1; i < norm.size(); ++i) c_[i] = c_[i - 1] + norm[i - 1];
  }
  void Encode(uint32_t s, std::vector<uint8_t> *out) {
    while (x_ >= (f_[s] << 20)) {
      out->push_back(static_cast<uint8_t>(x_ & 0xFF));
      out->push_back(static_cast<uint8_t>((x_ >> 8) & 0xFF));
      x_ >>= 16;
    }
    x_ = (x_ / f_[s]) * kANSTableSize + c_[s] + x_ % f_[s];
  }
  uint32_t GetState() const { return x_; }

 private:
  std::vector<uint32_t> f_, c_;
  uint32_t x_ = 1u << 15;
};

std::vector<uint8_t> GenerateSymbols(const std::vector<uint32_t> &freqs, size_t n, std::mt19937 *rng) {
  std::discrete_distribution<int> d(freqs.begin(), freqs.end());
  std::vector<uint8_t> v(n);
  for (auto &s : v) s = static_cast<uint8_t>(d(*rng));
  return v;
}

int Fail(const char *test, const char *what, size_t i) {
  printf("%s: %s differ at index %zu\n", test, what, i);
  return 1;
}

int CheckTable(const char *test, const HIPDecoder &decoder, const std::vector<uint32_t> &F) {
  const std::vector<uint32_t> norm = ans::hip::NormalizeFrequencies(F);
  std::vector<uint8_t> es(kANSTableSize, 0);
  std::vector<uint16_t> ef(kANSTableSize, 0), ec(kANSTableSize, 0);
  uint32_t sum = 0;
  for (size_t i = 0; i < norm.size(); ++i) {
    for (uint32_t j = 0; j < norm[i]; ++j) {
      es[sum + j] = static_cast<uint8_t>(i);
      ef[sum + j] = static_cast<uint16_t>(norm[i]);
      ec[sum + j] = static_cast<uint16_t>(sum);
    }
    sum += norm[i];
  }
  if (sum != kANSTableSize) return Fail(test, "normalised sum", sum);
  const auto s = decoder.GetSymbols();
  const auto f = decoder.GetFrequencies();
  const auto c = decoder.GetCumulativeFrequencies();
  for (size_t i = 0; i < kANSTableSize; ++i) {
    if (s[i] != es[i]) return Fail(test, "symbols", i);
    if (f[i] != ef[i]) return Fail(test, "frequencies", i);
    if (c[i] != ec[i]) return Fail(test, "cumulative frequencies", i);
  }
  return 0;
}

int Initialization(const std::unique_ptr<gpu::GPUContext> &ctx) {
  std::vector<uint32_t> F = {3, 2, 1, 4, 3};
  HIPDecoder decoder(ctx, F, 1);
  return CheckTable("Initialization", decoder, F);
}

int TableRebuilding(const std::unique_ptr<gpu::GPUContext> &ctx) {
  std::vector<uint32_t> F = {3, 2, 1, 4, 3, 406};
  HIPDecoder decoder(ctx, F, 1);
  std::vector<uint32_t> new_F = {80, 300, 2, 14, 1, 1, 1, 20};
  decoder.RebuildTable(new_F);
  return CheckTable("TableRebuilding", decoder, new_F);
}

int DecodeSingleStream(const std::unique_ptr<gpu::GPUContext> &ctx, std::mt19937 *rng) {
  std::vector<uint32_t> F = {12, 14, 17, 1, 1, 2, 372};
  std::vector<uint8_t> symbols = GenerateSymbols(F, kNumEncodedSymbols, rng);
  Encoder encoder(ans::hip::NormalizeFrequencies(F));
  std::vector<uint8_t> stream;
  for (auto s : symbols) encoder.Encode(s, &stream);
  for (int pass = 0; pass < 2; ++pass) {  // "Make sure we can decode it twice..."
    HIPDecoder decoder(ctx, F, 1);
    std::vector<uint8_t> decoded = decoder.Decode(encoder.GetState(), stream);
    if (decoded.size() != kNumEncodedSymbols) return Fail("DecodeSingleStream", "sizes", decoded.size());
    for (size_t i = 0; i < kNumEncodedSymbols; ++i)
      if (decoded[i] != symbols[i]) return Fail("DecodeSingleStream", "symbols", i);
  }
  return 0;
}

int DecodeInterleavedStreams(const std::unique_ptr<gpu::GPUContext> &ctx, std::mt19937 *rng) {
  std::vector<uint32_t> F = {32, 186, 54, 8, 94, 35, 13, 21, 456, 789, 33, 215, 6, 54, 987, 54, 65, 13, 2, 1};
  const size_t num_interleaved = 24;
  const std::vector<uint32_t> norm = ans::hip::NormalizeFrequencies(F);
  std::vector<std::vector<uint8_t>> symbols;
  std::vector<Encoder> encoders;
  for (size_t i = 0; i < num_interleaved; ++i) {
    symbols.push_back(GenerateSymbols(F, kNumEncodedSymbols, rng));
    encoders.emplace_back(norm);
  }
  std::vector<uint8_t> stream;
  for (size_t k = 0; k < kNumEncodedSymbols; ++k)
    for (size_t j = 0; j < num_interleaved; ++j) encoders[j].Encode(symbols[j][k], &stream);
  std::vector<uint32_t> states;
  for (const auto &e : encoders) states.push_back(e.GetState());
  HIPDecoder decoder(ctx, F, static_cast<int>(num_interleaved));
  auto decoded = decoder.Decode(states, stream);
  if (decoded.size() != num_interleaved) return Fail("DecodeInterleavedStreams", "stream counts", decoded.size());
  for (size_t j = 0; j < num_interleaved; ++j)
    for (size_t k = 0; k < kNumEncodedSymbols; ++k)
      if (decoded[j][k] != symbols[j][k]) return Fail("DecodeInterleavedStreams", "symbols", j * 256 + k);
  return 0;
}

int DecodeMultipleInterleavedStreams(const std::unique_ptr<gpu::GPUContext> &ctx, std::mt19937 *rng,
                                     size_t num_interleaved, size_t num_data) {
  std::vector<uint32_t> F = {65, 4, 6, 132, 135, 64, 879, 87, 456, 13, 2, 12, 33, 16, 546, 987,
                             98, 74, 65, 43, 21, 32, 1};
  const std::vector<uint32_t> norm = ans::hip::NormalizeFrequencies(F);
  std::vector<std::vector<uint8_t>> symbols, data(num_data);
  std::vector<uint32_t> states;
  for (size_t g = 0; g < num_data; ++g) {
    std::vector<Encoder> encoders(num_interleaved, Encoder(norm));
    for (size_t j = 0; j < num_interleaved; ++j) symbols.push_back(GenerateSymbols(F, kNumEncodedSymbols, rng));
    for (size_t k = 0; k < kNumEncodedSymbols; ++k)
      for (size_t j = 0; j < num_interleaved; ++j)
        encoders[j].Encode(symbols[g * num_interleaved + j][k], &data[g]);
    for (const auto &e : encoders) states.push_back(e.GetState());
  }
  HIPDecoder decoder(ctx, F, static_cast<int>(num_interleaved));
  auto decoded = decoder.Decode(states, data);
  if (decoded.size() != symbols.size()) return Fail("DecodeMultipleInterleavedStreams", "stream counts", 0);
  for (size_t s = 0; s < symbols.size(); ++s)
    for (size_t k = 0; k < kNumEncodedSymbols; ++k)
      if (decoded[s][k] != symbols[s][k]) return Fail("DecodeMultipleInterleavedStreams", "symbols", s * 256 + k);
  return 0;
}

}  // namespace

int main() {
  std::unique_ptr<gpu::GPUContext> ctx = gpu::GPUContext::InitializeHIP(0);
  if (!ctx) {
    fprintf(stderr, "no HIP device\n");
    return 2;
  }
  std::mt19937 rng(0);
  struct Case {
    const char *name;
    int rc;
  };
  try {
    const Case cases[] = {
        {"Initialization", Initialization(ctx)},
        {"TableRebuilding", TableRebuilding(ctx)},
        {"DecodeSingleStream", DecodeSingleStream(ctx, &rng)},
        {"DecodeInterleavedStreams", DecodeInterleavedStreams(ctx, &rng)},
        {"DecodeMultipleInterleavedStreams", DecodeMultipleInterleavedStreams(ctx, &rng, 1, 2)},
        {"DecodeMultipleInterleavedStreams (32 x 6)", DecodeMultipleInterleavedStreams(ctx, &rng, 32, 6)},
    };
    int failed = 0;
    for (const Case &c : cases) {
      printf("%-45s %s\n", c.name, c.rc ? "FAILED" : "ok");
      failed += c.rc != 0;
    }
    return failed ? 1 : 0;
  } catch (const GenTC::DecodeError &e) {
    printf("error: %s\n", e.what());
    return 1;
  }
}
