// gst_encode.cpp -- seeded synthetic textures and a CPU encoder for the GST format.
//
// Input tooling for tests and benchmarks (see include/gst_encode.h).  The format it
// writes is the reference's (GammaUNC/GST):
//   file    = GenTCHeader (7 x u32) | 4 x u16[256] frequencies | Y | chroma | palette | idx
//             (codec/encoder.cpp:122-144)
//   stream  = u32 end_offset[G] | G groups; group = [2 pad bytes if needed] u16 words
//             in emission order | u32 state[32]   (codec/entropy.cpp:174-265,
//             ans/encode.cpp:224-259); 8192 symbols per group, 256 per stream
//   planes  = Y1,Y2 | Co1,Cg1,Co2,Cg2 of the endpoint images in YCoCg667, each
//             forward 5/3-wavelet transformed per 32x32 tile (5 levels), +128,
//             tiles in raster order, raster inside a tile (codec/encoder.cpp:39-98)
//   palette = u32 DXT1 index words, zero-padded to 8192 B; idx = (i_k - i_{k-1}) + 128
//             (codec/dxt_image.cpp:597-619)
// The rANS parameters are the reference decoder's: M = 2^11, L = 2^15, 16-bit
// renormalisation (ans/ans_ocl_encode.cpp:10-18, ans/ans_decode.cl:4-8).
#include "gst_encode.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kM = 2048;
constexpr uint32_t kMLog = 11;
constexpr uint32_t kStateL = 32768;
constexpr uint32_t kLanes = 32;
constexpr uint32_t kSymsPerLane = 256;
constexpr uint32_t kGroupSyms = kLanes * kSymsPerLane;
constexpr int kTile = 32;

// ---------------------------------------------------------------------------
// Seeded generator (SplitMix64 stream + integer hash for the noise lattice).

struct SplitMix64 {
  uint64_t s;
  uint64_t Next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // Uniform integer in [lo, hi].
  int64_t Range(int64_t lo, int64_t hi) {
    const uint64_t span = static_cast<uint64_t>(hi - lo) + 1;
    return lo + static_cast<int64_t>(((Next() >> 32) * span) >> 32);
  }
};

uint32_t Hash4(uint64_t seed, uint32_t a, uint32_t b, uint32_t c) {
  uint64_t h = seed ^ (0x9E3779B97F4A7C15ull * (a + 1)) ^ (0xC2B2AE3D27D4EB4Full * (b + 7)) ^
               (0x165667B19E3779F9ull * (c + 13));
  h ^= h >> 33;
  h *= 0xFF51AFD7ED558CCDull;
  h ^= h >> 33;
  h *= 0xC4CEB9FE1A85EC53ull;
  h ^= h >> 33;
  return static_cast<uint32_t>(h);
}

// Smoothstep of t in Q16: t^2 (3 - 2t).
int64_t Smooth16(int64_t t) { return (((t * t) >> 16) * (3 * 65536 - 2 * t)) >> 16; }

// Value noise octave with a `period`-pixel lattice, result in [0, 65535].
int64_t ValueNoise(uint64_t seed, uint32_t oct, uint32_t ch, uint32_t x, uint32_t y,
                   uint32_t period) {
  const uint32_t ix = x / period, iy = y / period;
  const int64_t fx = Smooth16((static_cast<int64_t>(x % period) << 16) / period);
  const int64_t fy = Smooth16((static_cast<int64_t>(y % period) << 16) / period);
  auto lat = [&](uint32_t gx, uint32_t gy) -> int64_t {
    return Hash4(seed, oct * 4 + ch, gx, gy) & 0xFFFF;
  };
  const int64_t a = lat(ix, iy), b = lat(ix + 1, iy), c = lat(ix, iy + 1), d = lat(ix + 1, iy + 1);
  const int64_t top = a + (((b - a) * fx) >> 16);
  const int64_t bot = c + (((d - c) * fx) >> 16);
  return top + (((bot - top) * fy) >> 16);
}

// ---------------------------------------------------------------------------
// DXT1 block compression: principal axis from the colour covariance, extreme
// projections as endpoints, one least-squares refit, nearest-colour indices.

struct Rgb565 {
  static uint16_t Pack(int r, int g, int b) {
    return static_cast<uint16_t>((r << 11) | (g << 5) | b);
  }
  static void Expand(uint16_t c, int out[3]) {
    const int r = (c >> 11) & 31, g = (c >> 5) & 63, b = c & 31;
    out[0] = (r << 3) | (r >> 2);
    out[1] = (g << 2) | (g >> 4);
    out[2] = (b << 3) | (b >> 2);
  }
  static uint16_t Quantize(const float c[3]) {
    auto q = [](float v, int maxv) {
      int r = static_cast<int>(std::lround(std::min(255.f, std::max(0.f, v)) * maxv / 255.f));
      return std::min(maxv, std::max(0, r));
    };
    return Pack(q(c[0], 31), q(c[1], 63), q(c[2], 31));
  }
};

// DXT1 4-colour palette of endpoints c0 > c1.
void BlockPalette(uint16_t c0, uint16_t c1, int pal[4][3]) {
  Rgb565::Expand(c0, pal[0]);
  Rgb565::Expand(c1, pal[1]);
  for (int k = 0; k < 3; ++k) {
    pal[2][k] = (2 * pal[0][k] + pal[1][k]) / 3;
    pal[3][k] = (pal[0][k] + 2 * pal[1][k]) / 3;
  }
}

int Dist2(const int a[3], const uint8_t *p) {
  const int dr = a[0] - p[0], dg = a[1] - p[1], db = a[2] - p[2];
  return dr * dr + dg * dg + db * db;
}

// Indices (2 bits per texel, texel i = row*4+col at bits 2i) and squared error.
uint32_t AssignIndices(const uint8_t px[16][3], uint16_t c0, uint16_t c1, int *sse) {
  int pal[4][3];
  BlockPalette(c0, c1, pal);
  uint32_t word = 0;
  int err = 0;
  for (int i = 0; i < 16; ++i) {
    int best = 0, bd = Dist2(pal[0], px[i]);
    for (int k = 1; k < 4; ++k) {
      const int d = Dist2(pal[k], px[i]);
      if (d < bd) { bd = d; best = k; }
    }
    word |= static_cast<uint32_t>(best) << (2 * i);
    err += bd;
  }
  *sse = err;
  return word;
}

// Orders endpoints for 4-colour mode (c0 > c1), fixing the degenerate c0 == c1.
bool OrderEndpoints(uint16_t *c0, uint16_t *c1) {
  if (*c0 < *c1) std::swap(*c0, *c1);
  if (*c0 == *c1) {
    if (*c0 == 0) { *c0 = 1; return true; }  // all texels take index 1 (= colour 0)
    *c1 = static_cast<uint16_t>(*c0 - 1);
  }
  return false;
}

uint64_t CompressBlock(const uint8_t px[16][3]) {
  float mean[3] = {0, 0, 0};
  for (int i = 0; i < 16; ++i)
    for (int k = 0; k < 3; ++k) mean[k] += px[i][k];
  for (int k = 0; k < 3; ++k) mean[k] /= 16.f;
  float cov[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 16; ++i) {
    const float r = px[i][0] - mean[0], g = px[i][1] - mean[1], b = px[i][2] - mean[2];
    cov[0] += r * r; cov[1] += r * g; cov[2] += r * b;
    cov[3] += g * g; cov[4] += g * b; cov[5] += b * b;
  }
  float axis[3] = {1.f, 1.f, 1.f};
  for (int it = 0; it < 6; ++it) {
    const float x = cov[0] * axis[0] + cov[1] * axis[1] + cov[2] * axis[2];
    const float y = cov[1] * axis[0] + cov[3] * axis[1] + cov[4] * axis[2];
    const float z = cov[2] * axis[0] + cov[4] * axis[1] + cov[5] * axis[2];
    const float n = std::max(std::fabs(x), std::max(std::fabs(y), std::fabs(z)));
    if (n < 1e-6f) break;
    axis[0] = x / n; axis[1] = y / n; axis[2] = z / n;
  }
  const float an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  for (int k = 0; k < 3; ++k) axis[k] /= an;
  float tmin = 1e30f, tmax = -1e30f;
  for (int i = 0; i < 16; ++i) {
    const float t = (px[i][0] - mean[0]) * axis[0] + (px[i][1] - mean[1]) * axis[1] +
                    (px[i][2] - mean[2]) * axis[2];
    tmin = std::min(tmin, t);
    tmax = std::max(tmax, t);
  }
  float e0[3], e1[3];
  for (int k = 0; k < 3; ++k) {
    e0[k] = mean[k] + tmax * axis[k];
    e1[k] = mean[k] + tmin * axis[k];
  }
  uint16_t c0 = Rgb565::Quantize(e0), c1 = Rgb565::Quantize(e1);
  int sse = 0;
  uint32_t word;
  if (OrderEndpoints(&c0, &c1)) {
    word = 0x55555555u;
    int pal[4][3];
    BlockPalette(c0, c1, pal);
    for (int i = 0; i < 16; ++i) sse += Dist2(pal[1], px[i]);
  } else {
    word = AssignIndices(px, c0, c1, &sse);
  }

  // One least-squares refit of the endpoints for the chosen indices.
  static const float kW0[4] = {1.f, 0.f, 2.f / 3.f, 1.f / 3.f};
  float aa = 0, bb = 0, ab = 0, ax[3] = {0, 0, 0}, bx[3] = {0, 0, 0};
  for (int i = 0; i < 16; ++i) {
    const float a = kW0[(word >> (2 * i)) & 3], b = 1.f - a;
    aa += a * a; bb += b * b; ab += a * b;
    for (int k = 0; k < 3; ++k) { ax[k] += a * px[i][k]; bx[k] += b * px[i][k]; }
  }
  const float det = aa * bb - ab * ab;
  if (std::fabs(det) > 1e-3f) {
    float p0[3], p1[3];
    for (int k = 0; k < 3; ++k) {
      p0[k] = (ax[k] * bb - bx[k] * ab) / det;
      p1[k] = (bx[k] * aa - ax[k] * ab) / det;
    }
    uint16_t r0 = Rgb565::Quantize(p0), r1 = Rgb565::Quantize(p1);
    if (!OrderEndpoints(&r0, &r1)) {
      int sse2 = 0;
      const uint32_t w2 = AssignIndices(px, r0, r1, &sse2);
      if (sse2 < sse) { c0 = r0; c1 = r1; word = w2; sse = sse2; }
    }
  }
  return static_cast<uint64_t>(c0) | (static_cast<uint64_t>(c1) << 16) |
         (static_cast<uint64_t>(word) << 32);
}

void LoadBlockPixels(const uint8_t *rgb, uint32_t width, uint32_t bx, uint32_t by,
                     uint8_t px[16][3]) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      const uint8_t *p = rgb + 3 * (static_cast<size_t>(by * 4 + r) * width + bx * 4 + c);
      px[r * 4 + c][0] = p[0];
      px[r * 4 + c][1] = p[1];
      px[r * 4 + c][2] = p[2];
    }
}

// ---------------------------------------------------------------------------
// Forward 5/3 lifting, the exact integer inverse of the decoder's inverse step
// (codec/inverse_wavelet.cl:28-64): high-pass first, then low-pass update.

int Mirror(int idx, int range) {
  const int v = idx - static_cast<int>(idx >= range) * (idx - range + 2);
  return v < 0 ? -v : v;
}

void Forward1D(const int *v, int *out, int len, int stride_in, int stride_out) {
  const int mid = len / 2;
  int d[kTile], s[kTile];
  for (int k = 0; k < mid; ++k)
    d[k] = v[(2 * k + 1) * stride_in] -
           (v[2 * k * stride_in] + v[Mirror(2 * k + 2, len) * stride_in]) / 2;
  for (int k = 0; k < mid; ++k) {
    const int prev = Mirror(2 * k - 1, len) / 2, next = Mirror(2 * k + 1, len) / 2;
    s[k] = v[2 * k * stride_in] + (d[prev] + d[next] + 2) / 4;
  }
  for (int k = 0; k < mid; ++k) {
    out[k * stride_out] = s[k];
    out[(mid + k) * stride_out] = d[k];
  }
}

// Five levels over one 32x32 tile (raster), columns then rows per level, levels
// 32 -> 2; returns how many coefficients fell outside [-128, 127].
uint32_t ForwardTile(int t[kTile * kTile]) {
  int tmp[kTile];
  for (int dim = kTile; dim > 1; dim /= 2) {
    for (int x = 0; x < dim; ++x) {
      Forward1D(t + x, tmp, dim, kTile, 1);
      for (int y = 0; y < dim; ++y) t[y * kTile + x] = tmp[y];
    }
    for (int y = 0; y < dim; ++y) {
      Forward1D(t + y * kTile, tmp, dim, 1, 1);
      for (int x = 0; x < dim; ++x) t[y * kTile + x] = tmp[x];
    }
  }
  uint32_t over = 0;
  for (int i = 0; i < kTile * kTile; ++i) over += (t[i] < -128 || t[i] > 127);
  return over;
}

// ---------------------------------------------------------------------------
// rANS (b = 2^16, L = 2^15, M = 2^11), 32 interleaved streams per group.

// Normalises symbol counts to sum kM with every present symbol >= 1.
void NormalizeCounts(const uint64_t counts[256], uint16_t freq[256]) {
  uint64_t total = 0;
  for (int i = 0; i < 256; ++i) total += counts[i];
  int64_t f[256];
  int64_t sum = 0;
  std::vector<std::pair<uint64_t, int>> rem;
  for (int i = 0; i < 256; ++i) {
    if (!counts[i]) { f[i] = 0; continue; }
    const uint64_t scaled = counts[i] * kM;
    f[i] = std::max<int64_t>(1, static_cast<int64_t>(scaled / total));
    sum += f[i];
    rem.emplace_back(scaled % total, i);
  }
  std::sort(rem.begin(), rem.end(), [](const std::pair<uint64_t, int> &a,
                                       const std::pair<uint64_t, int> &b) {
    return a.first != b.first ? a.first > b.first : a.second < b.second;
  });
  for (size_t j = 0; sum < kM; j = (j + 1) % rem.size()) { f[rem[j].second]++; sum++; }
  while (sum > kM) {
    int big = -1;
    for (int i = 0; i < 256; ++i)
      if (f[i] > 1 && (big < 0 || f[i] > f[big])) big = i;
    f[big]--;
    sum--;
  }
  for (int i = 0; i < 256; ++i) freq[i] = static_cast<uint16_t>(f[i]);
}

// Appends one encoded group of kGroupSyms symbols to `out`.
void EncodeGroup(const uint8_t *sym, const uint16_t freq[256], const uint16_t cum[256],
                 std::vector<uint8_t> *out) {
  uint32_t x[kLanes];
  for (uint32_t j = 0; j < kLanes; ++j) x[j] = kStateL;
  std::vector<uint16_t> words;
  words.reserve(kGroupSyms);
  for (uint32_t i = 0; i < kSymsPerLane; ++i) {
    for (uint32_t j = 0; j < kLanes; ++j) {
      const uint8_t s = sym[j * kSymsPerLane + i];
      const uint32_t f = freq[s];
      if (x[j] >= (f << 20)) {  // b * (L / M) * f
        words.push_back(static_cast<uint16_t>(x[j] & 0xFFFF));
        x[j] >>= 16;
      }
      x[j] = (x[j] / f) * kM + (x[j] % f) + cum[s];
    }
  }
  const size_t bytes = 2 * words.size() + 4 * kLanes;
  if (bytes & 3) { out->push_back(0); out->push_back(0); }
  for (uint16_t w : words) { out->push_back(static_cast<uint8_t>(w)); out->push_back(static_cast<uint8_t>(w >> 8)); }
  for (uint32_t j = 0; j < kLanes; ++j)
    for (int b = 0; b < 4; ++b) out->push_back(static_cast<uint8_t>(x[j] >> (8 * b)));
}

// One stream: frequency table (u16[256]) and bytes [u32 end_offset[G]][groups].
void EncodeStream(const std::vector<uint8_t> &symbols, uint16_t freq[256],
                  std::vector<uint8_t> *stream) {
  uint64_t counts[256] = {0};
  for (uint8_t s : symbols) counts[s]++;
  NormalizeCounts(counts, freq);
  uint16_t cum[256];
  uint32_t run = 0;
  for (int i = 0; i < 256; ++i) { cum[i] = static_cast<uint16_t>(run); run += freq[i]; }
  const size_t G = symbols.size() / kGroupSyms;
  std::vector<uint8_t> body;
  std::vector<uint32_t> ends(G);
  for (size_t g = 0; g < G; ++g) {
    EncodeGroup(symbols.data() + g * kGroupSyms, freq, cum, &body);
    ends[g] = static_cast<uint32_t>(4 * G + body.size());
  }
  stream->clear();
  stream->reserve(4 * G + body.size());
  for (uint32_t e : ends)
    for (int b = 0; b < 4; ++b) stream->push_back(static_cast<uint8_t>(e >> (8 * b)));
  stream->insert(stream->end(), body.begin(), body.end());
}

void Put32(uint8_t *p, uint32_t v) {
  for (int b = 0; b < 4; ++b) p[b] = static_cast<uint8_t>(v >> (8 * b));
}

}  // namespace

extern "C" {

void gste_generate_rgb(uint64_t seed, uint32_t width, uint32_t height, uint32_t noise_milli,
                       uint8_t *rgb) {
  SplitMix64 rng{seed};
  int64_t base[3], amp[3][2], sx[3][2], sy[3][2];
  for (int c = 0; c < 3; ++c) {
    base[c] = rng.Range(48, 208);
    for (int k = 0; k < 2; ++k) {
      amp[c][k] = rng.Range(-80, 80);
      sx[c][k] = rng.Range(-1024, 1024);
      sy[c][k] = rng.Range(-1024, 1024);
    }
  }
  static const uint32_t kPeriods[4] = {64, 32, 16, 8};
  static const int64_t kWeights[4] = {8, 4, 2, 1};
  const uint64_t nseed = rng.Next();
  for (uint32_t y = 0; y < height; ++y) {
    for (uint32_t x = 0; x < width; ++x) {
      for (uint32_t c = 0; c < 3; ++c) {
        int64_t v = base[c];
        for (int k = 0; k < 2; ++k)
          v += amp[c][k] * (sx[c][k] * x / width + sy[c][k] * y / height) / 1024;
        int64_t n = 0;
        for (uint32_t o = 0; o < 4; ++o) n += kWeights[o] * ValueNoise(nseed, o, c, x, y, kPeriods[o]);
        n /= 15;  // [0, 65535]
        v += (n - 32768) * 255 * static_cast<int64_t>(noise_milli) / (65536 * 1000);
        rgb[3 * (static_cast<size_t>(y) * width + x) + c] =
            static_cast<uint8_t>(std::min<int64_t>(255, std::max<int64_t>(0, v)));
      }
    }
  }
  for (int r = 0; r < 8; ++r) {
    const int64_t rw = rng.Range(width / 16, width / 3), rh = rng.Range(height / 16, height / 3);
    const int64_t x0 = rng.Range(0, width - 1), y0 = rng.Range(0, height - 1);
    int64_t col[3];
    for (int c = 0; c < 3; ++c) col[c] = rng.Range(0, 255);
    const uint32_t x1 = static_cast<uint32_t>(std::min<int64_t>(width, x0 + rw));
    const uint32_t y1 = static_cast<uint32_t>(std::min<int64_t>(height, y0 + rh));
    for (uint32_t y = static_cast<uint32_t>(y0); y < y1; ++y)
      for (uint32_t x = static_cast<uint32_t>(x0); x < x1; ++x)
        for (int c = 0; c < 3; ++c) {
          const int64_t n = (ValueNoise(nseed, 3, c, x, y, 8) - 32768) * 255 *
                            static_cast<int64_t>(noise_milli) / (65536 * 2000);
          rgb[3 * (static_cast<size_t>(y) * width + x) + c] =
              static_cast<uint8_t>(std::min<int64_t>(255, std::max<int64_t>(0, col[c] + n)));
        }
  }
}

void gste_compress_dxt1(const uint8_t *rgb, uint32_t width, uint32_t height, uint8_t *dxt) {
  const uint32_t bxn = width / 4, byn = height / 4;
  for (uint32_t by = 0; by < byn; ++by)
    for (uint32_t bx = 0; bx < bxn; ++bx) {
      uint8_t px[16][3];
      LoadBlockPixels(rgb, width, bx, by, px);
      const uint64_t blk = CompressBlock(px);
      memcpy(dxt + 8 * (static_cast<size_t>(by) * bxn + bx), &blk, 8);
    }
}

size_t gste_max_encoded_size(uint32_t width, uint32_t height) {
  const size_t nv = static_cast<size_t>(width) * height / 16;
  const size_t symbols = 7 * nv + 4 * nv;
  // <= one 16-bit word per symbol, plus 128 B states and 4 B offset per group.
  return 28 + 2048 + 2 * symbols + (symbols / kGroupSyms + 4) * (128 + 4 + 2);
}

size_t gste_encode(uint8_t *dxt, const uint8_t *rgb, uint32_t width, uint32_t height,
                   uint32_t reuse_threshold, uint8_t *out, size_t cap, gste_stats *stats) {
  if (!dxt || !out || width == 0 || height == 0 || width % 128 || height % 128 ||
      (static_cast<uint64_t>(width) * height) % 131072)
    return 0;
  if (reuse_threshold && !rgb) return 0;
  const uint32_t bxn = width / 4, byn = height / 4, nv = bxn * byn;
  gste_stats st = {0, 0, 0, 0};

  // Palette of index words with reuse among the last 127 entries
  // (codec/dxt_image.cpp:538-620 defines the diff coding; the search here is ours).
  std::vector<uint32_t> palette;
  std::vector<uint8_t> idx_diffs(nv);
  int last = 0;
  for (uint32_t b = 0; b < nv; ++b) {
    uint64_t blk;
    memcpy(&blk, dxt + 8 * static_cast<size_t>(b), 8);
    const uint16_t c0 = static_cast<uint16_t>(blk), c1 = static_cast<uint16_t>(blk >> 16);
    const uint32_t word = static_cast<uint32_t>(blk >> 32);
    int chosen = -1;
    const size_t look = std::min<size_t>(127, palette.size());
    if (reuse_threshold && look) {
      uint8_t px[16][3];
      LoadBlockPixels(rgb, width, b % bxn, b / bxn, px);
      int pal[4][3];
      BlockPalette(c0, c1, pal);
      int dist[16][4];
      for (int i = 0; i < 16; ++i)
        for (int k = 0; k < 4; ++k) dist[i][k] = Dist2(pal[k], px[i]);
      int orig = 0;
      for (int i = 0; i < 16; ++i) orig += dist[i][(word >> (2 * i)) & 3];
      orig /= 48;
      int best = 1 << 30;
      for (size_t k = 0; k < look; ++k) {
        const uint32_t cand = palette[palette.size() - 1 - k];
        int e = 0;
        for (int i = 0; i < 16; ++i) e += dist[i][(cand >> (2 * i)) & 3];
        const int diff = e / 48 - orig;
        if (diff < best) {
          best = diff;
          chosen = static_cast<int>(k);
          if (diff <= 0) break;
        }
      }
      if (best >= static_cast<int>(reuse_threshold)) chosen = -1;
    } else if (!reuse_threshold && look) {
      for (size_t k = 0; k < look; ++k)
        if (palette[palette.size() - 1 - k] == word) { chosen = static_cast<int>(k); break; }
    }
    int this_index;
    if (chosen >= 0) {
      this_index = static_cast<int>(palette.size()) - 1 - chosen;
      const uint32_t nw = palette[static_cast<size_t>(this_index)];
      if (nw != word) st.reused_blocks++;
      memcpy(dxt + 8 * static_cast<size_t>(b) + 4, &nw, 4);
    } else {
      this_index = static_cast<int>(palette.size());
      palette.push_back(word);
    }
    idx_diffs[b] = static_cast<uint8_t>(this_index - last + 128);
    last = this_index;
  }
  st.palette_entries = static_cast<uint32_t>(palette.size());
  const size_t P = (4 * palette.size() + kGroupSyms - 1) / kGroupSyms * kGroupSyms;

  // Endpoint planes in YCoCg667 (forward of codec/image_processing.cpp:29-46's
  // inverse), wavelet-transformed and tile-linearised.
  std::vector<int> planes(6 * static_cast<size_t>(nv));
  for (uint32_t b = 0; b < nv; ++b) {
    uint64_t blk;
    memcpy(&blk, dxt + 8 * static_cast<size_t>(b), 8);
    for (int e = 0; e < 2; ++e) {
      const uint16_t c = static_cast<uint16_t>(blk >> (16 * e));
      const int r = (c >> 11) & 31, g = (c >> 5) & 63, bl = c & 31;
      const int co = r - bl;
      const int t = r + bl + (bl >> 4);
      const int cg = g - t;
      const int y = t + cg / 2;
      planes[static_cast<size_t>(e) * nv + b] = y;
      planes[static_cast<size_t>(2 + 2 * e) * nv + b] = co;
      planes[static_cast<size_t>(3 + 2 * e) * nv + b] = cg;
    }
  }
  std::vector<uint8_t> ysym(2 * static_cast<size_t>(nv)), csym(4 * static_cast<size_t>(nv));
  const uint32_t tiles_x = bxn / kTile, tiles_y = byn / kTile;
  for (int p = 0; p < 6; ++p) {
    uint8_t *dst = p < 2 ? ysym.data() + static_cast<size_t>(p) * nv
                         : csym.data() + static_cast<size_t>(p - 2) * nv;
    const int *src = planes.data() + static_cast<size_t>(p) * nv;
    for (uint32_t ty = 0; ty < tiles_y; ++ty)
      for (uint32_t tx = 0; tx < tiles_x; ++tx) {
        int t[kTile * kTile];
        for (int y = 0; y < kTile; ++y)
          for (int x = 0; x < kTile; ++x)
            t[y * kTile + x] = src[static_cast<size_t>(ty * kTile + y) * bxn + tx * kTile + x];
        st.wavelet_overflows += ForwardTile(t);
        uint8_t *o = dst + static_cast<size_t>(ty * tiles_x + tx) * kTile * kTile;
        for (int i = 0; i < kTile * kTile; ++i) o[i] = static_cast<uint8_t>(t[i] + 128);
      }
  }
  std::vector<uint8_t> psym(P, 0);
  memcpy(psym.data(), palette.data(), 4 * palette.size());

  const std::vector<uint8_t> *sym[4] = {&ysym, &csym, &psym, &idx_diffs};
  std::vector<uint8_t> streams[4];
  uint16_t freq[4][256];
  for (int k = 0; k < 4; ++k) EncodeStream(*sym[k], freq[k], &streams[k]);
  size_t total = 28 + 2048;
  for (int k = 0; k < 4; ++k) total += streams[k].size();
  if (total > cap) return 0;
  Put32(out + 0, width);
  Put32(out + 4, height);
  Put32(out + 8, static_cast<uint32_t>(P));
  for (int k = 0; k < 4; ++k) Put32(out + 12 + 4 * k, static_cast<uint32_t>(streams[k].size()));
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 256; ++i) {
      out[28 + 512 * k + 2 * i] = static_cast<uint8_t>(freq[k][i]);
      out[28 + 512 * k + 2 * i + 1] = static_cast<uint8_t>(freq[k][i] >> 8);
    }
  size_t off = 28 + 2048;
  for (int k = 0; k < 4; ++k) {
    memcpy(out + off, streams[k].data(), streams[k].size());
    off += streams[k].size();
  }
  if (stats) *stats = st;
  return total;
}

size_t gste_make_texture(uint64_t seed, uint32_t width, uint32_t height, uint32_t noise_milli,
                         uint32_t reuse_threshold, uint8_t *out, size_t cap, uint8_t *dxt_out,
                         gste_stats *stats) {
  if (width == 0 || height == 0 || width % 128 || height % 128) return 0;
  std::vector<uint8_t> rgb(3 * static_cast<size_t>(width) * height);
  gste_generate_rgb(seed, width, height, noise_milli, rgb.data());
  std::vector<uint8_t> dxt(static_cast<size_t>(width) * height / 2);
  gste_compress_dxt1(rgb.data(), width, height, dxt.data());
  const size_t n = gste_encode(dxt.data(), rgb.data(), width, height, reuse_threshold, out, cap, stats);
  if (n && dxt_out) memcpy(dxt_out, dxt.data(), dxt.size());
  return n;
}

void gste_make_textures(uint32_t count, const uint64_t *seeds, const uint32_t *widths,
                        const uint32_t *heights, uint32_t noise_milli, uint32_t reuse_threshold,
                        uint8_t *out, const uint64_t *out_offsets, uint8_t *dxt,
                        const uint64_t *dxt_offsets, uint64_t *sizes, int nthreads) {
  std::atomic<uint32_t> next(0);
  auto work = [&]() {
    for (;;) {
      const uint32_t i = next.fetch_add(1);
      if (i >= count) break;
      sizes[i] = gste_make_texture(seeds[i], widths[i], heights[i], noise_milli, reuse_threshold,
                                   out + out_offsets[i], gste_max_encoded_size(widths[i], heights[i]),
                                   dxt ? dxt + dxt_offsets[i] : nullptr, nullptr);
    }
  };
  const int n = std::max(1, std::min(nthreads, static_cast<int>(count)));
  std::vector<std::thread> th;
  for (int t = 1; t < n; ++t) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
}

void gste_make_textures_gentenc(uint32_t count, const uint64_t *seeds, const uint32_t *widths,
                                const uint32_t *heights, uint32_t noise_milli, uint8_t *out,
                                const uint64_t *out_offsets, uint8_t *dxt, const uint64_t *dxt_offsets,
                                uint64_t *sizes, int nthreads) {
  // Largest first, so that the big textures do not start last.
  std::vector<uint32_t> order(count);
  for (uint32_t i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return static_cast<uint64_t>(widths[a]) * heights[a] > static_cast<uint64_t>(widths[b]) * heights[b];
  });
  std::atomic<uint32_t> next(0);
  auto work = [&]() {
    std::vector<uint8_t> rgb;
    for (;;) {
      const uint32_t k = next.fetch_add(1);
      if (k >= count) break;
      const uint32_t i = order[k];
      sizes[i] = 0;
      if (widths[i] == 0 || heights[i] == 0 || widths[i] % 128 || heights[i] % 128) continue;
      rgb.resize(3 * static_cast<size_t>(widths[i]) * heights[i]);
      gste_generate_rgb(seeds[i], widths[i], heights[i], noise_milli, rgb.data());
      size_t n = 0;
      if (gste_compress_dxt(rgb.data(), widths[i], heights[i], nullptr, out + out_offsets[i],
                            gste_max_encoded_size(widths[i], heights[i]), &n,
                            dxt ? dxt + dxt_offsets[i] : nullptr) == GSTE_OK)
        sizes[i] = n;
    }
  };
  const int n = std::max(1, std::min(nthreads, static_cast<int>(count)));
  std::vector<std::thread> th;
  for (int t = 1; t < n; ++t) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
}

}  // extern "C"
