// gst_gentc.cpp -- the reference encoder, byte for byte (gste_compress_dxt in
// include/gst_encode.h).
//
// GenTC::CompressDXT (codec/encoder.cpp:39-187) writes a .gtc file from an RGB8 image:
//   1. DXT1 blocks: given, or compressed block by block with stb_dxt v1.04 in its
//      high-quality mode (codec/dxt_image.cpp:362-385, lib/include/stb_dxt.h);
//   2. Reencode (codec/dxt_image.cpp:512-624): each block may take one of the 127 most
//      recent palette index words, its endpoints refit by least squares, when that
//      raises its error by less than 35; the palette and the index steps are recorded;
//   3. the endpoint planes: RGB565 -> YCoCg667, one 5-level 5/3 wavelet per 32x32 tile,
//      +128, tiles in raster order (codec/encoder.cpp:20-37);
//   4. four rANS streams with frequencies from GenerateHistogram (codec/entropy.cpp:174-265).
// Every step keeps the reference's integer and float arithmetic in the reference's
// order, so the file and the blocks are the ones gentenc writes for the same input.
// That includes its edge behaviour: a wavelet coefficient outside a byte wraps (the
// reference's range assertion compares after the cast to int8 and never fires, so
// such a file decodes to other endpoints than the blocks it reports); a stream made
// only of zero bytes is refused (GSTE_ERR_UNCODABLE) where gentenc stops on its rANS
// options assertion; an index step outside a byte, which the 127-entry lookback cannot
// produce, would be GSTE_ERR_RANGE where gentenc asserts.  Float to int conversions of
// out-of-range values (NaN or infinity where a refit's system is singular) give
// INT_MIN as on x86-64, the reference's host.  Build with -ffp-contract=off: no fused
// multiply-adds.
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "gst_encode.h"
#include "gst_histogram.h"

namespace {

constexpr uint32_t kTableM = 2048;     // ans::ocl::kANSTableSize
constexpr uint32_t kStateL = 1u << 15; // ans/ans_ocl_encode.cpp:10-18
constexpr uint32_t kLanes = 32;        // ans::ocl::kThreadsPerEncodingGroup
constexpr uint32_t kLaneSyms = 256;    // ans::ocl::kNumEncodedSymbols
constexpr uint32_t kGroupSyms = kLanes * kLaneSyms;
constexpr int kTile = 32;              // kWaveletBlockDim, codec/codec_base.h:22
constexpr int kErrThreshold = 35;      // codec/dxt_image.cpp:510
constexpr size_t kLookback = 127;      // kNumPrevLookup - 1, codec/dxt_image.cpp:511,569

// (int)x for a float or double as the x86-64 truncating conversion gives it: values
// outside int32 (and NaN) become INT_MIN.
template <typename F>
int TruncToInt(F x) {
  return (x > static_cast<F>(-2147483649.0) && x < static_cast<F>(2147483648.0)) ? static_cast<int>(x) : INT_MIN;
}

// ---------------------------------------------------------------------------
// stb_dxt v1.04 colour-block compression, mode STB_DXT_HIGHQUAL, no dithering, no
// rounding bias (lib/include/stb_dxt.h:64-541).

class StbDxt {
 public:
  static const StbDxt &Get() {
    static const StbDxt tables;
    return tables;
  }

  // One 4x4 block of RGBA (alpha ignored except by the constant-block test) to the
  // 8-byte DXT1 block {max16, min16, 2-bit indices} as a little-endian u64.
  uint64_t Compress(const uint8_t blk[64]) const {
    uint16_t hi, lo;
    uint32_t mask;
    int i = 1;
    while (i < 16 && std::memcmp(blk + 4 * i, blk, 4) == 0) ++i;
    if (i == 16) {  // stb_dxt.h:483-488
      mask = 0xaaaaaaaau;
      MatchConstant(blk[0], blk[1], blk[2], &hi, &lo);
    } else {
      PrincipalEndpoints(blk, &hi, &lo);  // stb_dxt.h:494-499
      mask = hi != lo ? Match(blk, hi, lo) : 0u;
      for (int pass = 0; pass < 2; ++pass) {  // stb_dxt.h:502-516, two refinements
        const uint32_t before = mask;
        if (Refine(blk, &hi, &lo, mask)) {
          if (hi == lo) {
            mask = 0;
            break;
          }
          mask = Match(blk, hi, lo);
        }
        if (mask == before) break;
      }
    }
    if (hi < lo) {  // stb_dxt.h:520-527
      std::swap(hi, lo);
      mask ^= 0x55555555u;
    }
    return static_cast<uint64_t>(hi) | static_cast<uint64_t>(lo) << 16 | static_cast<uint64_t>(mask) << 32;
  }

 private:
  uint8_t expand5_[32], expand6_[64];
  uint8_t best5_[256][2], best6_[256][2];  // per 8-bit value: {max, min} code of a constant

  StbDxt() {  // stb_dxt.h:593-611 (the dither tables are not needed)
    for (int v = 0; v < 32; ++v) expand5_[v] = static_cast<uint8_t>((v << 3) | (v >> 2));
    for (int v = 0; v < 64; ++v) expand6_[v] = static_cast<uint8_t>((v << 2) | (v >> 4));
    BestPairs(expand5_, 32, best5_);
    BestPairs(expand6_, 64, best6_);
  }

  static int Scale8(int a, int b) {  // stb__Mul8Bit: a * b / 255, rounded
    const int t = a * b + 128;
    return (t + (t >> 8)) >> 8;
  }
  static int Third(int a, int b) { return (2 * a + b) / 3; }  // stb__Lerp13 without bias
  static uint16_t To565(int r, int g, int b) {
    return static_cast<uint16_t>((Scale8(r, 31) << 11) + (Scale8(g, 63) << 5) + Scale8(b, 31));
  }

  // stb_dxt.h:111-137: for each 8-bit value, the code pair whose 1/3 point lies
  // nearest, with 3% of the pair's spread added as error; the first best pair wins.
  static void BestPairs(const uint8_t *expand, int n, uint8_t out[256][2]) {
    for (int v = 0; v < 256; ++v) {
      int best = 256;
      for (int lo = 0; lo < n; ++lo)
        for (int hi = 0; hi < n; ++hi) {
          const int e = std::abs(Third(expand[hi], expand[lo]) - v) + std::abs(expand[hi] - expand[lo]) * 3 / 100;
          if (e < best) {
            out[v][0] = static_cast<uint8_t>(hi);
            out[v][1] = static_cast<uint8_t>(lo);
            best = e;
          }
        }
    }
  }

  void MatchConstant(int r, int g, int b, uint16_t *hi, uint16_t *lo) const {
    *hi = static_cast<uint16_t>((best5_[r][0] << 11) | (best6_[g][0] << 5) | best5_[b][0]);
    *lo = static_cast<uint16_t>((best5_[r][1] << 11) | (best6_[g][1] << 5) | best5_[b][1]);
  }

  // stb_dxt.h:139-147: the four palette colours of endpoints (c0, c1).
  void Palette(uint16_t c0, uint16_t c1, int pal[4][3]) const {
    const uint16_t c[2] = {c0, c1};
    for (int k = 0; k < 2; ++k) {
      pal[k][0] = expand5_[(c[k] >> 11) & 31];
      pal[k][1] = expand6_[(c[k] >> 5) & 63];
      pal[k][2] = expand5_[c[k] & 31];
    }
    for (int ch = 0; ch < 3; ++ch) {
      pal[2][ch] = Third(pal[0][ch], pal[1][ch]);
      pal[3][ch] = Third(pal[1][ch], pal[0][ch]);
    }
  }

  // stb_dxt.h:176-211: project onto the endpoint axis and pick by the midpoints.
  uint32_t Match(const uint8_t blk[64], uint16_t c0, uint16_t c1) const {
    int pal[4][3];
    Palette(c0, c1, pal);
    const int dr = pal[0][0] - pal[1][0], dg = pal[0][1] - pal[1][1], db = pal[0][2] - pal[1][2];
    int stop[4];
    for (int k = 0; k < 4; ++k) stop[k] = pal[k][0] * dr + pal[k][1] * dg + pal[k][2] * db;
    const int to_c0 = (stop[1] + stop[3]) >> 1;
    const int half = (stop[3] + stop[2]) >> 1;
    const int to_c3 = (stop[2] + stop[0]) >> 1;
    uint32_t mask = 0;
    for (int i = 15; i >= 0; --i) {
      const int dot = blk[4 * i] * dr + blk[4 * i + 1] * dg + blk[4 * i + 2] * db;
      mask <<= 2;
      if (dot < half)
        mask |= dot < to_c0 ? 1u : 3u;
      else
        mask |= dot < to_c3 ? 2u : 0u;
    }
    return mask;
  }

  // stb_dxt.h:273-375: principal axis by four power iterations on the covariance
  // (float), then the pixels at the two ends of the axis become the endpoints.
  void PrincipalEndpoints(const uint8_t blk[64], uint16_t *hi, uint16_t *lo) const {
    int mean[3], mn[3], mx[3];
    for (int ch = 0; ch < 3; ++ch) {
      int sum = blk[ch], a = blk[ch], b = blk[ch];
      for (int i = 1; i < 16; ++i) {
        const int v = blk[4 * i + ch];
        sum += v;
        if (v < a)
          a = v;
        else if (v > b)
          b = v;
      }
      mean[ch] = (sum + 8) >> 4;
      mn[ch] = a;
      mx[ch] = b;
    }
    int cov[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 16; ++i) {
      const int r = blk[4 * i] - mean[0], g = blk[4 * i + 1] - mean[1], b = blk[4 * i + 2] - mean[2];
      cov[0] += r * r;
      cov[1] += r * g;
      cov[2] += r * b;
      cov[3] += g * g;
      cov[4] += g * b;
      cov[5] += b * b;
    }
    float c[6];
    for (int k = 0; k < 6; ++k) c[k] = static_cast<float>(cov[k]) / 255.0f;
    float vr = static_cast<float>(mx[0] - mn[0]);
    float vg = static_cast<float>(mx[1] - mn[1]);
    float vb = static_cast<float>(mx[2] - mn[2]);
    for (int it = 0; it < 4; ++it) {
      const float r = vr * c[0] + vg * c[1] + vb * c[2];
      const float g = vr * c[1] + vg * c[3] + vb * c[4];
      const float b = vr * c[2] + vg * c[4] + vb * c[5];
      vr = r;
      vg = g;
      vb = b;
    }
    double magn = std::fabs(vr);
    if (std::fabs(vg) > magn) magn = std::fabs(vg);
    if (std::fabs(vb) > magn) magn = std::fabs(vb);
    int ar, ag, ab;
    if (magn < 4.0f) {  // luma weights x 1000
      ar = 299;
      ag = 587;
      ab = 114;
    } else {
      magn = 512.0 / magn;
      ar = TruncToInt(vr * magn);
      ag = TruncToInt(vg * magn);
      ab = TruncToInt(vb * magn);
    }
    int dmin = INT_MAX, dmax = -INT_MAX, imin = 0, imax = 0;
    for (int i = 0; i < 16; ++i) {
      const int dot = blk[4 * i] * ar + blk[4 * i + 1] * ag + blk[4 * i + 2] * ab;
      if (dot < dmin) {
        dmin = dot;
        imin = i;
      }
      if (dot > dmax) {
        dmax = dot;
        imax = i;
      }
    }
    *hi = To565(blk[4 * imax], blk[4 * imax + 1], blk[4 * imax + 2]);
    *lo = To565(blk[4 * imin], blk[4 * imin + 1], blk[4 * imin + 2]);
  }

  static int Clamp(float y, int top) {  // stb__sclamp
    const int x = TruncToInt(y);
    return x < 0 ? 0 : (x > top ? top : x);
  }

  // stb_dxt.h:388-464: least-squares endpoints for the current indices (normal
  // equations, Cramer's rule).  Returns whether the endpoints changed.
  bool Refine(const uint8_t blk[64], uint16_t *hi, uint16_t *lo, uint32_t mask) const {
    static const int kWeightHi[4] = {3, 0, 2, 1};  // weight of the first endpoint, x3
    const uint16_t old_hi = *hi, old_lo = *lo;
    uint16_t nhi, nlo;
    if ((mask ^ (mask << 2)) < 4) {  // one index everywhere: singular, match the mean
      int r = 8, g = 8, b = 8;
      for (int i = 0; i < 16; ++i) {
        r += blk[4 * i];
        g += blk[4 * i + 1];
        b += blk[4 * i + 2];
      }
      MatchConstant(r >> 4, g >> 4, b >> 4, &nhi, &nlo);
    } else {
      int xx = 0, yy = 0, xy = 0;
      int s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
      uint32_t m = mask;
      for (int i = 0; i < 16; ++i, m >>= 2) {
        const int w1 = kWeightHi[m & 3], w2 = 3 - w1;
        xx += w1 * w1;
        yy += w2 * w2;
        xy += w1 * w2;
        for (int ch = 0; ch < 3; ++ch) {
          s1[ch] += w1 * blk[4 * i + ch];
          s2[ch] += blk[4 * i + ch];
        }
      }
      for (int ch = 0; ch < 3; ++ch) s2[ch] = 3 * s2[ch] - s1[ch];
      const float frb = 3.0f * 31.0f / 255.0f / static_cast<float>(xx * yy - xy * xy);
      const float fg = frb * 63.0f / 31.0f;
      nhi = static_cast<uint16_t>(Clamp(static_cast<float>(s1[0] * yy - s2[0] * xy) * frb + 0.5f, 31) << 11 |
                                  Clamp(static_cast<float>(s1[1] * yy - s2[1] * xy) * fg + 0.5f, 63) << 5 |
                                  Clamp(static_cast<float>(s1[2] * yy - s2[2] * xy) * frb + 0.5f, 31));
      nlo = static_cast<uint16_t>(Clamp(static_cast<float>(s2[0] * xx - s1[0] * xy) * frb + 0.5f, 31) << 11 |
                                  Clamp(static_cast<float>(s2[1] * xx - s1[1] * xy) * fg + 0.5f, 63) << 5 |
                                  Clamp(static_cast<float>(s2[2] * xx - s1[2] * xy) * frb + 0.5f, 31));
    }
    *hi = nhi;
    *lo = nlo;
    return old_hi != nhi || old_lo != nlo;
  }
};

// ---------------------------------------------------------------------------
// Reencode's block model (codec/dxt_image.cpp:38-133, 144-215, 262-360): a block as
// its two 8-bit endpoints, its four palette colours and its 16 indices.

struct Logical {
  uint8_t ep[2][4];
  uint8_t pal[4][4];
  uint8_t idx[16];
};

uint8_t AbsDiff(uint8_t a, uint8_t b) { return a > b ? a - b : b - a; }

void Lerp(const uint8_t *a, const uint8_t *b, uint8_t *out, int num, int div) {
  for (int c = 0; c < 3; ++c) out[c] = static_cast<uint8_t>((a[c] * (div - num) + b[c] * num) / div);
  out[3] = 0xFF;
}

void Expand565(uint16_t v, uint8_t out[4]) {
  const uint32_t r = v >> 11, g = (v >> 5) & 0x3F, b = v & 0x1F;
  out[0] = static_cast<uint8_t>((r << 3) | (r >> 2));
  out[1] = static_cast<uint8_t>((g << 2) | (g >> 4));
  out[2] = static_cast<uint8_t>((b << 3) | (b >> 2));
  out[3] = 0xFF;
}

uint16_t Pack565(const uint8_t c[4]) {
  return static_cast<uint16_t>(((c[0] & 0xF8) << 8) | ((c[1] & 0xFC) << 3) | ((c[2] & 0xF8) >> 3));
}

// codec/dxt_image.cpp:81-132: the 8-bit value a 5- (6-) bit channel snaps to.  The
// candidates are the truncated code and the values half a step above and below it,
// each with its top bits replicated; ties keep the lower candidate.
uint8_t Snap(uint8_t x, int bits) {
  const uint8_t keep = static_cast<uint8_t>(0xFF << (8 - bits));
  const uint8_t half = static_cast<uint8_t>(1 << (7 - bits));
  uint8_t mid = x & keep;
  uint8_t up = static_cast<uint8_t>(mid + half);
  uint8_t down = mid == 0 ? mid : static_cast<uint8_t>(mid - half);
  mid = static_cast<uint8_t>(mid | (mid >> bits));
  up = static_cast<uint8_t>(up | (up >> bits));
  down = static_cast<uint8_t>(down | (down >> bits));
  const uint8_t dm = AbsDiff(x, mid), du = AbsDiff(x, up), dd = AbsDiff(x, down);
  if (dm <= du) return dm < dd ? mid : down;
  return up;
}

Logical FromPhysical(uint64_t blk) {  // codec/dxt_image.cpp:144-173
  Logical l;
  const uint16_t e0 = static_cast<uint16_t>(blk), e1 = static_cast<uint16_t>(blk >> 16);
  Expand565(e0, l.ep[0]);
  Expand565(e1, l.ep[1]);
  std::memcpy(l.pal[0], l.ep[0], 4);
  std::memcpy(l.pal[1], l.ep[1], 4);
  if (e0 <= e1) {
    Lerp(l.ep[0], l.ep[1], l.pal[2], 1, 2);
    std::memset(l.pal[3], 0, 4);
  } else {
    Lerp(l.ep[0], l.ep[1], l.pal[2], 1, 3);
    Lerp(l.ep[0], l.ep[1], l.pal[3], 2, 3);
  }
  const uint32_t word = static_cast<uint32_t>(blk >> 32);
  for (int i = 0; i < 16; ++i) l.idx[i] = (word >> (2 * i)) & 3;
  return l;
}

uint64_t ToPhysical(const Logical &l) {  // codec/dxt_image.cpp:175-215
  uint16_t p0 = Pack565(l.ep[0]), p1 = Pack565(l.ep[1]);
  const bool swap = (p0 > p1 && l.pal[3][3] == 0) || (p0 <= p1 && l.pal[3][3] == 255);
  if (swap) std::swap(p0, p1);
  uint32_t word = 0;
  for (int i = 0; i < 16; ++i) word |= static_cast<uint32_t>(l.idx[i]) << (2 * i);
  if (swap) word ^= 0x55555555u;
  return static_cast<uint64_t>(p0) | static_cast<uint64_t>(p1) << 16 | static_cast<uint64_t>(word) << 32;
}

// codec/dxt_image.cpp:265-283: mean squared error over the 48 channel samples, floored.
int BlockError(const Logical &l, const uint8_t px[48]) {
  size_t err = 0;
  for (int i = 0; i < 16; ++i)
    for (int c = 0; c < 3; ++c) {
      const size_t d = AbsDiff(px[3 * i + c], l.pal[l.idx[i]][c]);
      err += d * d;
    }
  return static_cast<int>(err / 48);
}

void TakeIndices(Logical *l, uint32_t word) {  // codec/dxt_image.cpp:295-299
  const uint64_t p = ToPhysical(*l);
  *l = FromPhysical((p & 0xFFFFFFFFull) | static_cast<uint64_t>(word) << 32);
}

// codec/dxt_image.cpp:301-359: least-squares endpoints for the block's indices in
// float, in the reference's accumulation order, then snapped per channel.
void Refit(Logical *l, const uint8_t px[48]) {
  static const float kOrder[4] = {0.f, 3.f, 1.f, 2.f};  // index -> position on the line
  float aa = 0.0f, bb = 0.0f, ab = 0.0f;
  float ax[3] = {0.0f, 0.0f, 0.0f}, bx[3] = {0.0f, 0.0f, 0.0f};
  for (int i = 0; i < 16; ++i) {
    const float t = kOrder[l->idx[i]];
    const float a = (3.0f - t) / 3.0f, b = t / 3.0f;
    aa += a * a;
    bb += b * b;
    ab += a * b;
    for (int c = 0; c < 3; ++c) {
      ax[c] += static_cast<float>(px[3 * i + c]) * a;
      bx[c] += static_cast<float>(px[3 * i + c]) * b;
    }
  }
  const float f = 1.0f / (aa * bb - ab * ab);
  for (int c = 0; c < 3; ++c) {
    const float e0 = f * (ax[c] * bb - bx[c] * ab);
    const float e1 = f * (bx[c] * aa - ax[c] * ab);
    l->ep[0][c] = static_cast<uint8_t>(std::max(0, std::min(255, TruncToInt(e0 + 0.5f))));
    l->ep[1][c] = static_cast<uint8_t>(std::max(0, std::min(255, TruncToInt(e1 + 0.5f))));
  }
  for (int e = 0; e < 2; ++e) {
    l->ep[e][0] = Snap(l->ep[e][0], 5);
    l->ep[e][1] = Snap(l->ep[e][1], 6);
    l->ep[e][2] = Snap(l->ep[e][2], 5);
  }
  std::memcpy(l->pal[0], l->ep[0], 4);
  std::memcpy(l->pal[1], l->ep[1], 4);
  Lerp(l->ep[0], l->ep[1], l->pal[2], 1, 3);
  Lerp(l->ep[0], l->ep[1], l->pal[3], 2, 3);
}

void BlockPixels(const uint8_t *rgb, uint32_t width, uint32_t bx, uint32_t by, uint8_t px[48]) {
  for (int r = 0; r < 4; ++r)
    std::memcpy(px + 12 * r, rgb + 3 * ((static_cast<size_t>(by) * 4 + r) * width + bx * 4), 12);
}

// ---------------------------------------------------------------------------
// Entropy coding (codec/entropy.cpp:174-265, ans/encode.cpp:49-69,224-259).

// Returns a GSTE_* status.
int EncodeStream(const std::vector<uint8_t> &sym, std::vector<uint8_t> *table, std::vector<uint8_t> *body) {
  std::vector<uint32_t> counts(256, 0);
  for (uint8_t s : sym) counts[s]++;
  size_t used = 0;
  for (size_t i = 0; i < 256; ++i)
    if (counts[i]) used = i + 1;
  counts.resize(used);
  // ans/ans_utils.h:42-47, ans/encode.cpp:196-199: the coder needs more than one
  // frequency, so a stream of zero bytes only is refused.
  if (used <= 1) return GSTE_ERR_UNCODABLE;
  std::vector<uint32_t> freq;
  if (!gst::hist::GenerateHistogram(counts, static_cast<int>(kTableM), &freq)) return GSTE_ERR_UNCODABLE;
  uint32_t cum[256] = {0}, f[256] = {0};
  uint32_t run = 0;
  for (size_t i = 0; i < freq.size(); ++i) {
    f[i] = freq[i];
    cum[i] = run;
    run += freq[i];
  }
  if (run != kTableM) return GSTE_ERR_UNCODABLE;
  table->assign(512, 0);
  for (size_t i = 0; i < freq.size(); ++i) {
    (*table)[2 * i] = static_cast<uint8_t>(freq[i]);
    (*table)[2 * i + 1] = static_cast<uint8_t>(freq[i] >> 8);
  }
  const size_t groups = sym.size() / kGroupSyms;
  std::vector<uint8_t> data;
  std::vector<uint32_t> ends(groups);
  std::vector<uint16_t> words;
  for (size_t g = 0; g < groups; ++g) {
    const uint8_t *s = sym.data() + g * kGroupSyms;
    uint32_t x[kLanes];
    for (uint32_t j = 0; j < kLanes; ++j) x[j] = kStateL;
    words.clear();
    for (uint32_t i = 0; i < kLaneSyms; ++i)
      for (uint32_t j = 0; j < kLanes; ++j) {  // lane j codes s[j * 256 + i]
        const uint32_t v = s[j * kLaneSyms + i], fv = f[v];
        while (x[j] >= (fv << 20)) {  // b = 2^16, L / M = 2^4
          words.push_back(static_cast<uint16_t>(x[j]));
          x[j] >>= 16;
        }
        x[j] = (x[j] / fv) * kTableM + cum[v] + x[j] % fv;
      }
    if ((2 * words.size() + 4 * kLanes) & 3) data.insert(data.end(), 2, 0);  // entropy.cpp:208-221
    for (uint16_t w : words) {
      data.push_back(static_cast<uint8_t>(w));
      data.push_back(static_cast<uint8_t>(w >> 8));
    }
    for (uint32_t j = 0; j < kLanes; ++j)
      for (int b = 0; b < 4; ++b) data.push_back(static_cast<uint8_t>(x[j] >> (8 * b)));
    ends[g] = static_cast<uint32_t>(4 * groups + data.size());
  }
  body->clear();
  for (uint32_t e : ends)
    for (int b = 0; b < 4; ++b) body->push_back(static_cast<uint8_t>(e >> (8 * b)));
  body->insert(body->end(), data.begin(), data.end());
  return GSTE_OK;
}

int Mirror(int idx, int range) {  // codec/wavelet.cpp:11-15
  const int v = idx - static_cast<int>(idx >= range) * (idx - range + 2);
  return v < 0 ? -v : v;
}

// One forward 5/3 lifting pass over `len` values (codec/wavelet.cpp:29-58): odd
// samples become details, even ones the smoothed averages, averages first.
void Lift(const int *v, int stride, int len, int *out) {
  const int mid = len / 2;
  for (int i = 1; i < len; i += 2)
    out[mid + i / 2] = v[i * stride] - (v[(i - 1) * stride] + v[Mirror(i + 1, len) * stride]) / 2;
  for (int i = 0; i < len; i += 2)
    out[i / 2] = v[i * stride] + (out[mid + Mirror(i - 1, len) / 2] + out[mid + Mirror(i + 1, len) / 2] + 2) / 4;
}

// codec/image_processing.h:291-330 with codec/wavelet.cpp:99-134: per level, every
// column of the top-left dim x dim square, then every row.
void TileTransform(int t[kTile * kTile]) {
  int line[kTile];
  for (int dim = kTile; dim > 1; dim /= 2) {
    for (int x = 0; x < dim; ++x) {
      Lift(t + x, kTile, dim, line);
      for (int y = 0; y < dim; ++y) t[y * kTile + x] = line[y];
    }
    for (int y = 0; y < dim; ++y) {
      Lift(t + y * kTile, 1, dim, line);
      std::memcpy(t + y * kTile, line, sizeof(int) * dim);
    }
  }
}

}  // namespace

extern "C" {

int gste_compress_dxt(const uint8_t *rgb, uint32_t width, uint32_t height, const uint8_t *dxt_in,
                      uint8_t *out, size_t out_cap, size_t *out_size, uint8_t *dxt_out) {
  if (out_size) *out_size = 0;
  if (!rgb || !out_size || width == 0 || height == 0 || width % 128 || height % 128 ||
      (static_cast<uint64_t>(width) * height) % (16ull * kGroupSyms))
    return GSTE_ERR_ARGUMENT;
  const uint32_t bw = width / 4, bh = height / 4, nv = bw * bh;

  // 1. DXT1 blocks (codec/dxt_image.cpp:517-530).
  std::vector<uint64_t> phys(nv);
  if (dxt_in) {
    std::memcpy(phys.data(), dxt_in, 8 * static_cast<size_t>(nv));
  } else {
    const StbDxt &stb = StbDxt::Get();
    uint8_t blk[64];
    for (uint32_t b = 0; b < nv; ++b) {
      uint8_t px[48];
      BlockPixels(rgb, width, b % bw, b / bw, px);
      for (int i = 0; i < 16; ++i) {
        std::memcpy(blk + 4 * i, px + 3 * i, 3);
        blk[4 * i + 3] = 0xFF;
      }
      phys[b] = stb.Compress(blk);
    }
  }

  // 2. Reencode (codec/dxt_image.cpp:538-620).
  std::vector<uint32_t> palette;
  std::vector<uint8_t> steps(nv);
  int last = 0;
  for (uint32_t b = 0; b < nv; ++b) {
    uint8_t px[48];
    BlockPixels(rgb, width, b % bw, b / bw, px);
    const Logical orig = FromPhysical(phys[b]);
    const int orig_err = BlockError(orig, px);
    int best = INT_MAX;
    size_t best_k = 0;
    const size_t look = std::min(kLookback, palette.size());
    for (size_t k = 0; k < look; ++k) {
      const uint32_t word = palette[palette.size() - 1 - k];
      Logical cand = orig;
      TakeIndices(&cand, word);
      Refit(&cand, px);
      // Only refits whose endpoints keep the 4-colour order keep the word as is.
      if (static_cast<uint32_t>(ToPhysical(cand) >> 32) != word || cand.pal[3][3] != 0xFF) continue;
      const int diff = BlockError(cand, px) - orig_err;
      if (diff < best) {
        best = diff;
        best_k = k;
        if (diff <= 0) break;
      }
    }
    int index;
    if (best < kErrThreshold) {
      Logical l = orig;
      TakeIndices(&l, palette[palette.size() - 1 - best_k]);
      Refit(&l, px);
      phys[b] = ToPhysical(l);
      index = static_cast<int>(palette.size() - best_k - 1);
    } else {
      index = static_cast<int>(palette.size());
      palette.push_back(static_cast<uint32_t>(phys[b] >> 32));
    }
    const int step = index - last;
    if (step < -128 || step > 127) return GSTE_ERR_RANGE;
    steps[b] = static_cast<uint8_t>(step + 128);
    last = index;
  }

  // 3. Endpoint planes Y1 Y2 | Co1 Cg1 Co2 Cg2 (codec/encoder.cpp:39-98,
  // codec/image_processing.cpp:10-26).
  std::vector<uint8_t> ysym(2 * static_cast<size_t>(nv)), csym(4 * static_cast<size_t>(nv));
  std::vector<int> plane(nv);
  const uint32_t tx_n = bw / kTile, ty_n = bh / kTile;
  for (int p = 0; p < 6; ++p) {
    const int e = p < 2 ? p : (p - 2) / 2;                // endpoint
    const int comp = p < 2 ? 0 : 1 + (p - 2) % 2;         // 0 Y, 1 Co, 2 Cg
    for (uint32_t b = 0; b < nv; ++b) {
      const uint16_t c = static_cast<uint16_t>(phys[b] >> (16 * e));
      const int r = c >> 11, g = (c >> 5) & 63, bl = c & 31;
      const int t = r + bl + (bl >> 4), cg = g - t;
      plane[b] = comp == 0 ? t + cg / 2 : (comp == 1 ? r - bl : cg);
    }
    uint8_t *dst = p < 2 ? ysym.data() + static_cast<size_t>(p) * nv : csym.data() + static_cast<size_t>(p - 2) * nv;
    for (uint32_t ty = 0; ty < ty_n; ++ty)
      for (uint32_t tx = 0; tx < tx_n; ++tx) {
        int t[kTile * kTile];
        for (int y = 0; y < kTile; ++y)
          for (int x = 0; x < kTile; ++x) t[y * kTile + x] = plane[(ty * kTile + y) * bw + tx * kTile + x];
        TileTransform(t);
        // int8 cast, then +128 (codec/image_processing.h:322-326, MakeUnsigned): mod 256.
        uint8_t *o = dst + static_cast<size_t>(ty * tx_n + tx) * kTile * kTile;
        for (int i = 0; i < kTile * kTile; ++i) o[i] = static_cast<uint8_t>(t[i] + 128);
      }
  }

  // 4. Streams and the file (codec/encoder.cpp:100-144).
  const size_t pal_bytes = (4 * palette.size() + kGroupSyms - 1) / kGroupSyms * kGroupSyms;
  std::vector<uint8_t> psym(pal_bytes, 0);
  for (size_t i = 0; i < palette.size(); ++i)
    for (int b = 0; b < 4; ++b) psym[4 * i + b] = static_cast<uint8_t>(palette[i] >> (8 * b));
  const std::vector<uint8_t> *sym[4] = {&ysym, &csym, &psym, &steps};
  std::vector<uint8_t> table[4], body[4];
  size_t total = 28;
  for (int k = 0; k < 4; ++k) {
    const int rc = EncodeStream(*sym[k], &table[k], &body[k]);
    if (rc != GSTE_OK) return rc;
    total += 512 + body[k].size();
  }
  *out_size = total;
  if (!out || out_cap < total) return GSTE_ERR_CAPACITY;
  const uint32_t hdr[7] = {width, height, static_cast<uint32_t>(pal_bytes),
                           static_cast<uint32_t>(body[0].size()), static_cast<uint32_t>(body[1].size()),
                           static_cast<uint32_t>(body[2].size()), static_cast<uint32_t>(body[3].size())};
  uint8_t *o = out;
  for (uint32_t v : hdr)
    for (int b = 0; b < 4; ++b) *o++ = static_cast<uint8_t>(v >> (8 * b));
  for (int k = 0; k < 4; ++k) o = std::copy(table[k].begin(), table[k].end(), o);
  for (int k = 0; k < 4; ++k) o = std::copy(body[k].begin(), body[k].end(), o);
  if (dxt_out)
    for (uint32_t b = 0; b < nv; ++b)
      for (int k = 0; k < 8; ++k) dxt_out[8 * static_cast<size_t>(b) + k] = static_cast<uint8_t>(phys[b] >> (8 * k));
  return GSTE_OK;
}

}  // extern "C"
