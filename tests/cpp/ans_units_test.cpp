// Host-only check of the rANS work-unit enumeration (gst_internal.h): for every
// combination of per-stream group counts in a range, the units of an image cover every
// group of every stream exactly once, no unit holds more groups than its work-group
// decodes, a paired unit holds two short remainders (each fits half the waves), and
// without pairing the count is the plain ceil(G / per) per stream.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gst_internal.h"

namespace {

struct Groups {
  uint32_t groups[4];
};

int Check(const Groups &d, uint32_t code) {
  const uint32_t per = code & 0xFFu;
  const bool pairing = (code & gst::kAnsPairShort) != 0;
  std::vector<int> seen[4];
  for (uint32_t k = 0; k < 4; ++k) seen[k].assign(d.groups[k], 0);
  const uint32_t n = gst::AnsUnitsOf(d, code);
  uint32_t plain = 0;
  for (uint32_t k = 0; k < 4; ++k) plain += (d.groups[k] + per - 1) / per;
  if (!pairing && n != plain) return 1;
  if (pairing && n > plain) return 2;
  for (uint32_t u = 0; u < n; ++u) {
    const gst::AnsUnit a = gst::AnsUnitAt(d, code, u);
    if (a.pair && !pairing) return 3;
    for (uint32_t s = 0; s < (a.pair ? 2u : 1u); ++s) {
      if (a.count[s] == 0) {
        if (s == 0) return 4;  // a unit always has a first segment
        continue;
      }
      if (a.k[s] > 3 || a.count[s] > (a.pair ? per / 2 : per)) return 5;
      if (a.pair && !gst::AnsShortRemainder(d.groups[a.k[s]], code)) return 6;
      for (uint32_t g = a.first[s]; g < a.first[s] + a.count[s]; ++g) {
        if (g >= d.groups[a.k[s]]) return 7;
        ++seen[a.k[s]][g];
      }
    }
    if (!a.pair && a.count[1] != 0) return 8;
  }
  for (uint32_t k = 0; k < 4; ++k)
    for (int c : seen[k])
      if (c != 1) return 9;
  return 0;
}

}  // namespace

int main() {
  const uint32_t codes[] = {8, 8 | gst::kAnsPairShort, 4, 16, 16 | gst::kAnsPairShort, 2, 1};
  const uint32_t kMax = 21;
  unsigned long checked = 0;
  for (uint32_t code : codes) {
    Groups d;
    for (d.groups[0] = 0; d.groups[0] < kMax; ++d.groups[0])
      for (d.groups[1] = 0; d.groups[1] < kMax; ++d.groups[1])
        for (d.groups[2] = 0; d.groups[2] < kMax; ++d.groups[2])
          for (d.groups[3] = 0; d.groups[3] < kMax; ++d.groups[3]) {
            const int rc = Check(d, code);
            ++checked;
            if (rc) {
              std::printf("FAIL code %#x groups %u %u %u %u: check %d\n", code, d.groups[0], d.groups[1],
                          d.groups[2], d.groups[3], rc);
              return 1;
            }
          }
  }
  // The C3 corpus's streams: 512x512 .. 2048x2048 images, palettes of 1-5 groups.
  Groups big = {{64, 128, 5, 32}};
  if (Check(big, 8 | gst::kAnsPairShort) || gst::AnsUnitsOf(big, 8 | gst::kAnsPairShort) != 8 + 16 + 1 + 4) {
    std::printf("FAIL 2048x2048 image\n");
    return 1;
  }
  Groups small = {{4, 8, 1, 2}};  // 512x512: Y 4, chroma 8, palette 1, indices 2 groups
  if (Check(small, 8 | gst::kAnsPairShort) || gst::AnsUnitsOf(small, 8 | gst::kAnsPairShort) != 1 + 2) {
    std::printf("FAIL 512x512 image\n");
    return 1;
  }
  std::printf("ans units: %lu group-count combinations checked\n", checked);
  return 0;
}
