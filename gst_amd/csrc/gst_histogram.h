// gst_histogram.h -- frequency normalisation of the reference's rANS coder
// (ans/histogram.cpp:40-121, GenerateHistogram), shared by the decoder library
// (table rebuilds of the stand-alone rANS decoder) and the encoder library (the
// frequency tables it writes).  Header-only, host code.
#ifndef GST_HISTOGRAM_H_
#define GST_HISTOGRAM_H_

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <vector>

namespace gst {
namespace hist {

// ans/histogram.cpp:11-19,27-31: a symbol's cost of moving its frequency one step in
// the correction's direction, count * log2(f / (f + sign)); the smallest goes first.
struct Candidate {
  int symbol;
  double rank;
  bool operator>(const Candidate &o) const { return rank > o.rank; }
};

inline double StepCost(uint32_t count, int f, int sign) {
  return std::log2(static_cast<double>(f) / static_cast<double>(f + sign)) * static_cast<double>(count);
}

inline int SumAsInt(const std::vector<uint32_t> &v) {
  uint32_t s = 0;
  for (uint32_t x : v) s += x;
  return static_cast<int>(s);
}

// ans/histogram.cpp:40-121 step by step, keeping its float/double mix and its 32-bit
// count * M product: each count scales to count * M / sum in float, rounds to the
// nearer of floor and floor + 1 by the geometric-mean test (at least 1), and the
// difference to M is then moved one unit at a time onto the symbols whose rank is
// smallest, through a binary heap kept by the same std::*_heap calls so that ties
// resolve as the reference's do.  False when M <= 0, every count is zero, or the
// correction cannot be placed.
inline bool GenerateHistogram(const std::vector<uint32_t> &counts, int M, std::vector<uint32_t> *out) {
  if (M <= 0) return false;
  std::vector<uint32_t> h;
  h.reserve(counts.size());
  const int sum = SumAsInt(counts);
  for (uint32_t c : counts) {
    if (c == 0) {
      h.push_back(0);
      continue;
    }
    const double scaled = static_cast<float>(c * static_cast<uint32_t>(M)) / static_cast<float>(sum);
    const int down = static_cast<int>(scaled);
    h.push_back(static_cast<uint32_t>(std::max(1, (scaled * scaled <= down * (down + 1)) ? down : down + 1)));
  }
  int correction = M - SumAsInt(h);
  if (correction == M) return false;  // every count was zero
  if (correction != 0) {
    const int sign = (correction >> 31) | static_cast<int>(correction > 0);
    std::vector<Candidate> heap;
    heap.reserve(counts.size());
    for (size_t i = 0; i < counts.size(); ++i) {
      if (counts[i] == 0) continue;
      if (h[i] > 1 || correction > 0)
        heap.push_back({static_cast<int>(i), StepCost(counts[i], static_cast<int>(h[i]), sign)});
    }
    std::make_heap(heap.begin(), heap.end(), std::greater<Candidate>());
    while (correction != 0) {
      if (heap.empty()) return false;
      std::pop_heap(heap.begin(), heap.end(), std::greater<Candidate>());
      const int i = heap.back().symbol;
      heap.pop_back();
      h[i] = static_cast<uint32_t>(static_cast<int>(h[i]) + sign);
      correction -= sign;
      if (h[i] > 1 || sign == 1) {
        heap.push_back({i, StepCost(counts[i], static_cast<int>(h[i]), sign)});
        std::push_heap(heap.begin(), heap.end(), std::greater<Candidate>());
      }
    }
  }
  *out = std::move(h);
  return true;
}

}  // namespace hist
}  // namespace gst

#endif  // GST_HISTOGRAM_H_
