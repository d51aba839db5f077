// gst_kernels.hip -- the five GST decode stages as gfx950 kernels (wave64, integer only).
//
// Every kernel is a grid-stride loop over "units" of its stage; a unit maps to
// (image, local unit) through the batch plan's per-stage prefix array, or to the
// single image passed by value.  Reference semantics are cited per kernel.
#include "gst_internal.h"

#include <mutex>

namespace gst {
namespace {

constexpr int32_t kBufferRsrcWord3 = 0x00020000;  // raw buffer, gfx9 family

// Pointers in the global address space.  Loads and stores through them compile to
// global_* instructions, which count in vmcnt only; through generic pointers they
// would be flat_* instructions, which also count in lgkmcnt, so every wait for an
// LDS access would wait for them as well (a sector store would stall the next
// rANS step until the store completes).  The host compilation pass only type-checks
// this code and has no address spaces for class-type copies, so there the aliases
// are plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using GPtr = __attribute__((address_space(1))) T *;
template <typename T>
using CPtr = __attribute__((address_space(4))) const T *;
#else
template <typename T>
using GPtr = T *;
template <typename T>
using CPtr = const T *;
#endif

template <typename T, typename U>
__device__ __forceinline__ GPtr<T> Gbl(U *p) {
  return (GPtr<T>)(reinterpret_cast<uintptr_t>(p));
}

// Element k of a 4-entry array without dynamic indexing (which would put the array
// in scratch memory).
template <typename T>
__device__ __forceinline__ T Pick4(const T (&v)[4], uint32_t k) {
  return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : v[3];
}

__device__ __forceinline__ uint32_t Popc64(uint64_t m) { return static_cast<uint32_t>(__popcll(m)); }

struct Located {
  uint32_t img;
  uint32_t local;
  bool valid;  // false: u is past the stage's last unit
};

// Work unit u of `stage` as (image, unit within the image).  Called by a whole wave
// with u uniform.  The image is the last index i with prefix[i] <= u, found by
// counting over a ballot instead of a dependent binary search: for plans of up to 255
// images the wave loads all prefix entries in one round trip (a 64-way narrowing
// per round trip beyond).  Padding past n_images compares above every u, so a u past
// the last unit lands on i = n_images, whose successor entry is padding (or absent
// when n_images = cap): the unit is invalid.  The image count itself is not needed.
__device__ __forceinline__ Located Locate(const KernelArgs &a, uint32_t stage, uint32_t u,
                                          uint32_t groups_per_unit) {
  if (!a.plan) return {0u, u, u < StageUnitsOf(a.single, stage, groups_per_unit)};
  const uint32_t len = a.plan_cap + 1;
  const GPtr<const uint32_t> pre =
      Gbl<const uint32_t>(reinterpret_cast<const uint8_t *>(a.plan) + PlanPrefixOffset(a.plan_cap, stage));
  const uint32_t lane = __lane_id();
  uint32_t lo, start, next = kPrefixPad;
  if (len <= 256) {
    uint32_t v[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) v[k] = lane + 64 * k < len ? pre[lane + 64 * k] : kPrefixPad;
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) cnt += Popc64(__builtin_amdgcn_ballot_w64(v[k] <= u));
    lo = cnt - 1;  // prefix[0] = 0 <= u
    start = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(Pick4(v, lo >> 6)),
                                                            static_cast<int>(lo & 63)));
    if (lo + 1 < len)
      next = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(Pick4(v, (lo + 1) >> 6)),
                                                             static_cast<int>((lo + 1) & 63)));
  } else {
    lo = 0;
    uint32_t m = len;  // the answer lies in [lo, lo + m)
    while (m > 64) {
      const uint32_t step = (m + 63) / 64;
      const uint32_t v = lane * step < m ? pre[lo + lane * step] : kPrefixPad;
      const uint32_t c = Popc64(__builtin_amdgcn_ballot_w64(v <= u));  // >= 1: lane 0 reads prefix[lo]
      lo += (c - 1) * step;
      m = min(step, m - (c - 1) * step);
    }
    const uint32_t v = lane < m ? pre[lo + lane] : kPrefixPad;
    const uint32_t c = Popc64(__builtin_amdgcn_ballot_w64(v <= u));
    start = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(c - 1)));
    lo += c - 1;
    if (lo + 1 < len) next = pre[lo + 1];
  }
  return {lo, u - start, next != kPrefixPad};
}

// An image descriptor as the kernels use it: the same fields with global pointers.
struct DevDesc {
  uint32_t width, height, nv, bx;
  uint32_t palette_bytes;
  uint32_t groups[4];
  uint32_t stream_len[4];
  uint32_t table_base;
  const uint8_t *freqs;
  const uint8_t *streams[4];  // read through buffer resources
  GPtr<uint8_t> symbols;
  GPtr<int8_t> planes;
  GPtr<int32_t> indices;
  GPtr<int32_t> group_sums;
  GPtr<int32_t> lane_sums;  // may be null
  GPtr<uint8_t> out;
};

// The descriptor of image img: scalar loads (the plan is read-only while kernels run).
__device__ __forceinline__ DevDesc LoadDesc(const KernelArgs &a, uint32_t img) {
  ImageDesc s;
  if (a.plan) {
    const CPtr<ImageDesc> p = (CPtr<ImageDesc>)(reinterpret_cast<uintptr_t>(a.plan) + PlanImagesOffset(a.plan_cap));
    s = p[img];
  } else {
    s = a.single;
  }
  DevDesc d;
  d.width = s.width;
  d.height = s.height;
  d.nv = s.nv;
  d.bx = s.bx;
  d.palette_bytes = s.palette_bytes;
  for (int k = 0; k < 4; ++k) {
    d.groups[k] = s.groups[k];
    d.stream_len[k] = s.stream_len[k];
    d.streams[k] = s.streams[k];
  }
  d.table_base = s.table_base;
  d.freqs = s.freqs;
  d.symbols = Gbl<uint8_t>(s.symbols);
  d.planes = Gbl<int8_t>(s.planes);
  d.indices = Gbl<int32_t>(s.indices);
  d.group_sums = Gbl<int32_t>(s.group_sums);
  d.lane_sums = Gbl<int32_t>(s.lane_sums);
  d.out = Gbl<uint8_t>(s.out);
  return d;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t MakeRsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, static_cast<int>(bytes),
                                           kBufferRsrcWord3);
}

template <typename T>
__device__ __forceinline__ T WaveSumWidth32(T v) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 32);
  return v;
}

// ---------------------------------------------------------------------------
// Stage 1: rANS slot tables (ans/build_table.cl:12-83).
//
// One work-group per table: a 256-entry scan of the frequencies, then every slot s
// takes the symbol x with cum[x] <= s < cum[x] + f[x].  The reference finds x with an
// 11-step branch-free search; for tables that sum to M its result is that closed
// form (the search path depends only on the target symbol, checked exhaustively
// against the oracle).  Entry packing, 4 bytes:
//   bits 0-11 f[x] (1..2048), bits 12-22 s - cum[x], bits 24-31 x.
// A table that does not sum to 2048 is flagged in table_status and replaced by the
// identity table (f = 2048, offset = slot, symbol 0), which never renormalises.
__global__ __launch_bounds__(256) void BuildTableKernel(KernelArgs a) {
  __shared__ uint32_t cum[256];
  __shared__ uint32_t freq[256];
  __shared__ uint32_t wsum[4];
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t total = a.plan ? a.plan->n_images * 4 : 4;
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    const uint32_t img = u >> 2, k = u & 3;
    const DevDesc d = LoadDesc(a, img);
    const uint32_t tix = d.table_base + k;
    const GPtr<const uint16_t> fr = Gbl<const uint16_t>(d.freqs + 512 * k);
    const uint32_t f = fr[t];
    uint32_t v = f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t n = __shfl_up(v, off);
      if (lane >= static_cast<uint32_t>(off)) v += n;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += wsum[w];
    cum[t] = v + add - f;
    freq[t] = f;
    const uint32_t sum = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    const bool valid = sum == kTableSize;
    const GPtr<uint32_t> tab = Gbl<uint32_t>(a.tables + static_cast<size_t>(tix) * kTableSize);
    for (uint32_t slot = t; slot < kTableSize; slot += 256) {
      uint32_t e;
      if (valid) {
        uint32_t x;
        if (cum[255] <= slot) {
          x = 255;
        } else {
          uint32_t pos = 0;  // number of cum[0..254] <= slot
#pragma unroll
          for (uint32_t s = 128; s; s >>= 1)
            if (cum[pos + s - 1] <= slot) pos += s;
          x = pos - 1;
        }
        e = freq[x] | ((slot - cum[x]) << 12) | (x << 24);
      } else {
        e = kTableSize | (slot << 12);
      }
      tab[slot] = e;
    }
    if (t == 0) Gbl<uint32_t>(a.table_status)[tix] = valid ? 1u : 0u;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stage 2: interleaved rANS decode (ans/ans_decode.cl:25-74,97-143).
//
// A work-group of WAVES wave64s decodes GPW*WAVES consecutive groups of one stream
// with that stream's table resident in LDS in the 4-byte packed form BuildTableKernel
// writes (one random LDS read per step: random table reads are where the LDS bank
// conflicts come from; 8 KiB per table, so a 4-wave work-group with its rings takes
// 16 KiB).  The state update is one v_mad_u32_u24 after two bit-field extracts.  With
// GPW = 2 a wave carries two 32-stream groups, one per half-wave; each lane holds
// one stream's state.  The reference's local-memory renormalisation mask
// (atomic_or / popcount / atomic_and between barriers) becomes one ballot: a lane
// that needs a word reads u16[p_half - total_in_half + rank_in_half], the rank
// coming from mbcnt over the 64-bit ballot, and each half-wave's pointer (a scalar)
// moves down by its total.
//
// The words are read from an LDS ring of two 256-word chunks per half-wave, so the
// serial chain of a step only touches LDS.  Chunk c covers stream words
// [256c, 256c + 256) and lives at ring slot c & 1.  With `top` the chunk holding the
// half's pointer - 1, chunks top and top-1 are resident and chunks top-2 and top-3
// are in flight in registers.  A step moves the pointer by at most 32 words, so 8
// steps read only chunks top and top-1; every 8 steps, if the pointer has dropped
// below 256*top, chunk top-2 is written over top's slot, top-3 moves up and top-4 is
// requested, which then has at least 16 steps to arrive.  Stream loads go through a
// buffer resource bounded by the stream length (corrupt offsets read zeros, never
// fault); a pointer that ends below the group's first byte raises kErrStreamOffsets.
// Each lane packs 64 symbols into 16 registers and writes them as one 64-byte sector,
// in the reference's output order (lane j, step i of the group at j*256 + 255 - i).
// Index-stream groups also emit
// the sum of (symbol - 128) per group and per lane, the totals the index scan
// starts from.
// LDS written by some lanes of a wave and read by others: the wave's LDS operations
// complete in order, so only the compiler needs a fence.
__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kChunkWords = 256;
constexpr uint32_t kRingWords = 2 * kChunkWords;
constexpr uint32_t kRefillSteps = kChunkWords / kLanesPerGroup;  // 8

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Lane l's 8 bytes of chunk c (words 4l .. 4l+3): one wave moves a whole chunk, so a
// refill is never divergent.  A chunk wholly below the group's first chunk `clo` (the
// ring asks up to three chunks past the group's start near its end) reads nothing: its
// offset is moved past the stream (c * 512 wraps to 0xFFFFFE00), as are chunks below
// the stream start, which wrap by themselves.
constexpr uint32_t kNoChunk = 0xFFFFFFFFu >> 9;
__device__ __forceinline__ u32x2 LoadChunk(__amdgpu_buffer_rsrc_t rs, uint32_t c, uint32_t clo, uint32_t l) {
  c = c >= clo ? c : kNoChunk;
  return __builtin_amdgcn_raw_buffer_load_b64(rs, static_cast<int>((c * kChunkWords + 4 * l) * 2), 0, 0);
}

__device__ __forceinline__ void StoreChunk(uint16_t *ring, uint32_t c, uint32_t l, const u32x2 &a) {
  *reinterpret_cast<u32x2 *>(ring + ((c * kChunkWords + 4 * l) & (kRingWords - 1))) = a;
}

// The first two chunks of each half's ring (top and top-1, top = the chunk holding
// the half's word pointer - 1), requested before the table is built and written to
// LDS after it.
struct AnsRing {
  u32x2 rA0, rA1, rB0, rB1;
  uint32_t topA, topB;
};

template <uint32_t GPW>
__device__ __forceinline__ AnsRing RequestRing(__amdgpu_buffer_rsrc_t rs, uint32_t pA, uint32_t pB,
                                               uint32_t loA, uint32_t loB, uint32_t lane) {
  AnsRing r;
  r.topA = pA ? (pA - 1) / kChunkWords : 0;
  r.topB = pB ? (pB - 1) / kChunkWords : 0;
  const uint32_t cloA = loA / kChunkWords, cloB = loB / kChunkWords;
  r.rA0 = LoadChunk(rs, r.topA, cloA, lane);
  r.rA1 = LoadChunk(rs, r.topA - 1, cloA, lane);
  r.rB0 = r.rB1 = u32x2{0u, 0u};
  if (GPW == 2) {
    r.rB0 = LoadChunk(rs, r.topB, cloB, lane);
    r.rB1 = LoadChunk(rs, r.topB - 1, cloB, lane);
  }
  return r;
}

// Stream-tail touch.  Every refill waits for all loads in flight, so a refill whose
// chunk is still on its way from HBM stalls the wave for the whole round trip, a dozen
// times per group when the compressed streams are not in any cache (the usual case in a
// decode pipeline, whose other kernel moves a few hundred MB between two rANS passes).
// One pass of 4-byte LDS-DMA loads, a lane per 64 bytes, pulls the rest of the half's
// stream (bytes [lo, 512 (top - 3)), below the four chunks already requested) into L2
// and the memory-side cache at the start; the refills then find their chunks there.
// The loads write a 256-byte LDS sink that nothing reads and hold no VGPR; offsets
// past the stream (and halves with nothing left) read nothing through the buffer's
// range check.  A group's stream is at most 2 * kGroupSyms bytes, so at most
// kTouchPasses passes of 32 lanes x 64 bytes.
constexpr uint32_t kTouchStride = 64;
constexpr uint32_t kTouchPasses = 2 * kGroupSyms / (kLanesPerGroup * kTouchStride);

__device__ __forceinline__ void TouchStreamTail(__amdgpu_buffer_rsrc_t rs, uint32_t top, uint32_t lo_byte,
                                                uint32_t j, uint32_t *sink) {
  const uint32_t hi = top >= 3 ? 2 * kChunkWords * (top - 3) : 0u;
  const uint32_t span = kLanesPerGroup * kTouchStride;
  const uint32_t need = hi > lo_byte ? (hi - lo_byte + span - 1) / span : 0u;
  const uint32_t n0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(need, 0));
  const uint32_t n1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(need, 32));
  const uint32_t n = min(n0 > n1 ? n0 : n1, kTouchPasses);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t back = kTouchStride * (j + 1) + span * i;
    const bool in = hi >= back && hi - back + kTouchStride > lo_byte;
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)sink, 4,
                                             static_cast<int>(in ? hi - back : 0x7FFFFFF0u), 0, 0, 0);
  }
}

template <bool kSum, uint32_t GPW>
__device__ __forceinline__ void DecodeHalfGroups(const uint32_t *tab, uint16_t *ring_wave, uint16_t *ring_half,
                                                 uint32_t *sink, __amdgpu_buffer_rsrc_t rs, AnsRing ring,
                                                 uint32_t lane,
                                                 bool active, uint32_t x, uint32_t pA, uint32_t pB,
                                                 uint32_t loA, uint32_t loB, GPtr<uint8_t> outp,
                                                 GPtr<int32_t> group_sum, GPtr<int32_t> lane_sum,
                                                 GPtr<int32_t> lane_suffix, uint32_t *err) {
  const uint32_t j = lane & 31;
  uint32_t upper = lane < 32 ? 0u : 0xFFFFFFFFu;
  asm volatile("" : "+v"(upper));  // keep the half-wave mask an opaque VGPR (one v_and below)
  // Renormalisation threshold per lane: L for a lane with a stream, 0 (never) for one
  // without, so the compare's mask is already the renormalising lanes and neither the
  // ballot nor the select needs a scalar AND with the active lanes.
  uint32_t thr = active ? kStateL : 0u;
  asm volatile("" : "+v"(thr));
  uint16_t *const ringA = ring_wave;
  uint16_t *const ringB = ring_wave + kRingWords;
  uint32_t topA = ring.topA, topB = ring.topB;
  const uint32_t startA = pA, startB = pB;  // a pointer above its start has wrapped below 0
  // Chunks top-2 and top-3 go in flight in two fixed registers by chunk parity (even
  // chunks in pXE, odd in pXO): a load lands in the register its chunk is later
  // stored from, so no register moves sit between in-flight loads.  The wait before a
  // refill's store is still vmcnt(0) (the number of refill loads between two crossings
  // is data-dependent and output stores are in flight), which TouchStreamTail makes
  // cheap.
  const uint32_t cloA = loA / kChunkWords, cloB = loB / kChunkWords;
  u32x2 pAE = LoadChunk(rs, (topA & 1u) ? topA - 3 : topA - 2, cloA, lane);
  u32x2 pAO = LoadChunk(rs, (topA & 1u) ? topA - 2 : topA - 3, cloA, lane);
  u32x2 pBE = {0u, 0u}, pBO = {0u, 0u};
  if (GPW == 2) {
    pBE = LoadChunk(rs, (topB & 1u) ? topB - 3 : topB - 2, cloB, lane);
    pBO = LoadChunk(rs, (topB & 1u) ? topB - 2 : topB - 3, cloB, lane);
  }
  TouchStreamTail(rs, lane < 32 ? topA : topB, 2u * (lane < 32 ? loA : loB), j, sink);
  StoreChunk(ringA, topA, lane, ring.rA0);
  StoreChunk(ringA, topA - 1, lane, ring.rA1);
  if (GPW == 2) {
    StoreChunk(ringB, topB, lane, ring.rB0);
    StoreChunk(ringB, topB - 1, lane, ring.rB1);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  int32_t sum = 0;
  // 64 steps fill one 64-byte sector of the lane's output; whole sectors leave in four
  // back-to-back 16-byte stores (smaller pieces of a lane's segment, 256 B apart
  // across lanes, are evicted half-written and written back several times).
  for (uint32_t sector = 0; sector < kSymsPerLane / 64; ++sector) {
    // A SIMD's waves compete for issue slots and the oldest win, so young waves
    // would be left with most of their steps for the end, when too few waves remain
    // to keep the SIMD busy.  Priority falls as a wave gets further ahead, which
    // keeps the waves of a SIMD within about a sector of each other.
    switch (sector) {
      case 0: __builtin_amdgcn_s_setprio(3); break;
      case 1: __builtin_amdgcn_s_setprio(2); break;
      case 2: __builtin_amdgcn_s_setprio(1); break;
      default: __builtin_amdgcn_s_setprio(0); break;
    }
    uint32_t acc[16];
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) acc[i] = 0;
#pragma unroll
    for (uint32_t blk = 0; blk < 64 / kRefillSteps; ++blk) {
#pragma unroll
      for (uint32_t s = 0; s < kRefillSteps; ++s) {
        // One LDS read per step: e = f | (slot - cum) << 12 | symbol << 24.
        const uint32_t fs = tab[x & (kTableSize - 1)];
        x = __umul24(x >> kTableLog, fs & 0xFFFu) + __builtin_amdgcn_ubfe(fs, 12, 11);
        const bool need = x < thr;
        const uint64_t m = __builtin_amdgcn_uicmp(x, thr, 36 /* ULT */);
        const uint32_t mlo = static_cast<uint32_t>(m), mhi = static_cast<uint32_t>(m >> 32);
        const uint32_t tlo = __popc(mlo), thi = __popc(mhi);
        pA -= tlo;
        pB -= thi;
        {
          // Word index p_half + rank: mbcnt over both halves counts the lower half's
          // bits for upper lanes, so upper lanes start from pB - tlo.  The rank needs
          // only the ballot, so it runs beside the scalar pointer updates; one
          // three-operand add joins them.  Every lane reads (a lane that keeps its
          // state reads a word of the ring it never uses) and a select on the ballot
          // keeps the new state: no exec-mask round trip through the scalar unit.
          const uint32_t rank = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u));
          const uint32_t widx = rank + pA + (upper & (pB - tlo - pA));
          const uint32_t xn = (x << 16) | ring_half[widx & (kRingWords - 1)];
          x = need ? xn : x;
        }
        // Step 8*blk + s lands at byte 63 - 8*blk - s of the sector.  The symbol (byte
        // 3 of fs) goes in with v_perm: selector codes 4-7 take bytes of the first
        // operand, 0-3 bytes of the second.
        const uint32_t pos = 63 - kRefillSteps * blk - s;
        const uint32_t sel = (0x03020100u & ~(0xFFu << (8 * (pos & 3)))) | (7u << (8 * (pos & 3)));
        acc[pos >> 2] = __builtin_amdgcn_perm(fs, acc[pos >> 2], sel);
        asm volatile("" : "+v"(acc[pos >> 2]));  // insert now: keeps 64 symbols out of VGPRs
      }
      // Ring refill at the 8-step boundary: the pointer moved at most 256 words, so it
      // crossed at most one chunk boundary, and the next 8 steps read only chunks
      // top and top-1 (top = chunk of pointer - 1).
      const bool crossA = pA < topA * kChunkWords, crossB = pB < topB * kChunkWords;
      // Per crossed half (uniform branches, all 64 lanes): store chunk top-2 from the
      // register of its parity and request chunk top-4 (same parity) into it.
      const bool oddA = (topA & 1u) != 0, oddB = (topB & 1u) != 0;
      if (crossA && oddA) {
        StoreChunk(ringA, topA - 2, lane, pAO);
        asm volatile("" ::: "memory");
        pAO = LoadChunk(rs, topA - 4, cloA, lane);
      }
      if (crossA && !oddA) {
        StoreChunk(ringA, topA - 2, lane, pAE);
        asm volatile("" ::: "memory");
        pAE = LoadChunk(rs, topA - 4, cloA, lane);
      }
      if (GPW == 2 && crossB && oddB) {
        StoreChunk(ringB, topB - 2, lane, pBO);
        asm volatile("" ::: "memory");
        pBO = LoadChunk(rs, topB - 4, cloB, lane);
      }
      if (GPW == 2 && crossB && !oddB) {
        StoreChunk(ringB, topB - 2, lane, pBE);
        asm volatile("" ::: "memory");
        pBE = LoadChunk(rs, topB - 4, cloB, lane);
      }
      topA -= crossA ? 1u : 0u;
      topB -= crossB ? 1u : 0u;
      if (crossA || crossB) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    if (kSum) {
      // Byte sums of the sector's two 32-symbol halves (v_sad_u8 against zero), less
      // the biases.  The sector covers lane positions 192 - 64 sector + [0, 64), so
      // its halves are sub-segments k = 6 - 2 sector and 7 - 2 sector, and the running
      // total from the top gives the suffix sums R[k] and R[k + 1].
      uint32_t lo = 0, hi = 0;
#pragma unroll
      for (uint32_t i = 0; i < 8; ++i) lo = __builtin_amdgcn_sad_u8(acc[i], 0u, lo);
#pragma unroll
      for (uint32_t i = 8; i < 16; ++i) hi = __builtin_amdgcn_sad_u8(acc[i], 0u, hi);
      const int32_t r_hi = sum + static_cast<int32_t>(hi) - 32 * 128;
      const int32_t r_lo = r_hi + static_cast<int32_t>(lo) - 32 * 128;
      sum = r_lo;
      if (active && lane_suffix)
        *(GPtr<int2>)(lane_suffix + 8 * j + 6 - 2 * sector) = make_int2(r_lo, r_hi);
    }
    if (active) {
      const GPtr<uint4> dst = (GPtr<uint4>)(outp + (kSymsPerLane - 64) - 64 * sector);
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i)
        dst[i] = make_uint4(acc[4 * i], acc[4 * i + 1], acc[4 * i + 2], acc[4 * i + 3]);
    }
  }
  const bool under = lane < 32 ? (pA < loA || pA > startA) : (pB < loB || pB > startB);
  if (active && under && j == 0) atomicOr(err, kErrStreamOffsets);
  if (kSum) {
    // Inclusive prefix of the lane totals over the half-wave's 32 lanes.
    int32_t incl = sum;
#pragma unroll
    for (uint32_t off = 1; off < 32; off <<= 1) {
      const int32_t n = __shfl_up(incl, off, 32);
      if (j >= off) incl += n;
    }
    if (active && lane_sum) lane_sum[j] = incl;
    const int32_t total = __shfl(incl, 31, 32);
    if (active && j == 0) *group_sum = total;
  }
}

// readfirstlane returns a signed int: every use converts through uint32_t so no
// 64-bit widening can sign-extend a value with bit 31 set.
__device__ __forceinline__ uint32_t Uniform(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// Work-groups are dispatched to the 8 XCDs round robin (group b on XCD b % 8), each XCD
// with its own L2.  This maps group b to a unit so that every XCD takes one contiguous
// range of the grid's first pass: neighbouring units, which share cache lines and an
// image's palette and sums, then meet in one L2.  A bijection on [0, n) for any n.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t XcdContiguous(uint32_t b, uint32_t n) {
  const uint32_t q = n / kXcds, r = n % kXcds, xcd = b % kXcds, idx = b / kXcds;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + idx;
}

// The same split by the stage's actual unit count, for grids of any size (persistent
// grids below it, plan grids above it): XCD x owns the x-th of 8 contiguous ranges of
// [0, units) and its work-groups stride through that range, so every XCD gets an equal
// share whatever the grid (a split by the grid's size would leave whole XCDs idle when
// the grid exceeds the units, and one XCD with the last round when it falls short).
// Grids under 8 work-groups leave XCDs without work-groups: they stride over all units.
struct XcdSpan {
  uint32_t first, end, stride;
};
__device__ __forceinline__ XcdSpan XcdBalanced(uint32_t b, uint32_t grid, uint32_t units) {
  if (grid < kXcds) return {b, units, grid};
  const uint32_t q = units / kXcds, r = units % kXcds, xcd = b % kXcds, idx = b / kXcds;
  const uint32_t lo = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  const uint32_t n = q + (xcd < r ? 1u : 0u);
  const uint32_t groups_here = (grid - xcd + kXcds - 1) / kXcds;  // work-groups with b % 8 == xcd
  return {lo + idx, lo + n, groups_here};
}

__device__ __forceinline__ const uint8_t *Uniform(const uint8_t *p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint64_t lo = Uniform(static_cast<uint32_t>(v));
  const uint64_t hi = Uniform(static_cast<uint32_t>(v >> 32));
  return reinterpret_cast<const uint8_t *>(lo | (hi << 32));
}

// Work-group barrier that orders LDS only: waits for the wave's own LDS operations,
// not for its global loads in flight (a __syncthreads fence would wait for those too).
__device__ __forceinline__ void LdsBarrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// The work of BuildTableKernel for one table, straight into the rANS kernel's LDS
// (256 <= THREADS): thread t < 256 brings frequency f[t].  The slot -> symbol map is a
// count, not a search.  Wave w owns slots [w S, (w + 1) S), S = 2048 / waves, lane l
// slot w S + 64 j + l of chunk j:
//   A. the waves clear their slots and the marks-before counters; the first four waves
//      scan f + (f != 0) << 24, giving every symbol its cum (low 24 bits: 256 u16 sum
//      below 2^24) and its rank among the nonzero symbols (high bits);
//   B. every nonzero symbol x writes fc[x] = f | cum << 12, sym_of_rank[rank] = x and a
//      start mark at slot cum[x], and each wave of symbols adds to before[w] how many
//      of its marks fall below slot w S (a ballot per boundary, one LDS add);
//   C. per chunk, a ballot of the marks and mbcnt give the marks up to each slot, so
//      slot s gets symbol sym_of_rank[before + marks in the wave's slots up to s - 1].
// Every LDS access of A and C is one word per lane at consecutive addresses.  A table
// that does not sum to 2048 becomes the identity table.  Returns whether the table is
// valid; ends with a work-group barrier.
template <uint32_t THREADS>
__device__ __forceinline__ bool BuildTableInLds(uint32_t f, uint32_t *tab, uint32_t *fc, uint8_t *sym_of_rank,
                                                uint32_t *wsum, uint32_t *before) {
  constexpr uint32_t kWaves = THREADS / 64, kSlotsPerWave = kTableSize / kWaves, kChunks = kSlotsPerWave / 64;
  static_assert(THREADS >= 256 && kWaves <= 8 && kChunks * 64 * kWaves == kTableSize, "launch shape");
  const uint32_t t = threadIdx.x, lane = t & 63, wave = Uniform(t >> 6);
  uint32_t *const mine = tab + wave * kSlotsPerWave + lane;
#pragma unroll
  for (uint32_t j = 0; j < kChunks; ++j) mine[64 * j] = 0;
  if (t < kWaves) before[t] = 0;
  const uint32_t pk = f + (static_cast<uint32_t>(f != 0) << 24);
  uint32_t v = 0;
  if (t < 256) {
    v = pk;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t n = __shfl_up(v, off);
      if (lane >= static_cast<uint32_t>(off)) v += n;
    }
    if (lane == 63) wsum[wave] = v;
  }
  LdsBarrier();
  const bool valid = ((wsum[0] + wsum[1] + wsum[2] + wsum[3]) & 0xFFFFFFu) == kTableSize;
  if (t < 256 && valid) {
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += wsum[w];
    const uint32_t excl = add + v - pk, cum = excl & 0xFFFFFFu;
    fc[t] = f | (cum << 12);
    if (f != 0) {  // the table sums to 2048, so cum + f <= 2048: the mark is inside it
      sym_of_rank[excl >> 24] = static_cast<uint8_t>(t);
      tab[cum] = 1;
    }
#pragma unroll
    for (uint32_t w = 1; w < kWaves; ++w) {
      const uint32_t below = static_cast<uint32_t>(__popcll(__builtin_amdgcn_ballot_w64(f != 0 && cum < w * kSlotsPerWave)));
      if (lane == 0 && below) atomicAdd(&before[w], below);
    }
  }
  LdsBarrier();
  if (valid) {
    uint32_t base = before[wave];
#pragma unroll
    for (uint32_t j = 0; j < kChunks; ++j) {
      const uint32_t mark = mine[64 * j];
      const uint64_t bal = __builtin_amdgcn_ballot_w64(mark != 0);
      const uint32_t upto = base + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(bal >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), mark));
      const uint32_t x = sym_of_rank[(upto - 1) & 255];
      const uint32_t w = fc[x];
      const uint32_t slot = wave * kSlotsPerWave + 64 * j + lane;
      mine[64 * j] = (w & 0xFFFu) | ((slot - (w >> 12)) << 12) | (x << 24);
      base += static_cast<uint32_t>(__popcll(bal));
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kChunks; ++j) mine[64 * j] = kTableSize | ((wave * kSlotsPerWave + 64 * j + lane) << 12);
  }
  LdsBarrier();
  return valid;
}

// kBuild: build the stream's table from its frequencies in LDS instead of loading
// the one BuildTableKernel wrote (the decode pipeline then skips that launch).
// kPair: units are numbered with short remainders paired (AnsUnitAt); a pair's two
// streams each get half of the waves and a table of their own, built one after the
// other.  The second table takes 8 KiB of LDS, which caps the kernel at 6 work-groups
// per CU: that is the point, a SIMD then never holds more than 6 decoding waves.
// A work-unit starts with a chain of dependent loads (plan -> descriptor -> group
// offsets -> lane states and ring chunks), all from HBM in a full pipeline; the
// frequencies go out with the offsets and the states and chunks before the table is
// built, so the LDS build overlaps their latency.
// LDS of the rANS kernels beside their tables and rings.
template <uint32_t WAVES>
struct AnsLdsSmall {
  uint32_t build_wsum[4];
  uint32_t build_before[WAVES];
  uint32_t touch_sink[64];
};

// The LDS of AnsDecodeUnits: tables (one, two with kPair) and the half-waves' rings.
// The table build's scratch lives in the rings: every wave has left the previous
// unit's ring (barrier) before the build, and the rings are filled after it.
template <uint32_t WAVES, bool kPair>
struct AnsLds {
  uint32_t tab[kPair ? 2 : 1][kTableSize];
  uint16_t ring[WAVES][2][kRingWords];
};

template <uint32_t WAVES, uint32_t GPW, bool kBuild, bool kPair>
__device__ __forceinline__ void AnsDecodeUnits(const KernelArgs &a, AnsLds<WAVES, kPair> &L,
                                               AnsLdsSmall<WAVES> &S) {
  static_assert(!kPair || (kBuild && WAVES % 2 == 0), "paired units build their tables in LDS");
  static_assert(WAVES * 2 * kRingWords * 2 >= 256 * 4 + 256, "build scratch fits the rings");
  auto &tab = L.tab;
  auto &ring = L.ring;
  uint32_t *const build_fc = reinterpret_cast<uint32_t *>(&ring[0][0][0]);
  uint8_t *const build_sym = reinterpret_cast<uint8_t *>(build_fc + 256);
  uint32_t *const build_wsum = S.build_wsum;
  uint32_t *const build_before = S.build_before;
  uint32_t *const touch_sink = S.touch_sink;
  constexpr uint32_t kGroupsPerUnit = WAVES * GPW;
  constexpr uint32_t kUnitCode = kGroupsPerUnit | (kPair ? kAnsPairShort : 0u);
  const uint32_t lane = threadIdx.x & 63, wave = Uniform(threadIdx.x >> 6);
  const uint32_t half = GPW == 2 ? (lane >> 5) : 0;
  const uint32_t j = lane & 31;
  for (uint32_t u = blockIdx.x;; u += gridDim.x) {
    const Located loc = Locate(a, kStageAns, u, kUnitCode);
    if (!loc.valid) break;
    const DevDesc d = LoadDesc(a, loc.img);
    const AnsUnit unit = AnsUnitAt(d, kUnitCode, Uniform(loc.local));
    // This wave's segment: the unit's only one, or its half of a pair.
    const uint32_t seg = kPair && unit.pair && wave >= WAVES / 2 ? 1u : 0u;
    const uint32_t k = Uniform(seg ? unit.k[1] : unit.k[0]);
    const uint32_t first = Uniform(seg ? unit.first[1] : unit.first[0]);
    const uint32_t count = Uniform(seg ? unit.count[1] : unit.count[0]);
    const uint32_t gi = (kPair && unit.pair ? wave % (WAVES / 2) : wave) * GPW + half;
    const uint32_t G = Uniform(Pick4(d.groups, k));
    const uint32_t g = first + gi;
    const uint32_t slen = Uniform(Pick4(d.stream_len, k));
    const __amdgpu_buffer_rsrc_t rs = MakeRsrc(Uniform(Pick4(d.streams, k)), slen);
    // Straight-line loads (out-of-range offsets read zeros): a load under a branch
    // makes the wait for it land at the join, after later loads have gone out.
    const uint32_t end = __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(4 * g), 0, 0);
    const uint32_t lo_word = __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(4 * (g - 1)), 0, 0);
    // Frequencies of the unit's table(s): thread t brings symbol t's.
    uint32_t f = 0, f2 = 0;
    const bool second = kPair && unit.pair && unit.count[1] != 0;
    if constexpr (kBuild) {
      const GPtr<const uint16_t> fr = Gbl<const uint16_t>(Uniform(d.freqs + 512 * unit.k[0]));
      f = fr[threadIdx.x & 255];
      if (kPair && second) {
        const GPtr<const uint16_t> fr2 = Gbl<const uint16_t>(Uniform(d.freqs + 512 * unit.k[1]));
        f2 = fr2[threadIdx.x & 255];
      }
    }
    const uint32_t lo = g == 0 ? 4 * G : lo_word;
    const bool ok = ((end | lo) & 3) == 0 && end <= slen && lo >= 4 * G && lo <= end &&
                    end - lo >= 4 * kLanesPerGroup;
    const bool in_stream = (GPW == 2 || lane < 32) && gi < count;
    if (in_stream && !ok && j == 0) atomicOr(a.err, kErrStreamOffsets);
    const bool active = in_stream && ok;
    const uint32_t x_word =
        __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(end - 4 * kLanesPerGroup + 4 * j), 0, 0);
    const uint32_t x = active ? x_word : 0u;
    const uint32_t p = active ? (end - 4 * kLanesPerGroup) >> 1 : 0u;
    const uint32_t plo = active ? lo >> 1 : 0u;
    const uint32_t pA = __builtin_amdgcn_readlane(p, 0);
    const uint32_t loA = __builtin_amdgcn_readlane(plo, 0);
    const uint32_t pB = GPW == 2 ? __builtin_amdgcn_readlane(p, 32) : 0u;
    const uint32_t loB = GPW == 2 ? __builtin_amdgcn_readlane(plo, 32) : 0u;
    const AnsRing chunks = RequestRing<GPW>(rs, pA, pB, loA, loB, lane);

    LdsBarrier();  // every wave is done with the previous unit's tables and rings
    bool table_ok;
    if constexpr (kBuild) {
      table_ok = BuildTableInLds<WAVES * 64>(f, tab[0], build_fc, build_sym, build_wsum, build_before);
      if (kPair && second)
        table_ok = BuildTableInLds<WAVES * 64>(f2, tab[kPair ? 1 : 0], build_fc, build_sym, build_wsum,
                                               build_before) && table_ok;
    } else {
      const uint32_t tix = Uniform(d.table_base + k);
      const GPtr<const uint32_t> src = Gbl<const uint32_t>(a.tables + static_cast<size_t>(tix) * kTableSize);
      for (uint32_t i = threadIdx.x; i < kTableSize; i += WAVES * 64) tab[0][i] = src[i];
      table_ok = Gbl<const uint32_t>(a.table_status)[tix] != 0;
      LdsBarrier();
    }
    if (!table_ok && threadIdx.x == 0) atomicOr(a.err, kErrFrequencies);

    // A wave past its stream's last group (the stream's group count is not a multiple
    // of the unit) has nothing to decode: it would only take issue slots from the waves
    // that do.  It goes straight to the next unit's barrier.
    if (__builtin_amdgcn_ballot_w64(active) == 0) continue;
    const uint32_t *const tw = tab[kPair ? seg : 0];
    const GPtr<uint8_t> outp = d.symbols + SymbolOffset(d, k) + static_cast<size_t>(g) * kGroupSyms + j * kSymsPerLane;
    uint16_t *ring_half = ring[wave][half];
    if (k == 3)
      DecodeHalfGroups<true, GPW>(tw, ring[wave][0], ring_half, touch_sink, rs, chunks, lane, active, x, pA, pB,
                                  loA, loB, outp, d.group_sums + g,
                                  d.lane_sums ? d.lane_sums + g * kLanesPerGroup : nullptr,
                                  d.lane_sums ? d.lane_sums + LaneSuffixBase(d.nv) + g * kLanesPerGroup * 8
                                              : nullptr,
                                  a.err);
    else
      DecodeHalfGroups<false, GPW>(tw, ring[wave][0], ring_half, touch_sink, rs, chunks, lane, active, x, pA, pB,
                                   loA, loB, outp, nullptr, nullptr, nullptr, a.err);
  }
}

template <uint32_t WAVES, uint32_t GPW, bool kBuild, bool kPair = false>
// With 4 or 8 waves LDS allows 7 waves per SIMD (6 with the second table); hold the
// VGPRs to that (at most 72, or 80).
__global__ __launch_bounds__(WAVES * 64, kPair ? 6 : WAVES >= 4 ? 7 : 1)
void AnsDecodeKernel(KernelArgs a) {
  __shared__ __align__(16) AnsLds<WAVES, kPair> lds;
  __shared__ AnsLdsSmall<WAVES> small;
  AnsDecodeUnits<WAVES, GPW, kBuild, kPair>(a, lds, small);
}

// The rANS stage whose work-unit form is data: the plan's unit code (or, for a single
// image, the one in the arguments) selects units of 8 groups of one stream or the
// paired form (short remainders share a work-group), both AnsDecodeUnits<4, 2> with
// tables built in LDS.  The two share the LDS of the larger, the paired form's two
// tables (6 work-groups per CU), and a captured graph of this launch replays either,
// so a batch can change forms from bind to bind.
constexpr uint32_t kAutoWaves = 4;
union AnsAutoLds {
  AnsLds<kAutoWaves, false> single;
  AnsLds<kAutoWaves, true> paired;
};

__global__ __launch_bounds__(kAutoWaves * 64, 6) void AnsAutoKernel(KernelArgs a) {
  __shared__ __align__(16) AnsAutoLds lds;
  __shared__ AnsLdsSmall<kAutoWaves> small;
  const uint32_t code = a.plan ? Uniform(Gbl<const uint32_t>(&a.plan->ans_unit_code)[0]) : a.ans_unit_code;
  if (code & kAnsPairShort)
    AnsDecodeUnits<kAutoWaves, 2, true, true>(a, lds.paired, small);
  else
    AnsDecodeUnits<kAutoWaves, 2, true, false>(a, lds.single, small);
}

// ---------------------------------------------------------------------------
// Stand-alone interleaved rANS groups of 1..32 streams, in the layout the reference's
// ANS decoder object packs (ans/ans_ocl.cpp:154-427) and its kernel reads
// (ans/ans_decode.cl:25-95): `groups` u32 end offsets, then per group its 16-bit
// renormalisation words in emission order followed by its n final states.  Same
// step as the texture path: slot = x & 2047, x = (x >> 11) f + slot - cum, and lanes
// with x < 2^15 take the next words in lane order.
//
// One wave per group; lanes >= n sit out of the ballot.  The table and the group's
// last 8192 words (a group of 32 streams x 256 steps reads at most that many) are
// copied to LDS first, so the serial chain reads only LDS; every global read is
// range-checked through a buffer resource.  Stream j of group g leaves its 256
// symbols at out[(g n + j) 256 ..], the last decoded first, 4 bytes per store.  A
// malformed end offset, or a group that does not end in the encoder's initial state
// 2^15, raises GST_DEVERR_STREAM_OFFSETS.
constexpr uint32_t kGroupWindow = kLanesPerGroup * kSymsPerLane;  // 8192 words

__global__ __launch_bounds__(64) void AnsGroupsKernel(const uint32_t *table, const uint8_t *data,
                                                      uint32_t data_len, uint32_t n, uint8_t *out,
                                                      uint32_t *err) {
  __shared__ uint32_t tab[kTableSize];
  __shared__ uint16_t words[kGroupWindow];
  const uint32_t lane = threadIdx.x, g = blockIdx.x;
  const GPtr<const uint32_t> t = Gbl<const uint32_t>(table);
  for (uint32_t i = lane; i < kTableSize; i += 64) tab[i] = t[i];
  const __amdgpu_buffer_rsrc_t rs = MakeRsrc(data, data_len);
  const uint32_t end = __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(4 * g), 0, 0);
  const bool ok = (end & 3) == 0 && end <= data_len && end >= 4 * n;
  const uint32_t p0 = ok ? (end - 4 * n) / 2 : 0u;
  const uint32_t wbase = p0 > kGroupWindow ? p0 - kGroupWindow : 0u;
  for (uint32_t i = lane; i < kGroupWindow; i += 64) {
    const uint32_t w = wbase + i;
    words[i] = w < p0 ? static_cast<uint16_t>(__builtin_amdgcn_raw_buffer_load_b16(rs, static_cast<int>(2 * w), 0, 0))
                      : static_cast<uint16_t>(0);
  }
  const bool active = ok && lane < n;
  uint32_t x = __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(active ? end - 4 * n + 4 * lane : 0u), 0, 0);
  x = active ? x : 0u;
  __syncthreads();
  uint32_t p = p0, acc = 0;
  const GPtr<uint32_t> dst = Gbl<uint32_t>(reinterpret_cast<uint32_t *>(out + (static_cast<size_t>(g) * n + lane) *
                                                                                 kSymsPerLane));
  for (uint32_t i = 0; i < kSymsPerLane; ++i) {
    const uint32_t e = tab[x & (kTableSize - 1)];
    x = __umul24(x >> kTableLog, e & 0xFFFu) + __builtin_amdgcn_ubfe(e, 12, 11);
    const bool need = active && x < kStateL;
    const uint64_t m = __builtin_amdgcn_ballot_w64(need);
    const uint32_t total = static_cast<uint32_t>(__popcll(m));
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
    if (need) x = (x << 16) | words[(p - total + rank - wbase) & (kGroupWindow - 1)];
    p -= total;
    acc = (acc << 8) | (e >> 24);
    if ((i & 3) == 3 && active) dst[(kSymsPerLane - 1 - i) / 4] = acc;
  }
  const bool bad = (lane == 0 && !ok) || (active && x != kStateL);
  if (__builtin_amdgcn_ballot_w64(bad) != 0 && lane == 0) atomicOr(err, kErrStreamOffsets);
}

// ---------------------------------------------------------------------------
// Stage 3: palette indices (codec/decode_indices.cl:6-84 composed by
// codec/decoder.cpp:304-393): idx[i] = sum_{j<=i} (diff[j] - 128), int32.
//
// One work-group per 8192-entry index group.  The group's starting value is the
// sum of the preceding groups' totals, which the rANS stage already produced, so
// the whole scan is one launch with no inter-work-group traffic.  Each thread
// scans 32 entries in registers, then wave and work-group scans join them.
__global__ __launch_bounds__(256) void DecodeIndicesKernel(KernelArgs a) {
  __shared__ uint32_t wtot[4];
  __shared__ uint32_t gprefix;
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (uint32_t u = blockIdx.x;; u += gridDim.x) {
    const Located loc = Locate(a, kStageIndices, u, 1);
    if (!loc.valid) break;
    const DevDesc d = LoadDesc(a, loc.img);
    const uint32_t g = loc.local;
    if (wave == 0) {
      uint32_t s = 0;
      for (uint32_t i = lane; i < g; i += 64) s += static_cast<uint32_t>(d.group_sums[i]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) gprefix = s;
    }
    const GPtr<const uint4> diffs =
        (GPtr<const uint4>)(d.symbols + SymbolOffset(d, 3) + static_cast<size_t>(g) * kGroupSyms + 32 * t);
    const uint4 q0 = diffs[0];
    const uint4 q1 = diffs[1];
    const uint32_t words[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    uint32_t run[32];
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      acc += ((words[i >> 2] >> (8 * (i & 3))) & 0xFFu) - 128u;
      run[i] = acc;
    }
    uint32_t incl = acc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t n = __shfl_up(incl, off);
      if (lane >= static_cast<uint32_t>(off)) incl += n;
    }
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    uint32_t offset = gprefix + incl - acc;
    for (uint32_t w = 0; w < wave; ++w) offset += wtot[w];
    const GPtr<uint4> dst = (GPtr<uint4>)(d.indices + static_cast<size_t>(g) * kGroupSyms + 32 * t);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      dst[i] = make_uint4(run[4 * i] + offset, run[4 * i + 1] + offset, run[4 * i + 2] + offset,
                          run[4 * i + 3] + offset);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stage 4: inverse 5/3 wavelet (codec/inverse_wavelet.cl:14-192).
//
// A wave64 transforms four 32x32 tiles (two work units) at once: lane r of half-wave
// h holds row r of tile h of the first unit in the low 16 bits and row r of tile h
// of the second unit in the high 16 bits of 32 registers, so every add and shift is
// a packed 16-bit op over two tiles.  16 bits reproduce the reference's int32
// arithmetic exactly: inputs are symbol - 128 in [-128, 127]; a 1-D pass maps
// low-band bound Lo and high-band bound Hi to at most Lo + 1.5 Hi + 0.5 (the 5/3
// synthesis taps sum to 1 on the low band and 1.5 on the high band, truncation adds
// under 1), so a level (rows, then columns) adds at most 674 to the bound, every
// value after 5 levels stays within 128 + 5 * 674 = 3498 and every sum formed on the
// way within 2 * 3498 + 2.
//
// The row pass of each level is straight-line code with compile-time mirror
// indices; the column pass runs after a transpose through LDS (row stride 33
// words: conflict-free in both directions).  Levels len = 2..32 each transform rows
// y < len over x < len, then columns x < len over y < len, exactly the reference's
// transposed ping-pong.  A last transpose returns rows, whose low bytes leave as
// int8 in two 16-byte stores per tile row.
typedef int16_t s16x2 __attribute__((ext_vector_type(2)));
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 AsPair(uint32_t w) { return __builtin_bit_cast(s16x2, w); }
__device__ __forceinline__ uint32_t Bits(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// C's truncating division of each half: negative values take divisor - 1 first.  Every
// sum divided here is within 2 * 3498 + 2 < 2^14 in magnitude (bound above), so bits
// 14 and 15 agree and the top two bits are the divisor - 1 (3) exactly when negative.
__device__ __forceinline__ s16x2 TruncDiv4(s16x2 s) {
  const u16x2 fix = __builtin_bit_cast(u16x2, s) >> static_cast<u16x2>(14);  // 0 or 3
  return (s + __builtin_bit_cast(s16x2, fix)) >> static_cast<s16x2>(2);
}
__device__ __forceinline__ s16x2 TruncDiv2(s16x2 s) {
  const u16x2 sign = __builtin_bit_cast(u16x2, s) >> static_cast<u16x2>(15);  // 0 or 1
  return (s + __builtin_bit_cast(s16x2, sign)) >> static_cast<s16x2>(1);
}

__host__ __device__ constexpr int Mirror(int idx, int range) {
  return (idx - static_cast<int>(idx >= range) * (idx - range + 2)) < 0
             ? -(idx - static_cast<int>(idx >= range) * (idx - range + 2))
             : (idx - static_cast<int>(idx >= range) * (idx - range + 2));
}

template <int LEN>
__device__ __forceinline__ void Inverse1D(s16x2 (&v)[kTileDim]) {
  constexpr int MID = LEN / 2;
  s16x2 e[LEN];
#pragma unroll
  for (int k = 0; k < MID; ++k) {
    const int p = MID + Mirror(2 * k - 1, LEN) / 2;
    const int n = MID + Mirror(2 * k + 1, LEN) / 2;
    e[2 * k] = v[k] - TruncDiv4(v[p] + v[n] + static_cast<s16x2>(2));
  }
#pragma unroll
  for (int k = 0; k < MID; ++k) {
    const int p = Mirror(2 * k, LEN);
    const int n = Mirror(2 * k + 2, LEN);
    e[2 * k + 1] = v[MID + k] + TruncDiv2(e[p] + e[n]);
  }
#pragma unroll
  for (int i = 0; i < LEN; ++i) v[i] = e[i];
}

constexpr int kLdsStride = kTileDim + 1;

// One level on the half-wave's tile pair; lanes r >= LEN sit out (their rows and
// columns are outside this level's square).  Each half-wave owns its LDS tile, so
// wave-level ordering is all the transposes need.
template <int LEN>
__device__ __forceinline__ void WaveletLevel(s16x2 (&v)[kTileDim], uint32_t *L, int r) {
  if (r < LEN) {
    Inverse1D<LEN>(v);
#pragma unroll
    for (int c = 0; c < LEN; ++c) L[r * kLdsStride + c] = Bits(v[c]);
  }
  WaveSync();
  if (r < LEN) {
#pragma unroll
    for (int c = 0; c < LEN; ++c) v[c] = AsPair(L[c * kLdsStride + r]);
  }
  WaveSync();
  if (r < LEN) {
    Inverse1D<LEN>(v);
#pragma unroll
    for (int c = 0; c < LEN; ++c) L[c * kLdsStride + r] = Bits(v[c]);
  }
  WaveSync();
  if (r < LEN) {
#pragma unroll
    for (int c = 0; c < LEN; ++c) v[c] = AsPair(L[r * kLdsStride + c]);
  }
  WaveSync();
}

struct TileRef {
  GPtr<const uint8_t> src;  // tile-linear symbols (1024 bytes)
  GPtr<int8_t> dst;         // top-left of the tile in the raster plane
  uint32_t pitch;           // plane row length (blocks)
};

__device__ __forceinline__ TileRef TileAt(const DevDesc &d, uint32_t tile) {
  const uint32_t per_plane = d.nv / kTileSyms, tiles_x = d.bx / kTileDim;
  const uint32_t plane = tile / per_plane, t = tile - plane * per_plane;
  const uint32_t ty = t / tiles_x, tx = t - ty * tiles_x;
  const size_t base = static_cast<size_t>(plane) * d.nv;
  return {d.symbols + base + static_cast<size_t>(t) * kTileSyms,
          d.planes + base + static_cast<size_t>(ty) * kTileDim * d.bx + tx * kTileDim, d.bx};
}

// Tile `half` of wavelet unit `u`, if the unit exists (uniform address math, one
// select per lane; whole-wave call).
__device__ __forceinline__ bool WaveletTile(const KernelArgs &a, uint32_t u, uint32_t half, TileRef *t) {
  const Located loc = Locate(a, kStageWavelet, u, 1);
  if (!loc.valid) return false;
  const DevDesc d = LoadDesc(a, loc.img);
  const TileRef t0 = TileAt(d, loc.local * kWaveletTilesPerUnit);
  const TileRef t1 = TileAt(d, loc.local * kWaveletTilesPerUnit + 1);
  *t = half ? t1 : t0;
  return true;
}

// Row r of two tiles (a: 32 symbols of the first, b: of the second) through the five
// levels, with this half-wave's transpose tile L; returns the rows' int8 results,
// 32 bytes each.
__device__ __forceinline__ void TransformRowPair(const uint4 &a0, const uint4 &a1, const uint4 &b0,
                                                 const uint4 &b1, uint32_t *L, uint32_t r, uint32_t (&oa)[8],
                                                 uint32_t (&ob)[8]) {
  const uint32_t wa[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  const uint32_t wb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
  s16x2 v[kTileDim];
#pragma unroll
  for (int i = 0; i < static_cast<int>(kTileDim); ++i) {
    // Byte i&3 of the first tile's word to bits 0-7, of the second's to bits
    // 16-23 (selector 0x0c is a zero byte), then remove the +128 bias.
    const uint32_t q = static_cast<uint32_t>(i & 3);
    const uint32_t sel = 0x0C000C00u | q | ((4u + q) << 16);
    v[i] = AsPair(__builtin_amdgcn_perm(wb[i >> 2], wa[i >> 2], sel)) - static_cast<s16x2>(128);
  }
  WaveletLevel<2>(v, L, static_cast<int>(r));
  WaveletLevel<4>(v, L, static_cast<int>(r));
  WaveletLevel<8>(v, L, static_cast<int>(r));
  WaveletLevel<16>(v, L, static_cast<int>(r));
  WaveletLevel<32>(v, L, static_cast<int>(r));
#pragma unroll
  for (int w = 0; w < 8; ++w) {
    // Low bytes of four pairs: t01 = (A0, A1, B0, B1), t23 = (A2, A3, B2, B3).
    const uint32_t t01 = __builtin_amdgcn_perm(Bits(v[4 * w + 1]), Bits(v[4 * w]), 0x06020400u);
    const uint32_t t23 = __builtin_amdgcn_perm(Bits(v[4 * w + 3]), Bits(v[4 * w + 2]), 0x06020400u);
    oa[w] = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
    ob[w] = __builtin_amdgcn_perm(t23, t01, 0x07060302u);
  }
}

template <uint32_t WAVES>
__global__ __launch_bounds__(WAVES * 64) void InvWaveletKernel(KernelArgs a) {
  __shared__ uint32_t lds[WAVES][2][kTileDim * kLdsStride];
  const uint32_t lane = threadIdx.x & 63, half = lane >> 5, r = lane & 31;
  const uint32_t wave = Uniform(threadIdx.x >> 6);
  uint32_t *L = lds[wave][half];
  const uint32_t stride = gridDim.x * WAVES;
  // The rows of the next pair are requested before the current pair is transformed,
  // so a wave's loads overlap its own lifting when it takes more than one pair.
  uint32_t pair = blockIdx.x * WAVES + wave;
  TileRef nlo{}, nhi{};
  uint4 na0{}, na1{}, nb0{}, nb1{};
  bool nhas_lo = WaveletTile(a, 2 * pair, half, &nlo);
  bool nhas_hi = nhas_lo && WaveletTile(a, 2 * pair + 1, half, &nhi);
  if (nhas_lo) {
    na0 = ((GPtr<const uint4>)(nlo.src + r * kTileDim))[0];
    na1 = ((GPtr<const uint4>)(nlo.src + r * kTileDim))[1];
    if (nhas_hi) {
      nb0 = ((GPtr<const uint4>)(nhi.src + r * kTileDim))[0];
      nb1 = ((GPtr<const uint4>)(nhi.src + r * kTileDim))[1];
    }
  }
  for (; nhas_lo; pair += stride) {
    const TileRef lo = nlo, hi = nhi;
    const bool has_hi = nhas_hi;
    const uint4 a0 = na0, a1 = na1;
    const uint4 b0 = nhas_hi ? nb0 : make_uint4(0, 0, 0, 0), b1 = nhas_hi ? nb1 : make_uint4(0, 0, 0, 0);
    const uint32_t next = pair + stride;
    nhas_lo = WaveletTile(a, 2 * next, half, &nlo);
    nhas_hi = nhas_lo && WaveletTile(a, 2 * next + 1, half, &nhi);
    if (nhas_lo) {
      na0 = ((GPtr<const uint4>)(nlo.src + r * kTileDim))[0];
      na1 = ((GPtr<const uint4>)(nlo.src + r * kTileDim))[1];
      if (nhas_hi) {
        nb0 = ((GPtr<const uint4>)(nhi.src + r * kTileDim))[0];
        nb1 = ((GPtr<const uint4>)(nhi.src + r * kTileDim))[1];
      }
    }
    uint32_t oa[8], ob[8];
    TransformRowPair(a0, a1, b0, b1, L, r, oa, ob);
    const GPtr<uint4> da = (GPtr<uint4>)(lo.dst + static_cast<size_t>(r) * lo.pitch);
    da[0] = make_uint4(oa[0], oa[1], oa[2], oa[3]);
    da[1] = make_uint4(oa[4], oa[5], oa[6], oa[7]);
    if (has_hi) {
      const GPtr<uint4> db = (GPtr<uint4>)(hi.dst + static_cast<size_t>(r) * hi.pitch);
      db[0] = make_uint4(ob[0], ob[1], ob[2], ob[3]);
      db[1] = make_uint4(ob[4], ob[5], ob[6], ob[7]);
    }
  }
}

// ---------------------------------------------------------------------------
// Stage 5: DXT1 assembly (codec/assemble.cl:14-81) and RGB8 assembly (:83-130).
//
// A thread assembles 4 consecutive blocks: one 4-byte load per plane, one 16-byte
// index load, 4 palette gathers and two 16-byte stores.  YCoCg -> RGB565 uses C
// truncating division; the pack keeps the low 16 bits of the int shifts.  A palette
// index outside the palette (corrupt input only) yields a zero word and raises
// kErrPaletteIndex instead of reading out of bounds.
__device__ __forceinline__ uint32_t Rgb565(int y, int co, int cg) {
  const int t = y - cg / 2;
  const int g = cg + t;
  const int b = (t - co) / 2;
  const int r = b + co;
  return ((static_cast<uint32_t>(r) << 11) | (static_cast<uint32_t>(g) << 5) |
          static_cast<uint32_t>(b)) & 0xFFFFu;
}

__device__ __forceinline__ int SByte(uint32_t w, int q) {
  return static_cast<int>(static_cast<int8_t>((w >> (8 * q)) & 0xFFu));
}

__global__ __launch_bounds__(256) void AssembleDxtKernel(KernelArgs a) {
  for (uint32_t u = blockIdx.x;; u += gridDim.x) {
    const Located loc = Locate(a, kStageAssemble, u, 1);
    if (!loc.valid) break;
    const DevDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const size_t b0 = static_cast<size_t>(loc.local) * kAssembleBlocksPerUnit + 4 * threadIdx.x;
    uint32_t pw[6];
#pragma unroll
    for (int p = 0; p < 6; ++p)
      pw[p] = *(GPtr<const uint32_t>)(d.planes + static_cast<size_t>(p) * nv + b0);
    const int4 ix = *(GPtr<const int4>)(d.indices + b0);
    const GPtr<const uint32_t> pal = (GPtr<const uint32_t>)(d.symbols + SymbolOffset(d, 2));
    const uint32_t words = d.palette_bytes / 4;
    const int idx[4] = {ix.x, ix.y, ix.z, ix.w};
    uint32_t o[8];
    bool bad = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t ep1 = Rgb565(SByte(pw[0], q), SByte(pw[2], q), SByte(pw[3], q));
      const uint32_t ep2 = Rgb565(SByte(pw[1], q), SByte(pw[4], q), SByte(pw[5], q));
      const uint32_t pi = static_cast<uint32_t>(idx[q]);
      uint32_t word = 0;
      if (pi < words) word = pal[pi]; else bad = true;
      o[2 * q] = ep1 | (ep2 << 16);
      o[2 * q + 1] = word;
    }
    const GPtr<uint4> dst = (GPtr<uint4>)(d.out + 8 * b0);
    dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
    dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
    if (bad) atomicOr(a.err, kErrPaletteIndex);
  }
}

// Index scan fused into DXT1 assembly: one wave per 1024-block unit, which covers
// four 256-block segments of the index stream (one rANS lane each).  The value
// before the unit is the sum of the earlier groups' totals plus the sum before it
// within its group, both from the partial sums the rANS stage wrote; a segment is a
// 4-per-lane register prefix plus a wave scan.  The palette indices never go to
// memory, so a block costs 1 diff byte instead of 1 + 4 + 4 bytes of index traffic.
// Results are those of DecodeIndicesKernel followed by AssembleDxtKernel.
__device__ __forceinline__ int32_t WaveSum64(int32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Sum of (diff - 128) over the blocks of b's rANS group that come before block b
// (b a multiple of 32): the lane's inclusive prefix less the lane's suffix from b.
__device__ __forceinline__ int32_t GroupLocalPrefix(const DevDesc &d, uint32_t b) {
  const uint32_t lane_ix = b / kSymsPerLane;  // group * 32 + lane
  const uint32_t k = (b % kSymsPerLane) / kRowBlocks;
  return d.lane_sums[lane_ix] - d.lane_sums[LaneSuffixBase(d.nv) + lane_ix * 8 + k];
}

template <uint32_t WAVES>
__global__ __launch_bounds__(WAVES * 64) void AssembleFusedKernel(KernelArgs a) {
  constexpr uint32_t kSegments = kAssembleBlocksPerUnit / kSymsPerLane;  // 4
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = Uniform(threadIdx.x >> 6);
  for (uint32_t u = blockIdx.x * WAVES + wave;; u += gridDim.x * WAVES) {
    const Located loc = Locate(a, kStageAssemble, u, 1);
    if (!loc.valid) break;
    const DevDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const uint32_t b0 = loc.local * kAssembleBlocksPerUnit;
    const uint32_t g = b0 / kGroupSyms;
    int32_t base = 0;
    for (uint32_t i = lane; i < g; i += 64) base += d.group_sums[i];
    base = WaveSum64(base) + GroupLocalPrefix(d, b0);
    const GPtr<const uint8_t> diffs = d.symbols + SymbolOffset(d, 3);
    const GPtr<const uint32_t> pal = (GPtr<const uint32_t>)(d.symbols + SymbolOffset(d, 2));
    const uint32_t words = d.palette_bytes / 4;
    bool bad = false;
#pragma unroll
    for (uint32_t seg = 0; seg < kSegments; ++seg) {
      const uint32_t blk = b0 + seg * kSymsPerLane + 4 * lane;
      const uint32_t dw = *(GPtr<const uint32_t>)(diffs + blk);
      uint32_t pw[6];
#pragma unroll
      for (int p = 0; p < 6; ++p)
        pw[p] = *(GPtr<const uint32_t>)(d.planes + static_cast<size_t>(p) * nv + blk);
      int32_t run[4];
      int32_t acc = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        acc += static_cast<int32_t>((dw >> (8 * q)) & 0xFFu) - 128;
        run[q] = acc;
      }
      int32_t incl = acc;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int32_t n = __shfl_up(incl, off);
        if (lane >= static_cast<uint32_t>(off)) incl += n;
      }
      const int32_t before = base + incl - acc;
      base += __shfl(incl, 63);
      uint32_t o[8];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t ep1 = Rgb565(SByte(pw[0], q), SByte(pw[2], q), SByte(pw[3], q));
        const uint32_t ep2 = Rgb565(SByte(pw[1], q), SByte(pw[4], q), SByte(pw[5], q));
        const uint32_t pi = static_cast<uint32_t>(before + run[q]);
        uint32_t word = 0;
        if (pi < words) word = pal[pi]; else bad = true;
        o[2 * q] = ep1 | (ep2 << 16);
        o[2 * q + 1] = word;
      }
      const GPtr<uint4> dst = (GPtr<uint4>)(d.out + 8 * static_cast<size_t>(blk));
      dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
      dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
    }
    if (bad) atomicOr(a.err, kErrPaletteIndex);
  }
}

// The same unit with 16 consecutive blocks per lane: 16-byte loads of the diffs and
// of each plane, one wave scan per unit, 128 contiguous output bytes per lane.
__device__ __forceinline__ uint32_t ByteOf(const uint4 &v, int q) {
  const uint32_t w = q < 4 ? v.x : q < 8 ? v.y : q < 12 ? v.z : v.w;
  return (w >> (8 * (q & 3))) & 0xFFu;
}

template <uint32_t WAVES>
__global__ __launch_bounds__(WAVES * 64) void AssembleFusedWideKernel(KernelArgs a) {
  constexpr uint32_t kPerLane = kAssembleBlocksPerUnit / 64;  // 16
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = Uniform(threadIdx.x >> 6);
  for (uint32_t u = blockIdx.x * WAVES + wave;; u += gridDim.x * WAVES) {
    const Located loc = Locate(a, kStageAssemble, u, 1);
    if (!loc.valid) break;
    const DevDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const uint32_t b0 = loc.local * kAssembleBlocksPerUnit;
    const uint32_t blk = b0 + kPerLane * lane;
    const uint4 dv = *(GPtr<const uint4>)(d.symbols + SymbolOffset(d, 3) + blk);
    uint4 pv[6];
#pragma unroll
    for (int p = 0; p < 6; ++p)
      pv[p] = *(GPtr<const uint4>)(d.planes + static_cast<size_t>(p) * nv + blk);
    const uint32_t g = b0 / kGroupSyms;
    int32_t base = 0;
    for (uint32_t i = lane; i < g; i += 64) base += d.group_sums[i];
    base = WaveSum64(base) + GroupLocalPrefix(d, b0);
    int32_t run[kPerLane];
    int32_t acc = 0;
#pragma unroll
    for (int q = 0; q < static_cast<int>(kPerLane); ++q) {
      acc += static_cast<int32_t>(ByteOf(dv, q)) - 128;
      run[q] = acc;
    }
    int32_t incl = acc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t n = __shfl_up(incl, off);
      if (lane >= static_cast<uint32_t>(off)) incl += n;
    }
    const int32_t before = base + incl - acc;
    const GPtr<const uint32_t> pal = (GPtr<const uint32_t>)(d.symbols + SymbolOffset(d, 2));
    const uint32_t words = d.palette_bytes / 4;
    bool bad = false;
    const GPtr<uint4> dst = (GPtr<uint4>)(d.out + 8 * static_cast<size_t>(blk));
#pragma unroll
    for (int i = 0; i < static_cast<int>(kPerLane) / 2; ++i) {
      uint32_t o[4];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int q = 2 * i + h;
        const int32_t y1 = static_cast<int8_t>(ByteOf(pv[0], q)), y2 = static_cast<int8_t>(ByteOf(pv[1], q));
        const int32_t co1 = static_cast<int8_t>(ByteOf(pv[2], q)), cg1 = static_cast<int8_t>(ByteOf(pv[3], q));
        const int32_t co2 = static_cast<int8_t>(ByteOf(pv[4], q)), cg2 = static_cast<int8_t>(ByteOf(pv[5], q));
        const uint32_t pi = static_cast<uint32_t>(before + run[q]);
        uint32_t word = 0;
        if (pi < words) word = pal[pi]; else bad = true;
        o[2 * h] = Rgb565(y1, co1, cg1) | (Rgb565(y2, co2, cg2) << 16);
        o[2 * h + 1] = word;
      }
      dst[i] = make_uint4(o[0], o[1], o[2], o[3]);
    }
    if (bad) atomicOr(a.err, kErrPaletteIndex);
  }
}

// Byte k of a and of b, sign-extended into the low and high 16-bit halves.
__device__ __forceinline__ s16x2 SBytePair(uint32_t a, uint32_t b, uint32_t k) {
  // Selector codes 0-3 take bytes of the second operand, 4-7 of the first, 0x0c a zero:
  // a's byte to bits 8-15, b's to bits 24-31, then an arithmetic shift of each half.
  const uint32_t sel = 0x0C | (k << 8) | (0x0Cu << 16) | ((4u + k) << 24);
  return AsPair(__builtin_amdgcn_perm(b, a, sel)) >> static_cast<s16x2>(8);
}

// Rgb565 of both endpoints of a block at once, endpoint 1 in the low and endpoint 2 in
// the high half: the same truncating divisions and the same low 16 bits of the shifts
// (|y|, |co|, |cg| <= 128, so every intermediate fits 16 bits).  Arguments: the four
// blocks' plane words (y1, y2, co1, co2, cg1, cg2) and the block's byte k.
__device__ __forceinline__ uint32_t Rgb565Pair(uint32_t y1, uint32_t y2, uint32_t co1, uint32_t co2,
                                               uint32_t cg1, uint32_t cg2, uint32_t k) {
  const s16x2 y = SBytePair(y1, y2, k), co = SBytePair(co1, co2, k);
  uint32_t cgw = Bits(SBytePair(cg1, cg2, k));
  asm volatile("" : "+v"(cgw));  // else cg / 2 is split into two 16-bit halves and re-packed
  const s16x2 cg = AsPair(cgw);
  const s16x2 t = y - TruncDiv2(cg);
  const s16x2 g = cg + t;
  const s16x2 b = TruncDiv2(t - co);
  const s16x2 r = b + co;
  return Bits(r << static_cast<s16x2>(11)) | Bits(g << static_cast<s16x2>(5)) | Bits(b);
}

// Stages 4 and 5 in one kernel: inverse wavelet, index scan and DXT1 assembly of two
// tile positions per work-group, so the int8 planes never leave LDS (the planes
// buffer's write and read, 12 bytes a block, are the traffic this saves).
//
// A work unit is two consecutive assembly units of one image, i.e. tile positions
// t0 and t0 + 1 of the per-plane raster tile order (every plane has a multiple of 8
// tiles, so pairs never straddle images).  Wave w transforms planes 2w and 2w + 1,
// one per half-wave, with t0 in the low and t0 + 1 in the high 16 bits, exactly as
// InvWaveletKernel does, and leaves the int8 rows of its four tiles in its own
// transpose space (1 KiB a tile).  After a work-group barrier each half-wave
// assembles tile rows: lane x of row y of position P is block (32 ty + y, 32 tx + x)
// of that position; its six plane values are LDS bytes, its palette index is the
// row's base plus a 32-lane scan of the diffs, and its 8 bytes join the row's 256
// contiguous output bytes.  A row's base is the earlier groups' totals (one 64-lane
// scan of the group sums of the groups the unit's rows touch, when at most 64; else a
// per-row reduction) plus the sum before the row within its group.
constexpr uint32_t kTileFusedWaves = 3;

// At most 96 VGPRs: 5 waves per SIMD, about what the 25 KiB of LDS allows.
__global__ __launch_bounds__(kTileFusedWaves * 64, 5) void WaveletAssembleKernel(KernelArgs a) {
  __shared__ uint32_t lds[kTileFusedWaves][2][kTileDim * kLdsStride];
  constexpr uint32_t kWaveLdsBytes = sizeof(lds[0]);
  const uint32_t lane = threadIdx.x & 63, half = lane >> 5, r = lane & 31;
  const uint32_t wave = Uniform(threadIdx.x >> 6);
  uint32_t *const L = lds[wave][half];
  // Work units are pairs of assembly units (every image has an even number of them).
  const uint32_t pairs = (a.plan ? Uniform(Gbl<const uint32_t>(&a.plan->units[kStageAssemble])[0])
                                 : StageUnitsOf(a.single, kStageAssemble, 1)) / 2;
  const XcdSpan span = XcdBalanced(blockIdx.x, gridDim.x, pairs);
  for (uint32_t u = span.first; u < span.end; u += span.stride) {
    const Located loc = Locate(a, kStageAssemble, 2 * u, 1);
    if (!loc.valid) break;
    const DevDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv, bx = d.bx, tiles_x = bx / kTileDim;
    const uint32_t t0 = loc.local, t1 = t0 + 1;
    // Rows of plane 2 wave + half, tiles t0 and t1 (adjacent in the tile-linear symbols).
    const GPtr<const uint8_t> src = d.symbols + static_cast<size_t>(2 * wave + half) * nv +
                                    static_cast<size_t>(t0) * kTileSyms + r * kTileDim;
    const uint4 a0 = ((GPtr<const uint4>)src)[0], a1 = ((GPtr<const uint4>)src)[1];
    const uint4 b0 = ((GPtr<const uint4>)(src + kTileSyms))[0], b1 = ((GPtr<const uint4>)(src + kTileSyms))[1];

    // Totals of the groups before each of the (at most 64) groups the rows touch:
    // lane i holds the sum over groups < g_min + i.
    const uint32_t ty0 = t0 / tiles_x, tx0 = t0 - ty0 * tiles_x;
    const uint32_t ty1 = t1 / tiles_x, tx1 = t1 - ty1 * tiles_x;
    const uint32_t g_min = (ty0 * kTileDim * bx + tx0 * kTileDim) / kGroupSyms;
    const uint32_t g_max = ((ty1 * kTileDim + kTileDim - 1) * bx + tx1 * kTileDim) / kGroupSyms;
    const bool few_groups = g_max - g_min < 64;
    int32_t gpre = 0;
    if (few_groups) {
      int32_t before = 0;
      for (uint32_t i = lane; i < g_min; i += 64) before += d.group_sums[i];
      before = WaveSum64(before);
      const int32_t gs = lane <= g_max - g_min ? d.group_sums[g_min + lane] : 0;
      int32_t incl = gs;
#pragma unroll
      for (uint32_t off = 1; off < 64; off <<= 1) {
        const int32_t n = __shfl_up(incl, off);
        if (lane >= off) incl += n;
      }
      gpre = before + incl - gs;
    }

    uint32_t oa[8], ob[8];
    TransformRowPair(a0, a1, b0, b1, L, r, oa, ob);
    // Tile (2 half + P) of this wave's space: row r at byte 32 r.
    uint4 *const da = reinterpret_cast<uint4 *>(&lds[wave][0][0] + (2 * half + 0) * (kTileSyms / 4) + r * 8);
    uint4 *const db = reinterpret_cast<uint4 *>(&lds[wave][0][0] + (2 * half + 1) * (kTileSyms / 4) + r * 8);
    da[0] = make_uint4(oa[0], oa[1], oa[2], oa[3]);
    da[1] = make_uint4(oa[4], oa[5], oa[6], oa[7]);
    db[0] = make_uint4(ob[0], ob[1], ob[2], ob[3]);
    db[1] = make_uint4(ob[4], ob[5], ob[6], ob[7]);
    LdsBarrier();

    const GPtr<const uint8_t> diffs = d.symbols + SymbolOffset(d, 3);
    const GPtr<const uint32_t> pal = (GPtr<const uint32_t>)(d.symbols + SymbolOffset(d, 2));
    const uint32_t words = d.palette_bytes / 4;
    const uint8_t *const tiles = reinterpret_cast<const uint8_t *>(&lds[0][0][0]);
    bool bad = false;
    // 64 tile rows (2 positions x 32) in row groups of 8: lane l takes blocks
    // 4 (l % 8) .. +3 of row l / 8 of its group, and wave w groups w, w + 3, w + 6
    // (< 8).  Each phase runs over all of a wave's groups, so their loads are in
    // flight together: diffs and row bases, then palette words, then the stores.
    constexpr uint32_t kGroupsOfRows = 2 * kTileDim / 8;
    constexpr uint32_t kSteps = (kGroupsOfRows + kTileFusedWaves - 1) / kTileFusedWaves;  // 3
    const uint32_t sub = lane >> 3, xb = 4 * (lane & 7);
    uint32_t blk[kSteps], dw[kSteps];
    int32_t base[kSteps];
#pragma unroll
    for (uint32_t i = 0; i < kSteps; ++i) {
      const uint32_t grp0 = wave + kTileFusedWaves * i;
      const uint32_t q = 8 * (grp0 < kGroupsOfRows ? grp0 : wave) + sub;  // a spare step repeats group `wave`
      const uint32_t P = q / kTileDim, y = q % kTileDim;
      const uint32_t ty = P ? ty1 : ty0, tx = P ? tx1 : tx0;
      const uint32_t row0 = (ty * kTileDim + y) * bx + tx * kTileDim;
      blk[i] = row0 + xb;
      dw[i] = *(GPtr<const uint32_t>)(diffs + blk[i]);
      const uint32_t g = row0 / kGroupSyms;
      int32_t before;
      if (few_groups) {
        before = __shfl(gpre, static_cast<int>(g - g_min));
      } else {
        before = 0;
        for (uint32_t k = xb / 4; k < g; k += 8) before += d.group_sums[k];
#pragma unroll
        for (int off = 4; off > 0; off >>= 1) before += __shfl_xor(before, off, 8);
      }
      base[i] = before + GroupLocalPrefix(d, row0);
    }
    uint32_t word[kSteps][4];
#pragma unroll
    for (uint32_t i = 0; i < kSteps; ++i) {
      int32_t run[4];
      int32_t acc = 0;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        acc += static_cast<int32_t>((dw[i] >> (8 * k)) & 0xFFu) - 128;
        run[k] = acc;
      }
      int32_t incl = acc;
#pragma unroll
      for (uint32_t off = 1; off < 8; off <<= 1) {
        const int32_t n = __shfl_up(incl, off, 8);
        if ((lane & 7) >= off) incl += n;
      }
      const int32_t start = base[i] + incl - acc;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t pi = static_cast<uint32_t>(start + run[k]);
        word[i][k] = 0;
        if (pi < words) word[i][k] = pal[pi]; else bad = true;
      }
    }
#pragma unroll
    for (uint32_t i = 0; i < kSteps; ++i) {
      const uint32_t grp = wave + kTileFusedWaves * i;
      if (grp >= kGroupsOfRows) continue;
      const uint32_t q = 8 * grp + sub;
      const uint32_t P = q / kTileDim, y = q % kTileDim;
      uint32_t pw[6];
#pragma unroll
      for (uint32_t p = 0; p < 6; ++p)
        pw[p] = *reinterpret_cast<const uint32_t *>(tiles + (p >> 1) * kWaveLdsBytes +
                                                    ((p & 1) * 2 + P) * kTileSyms + y * kTileDim + xb);
      uint32_t o[8];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        o[2 * k] = Rgb565Pair(pw[0], pw[1], pw[2], pw[4], pw[3], pw[5], k);
        o[2 * k + 1] = word[i][k];
      }
      const GPtr<uint4> dst = (GPtr<uint4>)(d.out + 8 * static_cast<size_t>(blk[i]));
      dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
      dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
    }
    if (bad) atomicOr(a.err, kErrPaletteIndex);
    LdsBarrier();  // the next unit's transposes reuse the tiles' space
  }
}

__global__ __launch_bounds__(256) void AssembleRgbKernel(KernelArgs a) {
  for (uint32_t u = blockIdx.x;; u += gridDim.x) {
    const Located loc = Locate(a, kStageAssemble, u, 1);
    if (!loc.valid) break;
    const DevDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const size_t b0 = static_cast<size_t>(loc.local) * kAssembleBlocksPerUnit + 4 * threadIdx.x;
    uint32_t pw[6];
#pragma unroll
    for (int p = 0; p < 6; ++p)
      pw[p] = *(GPtr<const uint32_t>)(d.planes + static_cast<size_t>(p) * nv + b0);
    const int4 ix = *(GPtr<const int4>)(d.indices + b0);
    const GPtr<const uint32_t> pal = (GPtr<const uint32_t>)(d.symbols + SymbolOffset(d, 2));
    const uint32_t words = d.palette_bytes / 4;
    const int idx[4] = {ix.x, ix.y, ix.z, ix.w};
    bool bad = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int pal8[4][3];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int y = SByte(pw[e], q), co = SByte(pw[2 + 2 * e], q), cg = SByte(pw[3 + 2 * e], q);
        const int t = y - cg / 2;
        const int g = cg + t;
        const int b = (t - co) / 2;
        const int r = b + co;
        pal8[e][0] = static_cast<int>((static_cast<uint32_t>(r) << 3) | static_cast<uint32_t>(r >> 2));
        pal8[e][1] = static_cast<int>((static_cast<uint32_t>(g) << 2) | static_cast<uint32_t>(g >> 4));
        pal8[e][2] = static_cast<int>((static_cast<uint32_t>(b) << 3) | static_cast<uint32_t>(b >> 2));
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        pal8[2][c] = (2 * pal8[0][c] + pal8[1][c]) / 3;
        pal8[3][c] = (pal8[0][c] + 2 * pal8[1][c]) / 3;
      }
      const uint32_t pi = static_cast<uint32_t>(idx[q]);
      uint32_t word = 0;
      if (pi < words) word = pal[pi]; else bad = true;
      const size_t blk = b0 + q;
      const uint32_t bxi = static_cast<uint32_t>(blk % d.bx), byi = static_cast<uint32_t>(blk / d.bx);
#pragma unroll
      for (int row = 0; row < 4; ++row) {
        uint32_t bytes[12];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int sel = (word >> (2 * (4 * row + c))) & 3;
          bytes[3 * c + 0] = static_cast<uint32_t>(pal8[sel][0]) & 0xFFu;
          bytes[3 * c + 1] = static_cast<uint32_t>(pal8[sel][1]) & 0xFFu;
          bytes[3 * c + 2] = static_cast<uint32_t>(pal8[sel][2]) & 0xFFu;
        }
        const GPtr<uint32_t> dst =
            (GPtr<uint32_t>)(d.out + 3 * (static_cast<size_t>(d.width) * (4 * byi + row) + 4 * bxi));
#pragma unroll
        for (int w = 0; w < 3; ++w)
          dst[w] = bytes[4 * w] | (bytes[4 * w + 1] << 8) | (bytes[4 * w + 2] << 16) |
                   (bytes[4 * w + 3] << 24);
      }
    }
    if (bad) atomicOr(a.err, kErrPaletteIndex);
  }
}

template <uint32_t WAVES, uint32_t GPW, bool kBuild = false, bool kPair = false>
hipError_t LaunchAns(const KernelArgs &args, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((AnsDecodeKernel<WAVES, GPW, kBuild, kPair>), dim3(grid), dim3(WAVES * 64), 0, stream, args);
  return hipGetLastError();
}

}  // namespace

hipError_t LaunchStage(uint32_t stage, const KernelArgs &args, uint32_t grid,
                       const StageConfig &cfg, hipStream_t stream) {
  if (grid == 0) return hipSuccess;
  switch (stage) {
    case kStageTables:
      if (cfg.ans_build_tables) return hipSuccess;  // the rANS kernel builds them in LDS
      hipLaunchKernelGGL(BuildTableKernel, dim3(grid), dim3(256), 0, stream, args);
      return hipGetLastError();
    case kStageAns: {
      const uint32_t w = cfg.ans_waves, g = cfg.ans_groups_per_wave;
      if (cfg.ans_auto) {
        // The form travels in the plan (or in args.ans_unit_code for a single image).
        if (!cfg.ans_build_tables || g != 2 || w != kAutoWaves) return hipErrorInvalidValue;
        hipLaunchKernelGGL(AnsAutoKernel, dim3(grid), dim3(kAutoWaves * 64), 0, stream, args);
        return hipGetLastError();
      }
      if (cfg.ans_pair_short) return hipErrorInvalidValue;  // paired units run in AnsAutoKernel
      if (cfg.ans_build_tables) {
        if (g == 2 && w == 4) return LaunchAns<4, 2, true>(args, grid, stream);
        if (g == 2 && w == 8) return LaunchAns<8, 2, true>(args, grid, stream);
        if (g == 1 && w == 4) return LaunchAns<4, 1, true>(args, grid, stream);
        if (g == 1 && w == 8) return LaunchAns<8, 1, true>(args, grid, stream);
        return hipErrorInvalidValue;
      }
      if (g == 2 && w == 1) return LaunchAns<1, 2>(args, grid, stream);
      if (g == 2 && w == 2) return LaunchAns<2, 2>(args, grid, stream);
      if (g == 2 && w == 4) return LaunchAns<4, 2>(args, grid, stream);
      if (g == 2 && w == 8) return LaunchAns<8, 2>(args, grid, stream);
      if (g == 1 && w == 2) return LaunchAns<2, 1>(args, grid, stream);
      if (g == 1 && w == 4) return LaunchAns<4, 1>(args, grid, stream);
      if (g == 1 && w == 8) return LaunchAns<8, 1>(args, grid, stream);
      return hipErrorInvalidValue;
    }
    case kStageIndices:
      if (cfg.fused_waves || cfg.tile_fused) return hipSuccess;  // the assembly scans them itself
      hipLaunchKernelGGL(DecodeIndicesKernel, dim3(grid), dim3(256), 0, stream, args);
      return hipGetLastError();
    case kStageWavelet:
      if (cfg.tile_fused) return hipSuccess;  // inside the assembly
      if (cfg.wavelet_waves == 1)
        hipLaunchKernelGGL(InvWaveletKernel<1>, dim3(grid), dim3(64), 0, stream, args);
      else if (cfg.wavelet_waves == 2)
        hipLaunchKernelGGL(InvWaveletKernel<2>, dim3(grid), dim3(128), 0, stream, args);
      else if (cfg.wavelet_waves == 4)
        hipLaunchKernelGGL(InvWaveletKernel<4>, dim3(grid), dim3(256), 0, stream, args);
      else
        return hipErrorInvalidValue;
      return hipGetLastError();
    case kStageAssemble:
      if (cfg.tile_fused) {
        if (cfg.rgb) return hipErrorInvalidValue;  // the fused kernel writes DXT1 only
        hipLaunchKernelGGL(WaveletAssembleKernel, dim3(grid), dim3(kTileFusedWaves * 64), 0, stream, args);
      } else if (cfg.rgb) {
        if (cfg.fused_waves) return hipErrorInvalidValue;  // the fused scan writes DXT1 only
        hipLaunchKernelGGL(AssembleRgbKernel, dim3(grid), dim3(256), 0, stream, args);
      } else if (cfg.fused_waves && cfg.fused_wide) {
        if (cfg.fused_waves == 1)
          hipLaunchKernelGGL(AssembleFusedWideKernel<1>, dim3(grid), dim3(64), 0, stream, args);
        else if (cfg.fused_waves == 2)
          hipLaunchKernelGGL(AssembleFusedWideKernel<2>, dim3(grid), dim3(128), 0, stream, args);
        else if (cfg.fused_waves == 4)
          hipLaunchKernelGGL(AssembleFusedWideKernel<4>, dim3(grid), dim3(256), 0, stream, args);
        else
          return hipErrorInvalidValue;
      } else if (cfg.fused_waves == 1) {
        hipLaunchKernelGGL(AssembleFusedKernel<1>, dim3(grid), dim3(64), 0, stream, args);
      } else if (cfg.fused_waves == 2) {
        hipLaunchKernelGGL(AssembleFusedKernel<2>, dim3(grid), dim3(128), 0, stream, args);
      } else if (cfg.fused_waves == 4) {
        hipLaunchKernelGGL(AssembleFusedKernel<4>, dim3(grid), dim3(256), 0, stream, args);
      } else if (cfg.fused_waves == 0) {
        hipLaunchKernelGGL(AssembleDxtKernel, dim3(grid), dim3(256), 0, stream, args);
      } else {
        return hipErrorInvalidValue;
      }
      return hipGetLastError();
    default:
      return hipErrorInvalidValue;
  }
}

// One work-group row per job: 16-byte loads grid-stride over the job's bytes (four
// pairs in flight per lane, so a large output is not one load round trip per 256 KiB),
// one ballot per wave, and a flag store only where a difference was seen.
__global__ __launch_bounds__(256) void CompareKernel(CompareArgs a) {
  const CompareJob j = a.job[blockIdx.y];
  const GPtr<const uint4> x = Gbl<const uint4>(j.out);
  const GPtr<const uint4> y = Gbl<const uint4>(j.expected);
  const uint64_t words = j.bytes / 16;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  bool diff = false;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < words; i += 4 * stride) {
    uint4 p[4], q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      p[k] = x[i + k * stride];
      q[k] = y[i + k * stride];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      diff |= ((p[k].x ^ q[k].x) | (p[k].y ^ q[k].y) | (p[k].z ^ q[k].z) | (p[k].w ^ q[k].w)) != 0;
  }
  for (; i < words; i += stride) {
    const uint4 p = x[i], q = y[i];
    diff |= ((p.x ^ q.x) | (p.y ^ q.y) | (p.z ^ q.z) | (p.w ^ q.w)) != 0;
  }
  if (__builtin_amdgcn_ballot_w64(diff) != 0 && __lane_id() == 0) Gbl<uint32_t>(a.flags)[j.flag] = 1u;
}

hipError_t LaunchCompare(const CompareArgs &a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  if (a.n > kCompareJobs) return hipErrorInvalidValue;
  hipLaunchKernelGGL(CompareKernel, dim3(64, a.n), dim3(256), 0, stream, a);
  return hipGetLastError();
}

// Work-group row per job, 16-byte words grid-stride with four loads in flight per
// lane before their stores (reads cross PCIe: latency wants many requests out); the
// job's last bytes past a whole word go one at a time.
__global__ __launch_bounds__(256) void GatherKernel(GatherArgs a) {
  const GatherJob j = a.job[blockIdx.y];
  const GPtr<const uint4> src = Gbl<const uint4>(j.src);
  const GPtr<uint4> dst = Gbl<uint4>(j.dst);
  const uint64_t words = j.bytes / 16;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < words; i += 4 * stride) {
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = src[i + k * stride];
#pragma unroll
    for (int k = 0; k < 4; ++k) dst[i + k * stride] = v[k];
  }
  for (; i < words; i += stride) dst[i] = src[i];
  if (blockIdx.x == 0 && threadIdx.x < (j.bytes & 15)) {
    const uint64_t b = words * 16 + threadIdx.x;
    Gbl<uint8_t>(j.dst)[b] = Gbl<const uint8_t>(j.src)[b];
  }
}

hipError_t LaunchGather(const GatherArgs &a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  if (a.n > kGatherJobs) return hipErrorInvalidValue;
  uint64_t most = 0;
  for (uint32_t i = 0; i < a.n; ++i) {
    if ((reinterpret_cast<uintptr_t>(a.job[i].src) | reinterpret_cast<uintptr_t>(a.job[i].dst)) & 15)
      return hipErrorInvalidValue;
    most = std::max(most, a.job[i].bytes);
  }
  // A few work-groups keep the link busy (each has 16 KiB of loads in flight; about
  // 2 us x 55 GB/s is needed), so the gather takes at most 128 work-groups in all and
  // leaves the CUs to the decodes it runs beside: at most 128 / n a job.  A small job gets
  // one work-group per 4 KiB (a word a lane), so all its loads go out at once: one link
  // round trip, shorter than a DMA copy's fixed cost.
  const uint64_t per_group = 4 << 10;
  const uint32_t gx = static_cast<uint32_t>(
      std::max<uint64_t>(1, std::min<uint64_t>(128 / a.n, (most + per_group - 1) / per_group)));
  hipLaunchKernelGGL(GatherKernel, dim3(gx, a.n), dim3(256), 0, stream, a);
  return hipGetLastError();
}

namespace {

// The no-allocation entry point's scratch.  In host code `scratch` names a host-side
// stand-in, not device memory: host code uses only the address ResolveNoAllocScratch
// obtained for each device.
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

hipError_t ResolveNoAllocScratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!p) return hipErrorInvalidSymbol;
  g_scratch_ptr[device] = static_cast<uint8_t *>(p);
  return hipSuccess;
}

uint8_t *NoAllocScratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

hipError_t LaunchAnsGroups(const uint32_t *table, const uint8_t *data, uint32_t data_len, uint32_t groups,
                           uint32_t n, uint8_t *out, uint32_t *err, hipStream_t stream) {
  if (groups == 0) return hipSuccess;
  if (n == 0 || n > kLanesPerGroup) return hipErrorInvalidValue;
  hipLaunchKernelGGL(AnsGroupsKernel, dim3(groups), dim3(64), 0, stream, table, data, data_len, n, out, err);
  return hipGetLastError();
}

}  // namespace gst
