// gst_kernels.hip -- the five GST decode stages as gfx950 kernels (wave64, integer only).
//
// Every kernel is a grid-stride loop over "units" of its stage; a unit maps to
// (image, local unit) through the batch plan's per-stage prefix array, or to the
// single image passed by value.  Reference semantics are cited per kernel.
#include "gst_internal.h"

namespace gst {
namespace {

constexpr int32_t kBufferRsrcWord3 = 0x00020000;  // raw buffer, gfx9 family

__device__ __forceinline__ uint32_t StageTotal(const KernelArgs &a, uint32_t stage,
                                               uint32_t groups_per_unit) {
  return a.plan ? a.plan->units[stage] : StageUnitsOf(a.single, stage, groups_per_unit);
}

struct Located {
  uint32_t img;
  uint32_t local;
};

// Largest image i with prefix[i] <= u (every image has at least one unit per stage).
__device__ __forceinline__ Located Locate(const KernelArgs &a, uint32_t stage, uint32_t u) {
  if (!a.plan) return {0u, u};
  const uint32_t n = a.plan->n_images;
  const uint32_t *pre = reinterpret_cast<const uint32_t *>(reinterpret_cast<const uint8_t *>(a.plan) +
                                                           kPlanHeaderBytes) + stage * (n + 1);
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= u) lo = mid; else hi = mid;
  }
  return {lo, u - pre[lo]};
}

__device__ __forceinline__ ImageDesc LoadDesc(const KernelArgs &a, uint32_t img) {
  if (!a.plan) return a.single;
  return reinterpret_cast<const ImageDesc *>(reinterpret_cast<const uint8_t *>(a.plan) +
                                             PlanImagesOffset(a.plan->n_images))[img];
}

// Element k of a 4-entry descriptor array without dynamic indexing (which would put
// the descriptor copy in scratch memory).
template <typename T>
__device__ __forceinline__ T Pick4(const T (&v)[4], uint32_t k) {
  return k == 0 ? v[0] : k == 1 ? v[1] : k == 2 ? v[2] : v[3];
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
    const ImageDesc d = LoadDesc(a, img);
    const uint32_t tix = d.table_base + k;
    const uint16_t *fr = reinterpret_cast<const uint16_t *>(d.freqs + 512 * k);
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
    uint32_t *tab = a.tables + static_cast<size_t>(tix) * kTableSize;
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
    if (t == 0) a.table_status[tix] = valid ? 1u : 0u;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stage 2: interleaved rANS decode (ans/ans_decode.cl:97-157).
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
constexpr uint32_t kChunkWords = 256;
constexpr uint32_t kRingWords = 2 * kChunkWords;
constexpr uint32_t kRefillSteps = kChunkWords / kLanesPerGroup;  // 8

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Lane j's 16 bytes of chunk c (words 8j .. 8j+7); chunks below the stream start
// wrap to huge offsets and read zeros.
__device__ __forceinline__ u32x4 LoadChunk(__amdgpu_buffer_rsrc_t rs, uint32_t c, uint32_t j) {
  return __builtin_amdgcn_raw_buffer_load_b128(rs, static_cast<int>((c * kChunkWords + 8 * j) * 2), 0, 0);
}

__device__ __forceinline__ void StoreChunk(uint16_t *ring, uint32_t c, uint32_t j, const u32x4 &a) {
  *reinterpret_cast<u32x4 *>(ring + ((c * kChunkWords + 8 * j) & (kRingWords - 1))) = a;
}

template <bool kSum, uint32_t GPW>
__device__ __forceinline__ void DecodeHalfGroups(const uint32_t *tab, uint16_t *ring_half,
                                                 __amdgpu_buffer_rsrc_t rs, uint32_t lane, bool active,
                                                 uint32_t x, uint32_t pA, uint32_t pB, uint32_t loA,
                                                 uint32_t loB, uint8_t *outp, int32_t *group_sum,
                                                 int32_t *lane_sum,
                                                 uint32_t *err) {
  const uint32_t j = lane & 31;
  uint32_t upper = lane < 32 ? 0u : 0xFFFFFFFFu;
  asm volatile("" : "+v"(upper));  // keep the half-wave mask an opaque VGPR (one v_and below)
  const uint64_t active_mask = __builtin_amdgcn_ballot_w64(active);
  // Ring prologue: chunks top and top-1 resident, top-2 and top-3 in flight.
  uint32_t topA = pA ? (pA - 1) / kChunkWords : 0;
  uint32_t topB = pB ? (pB - 1) / kChunkWords : 0;
  const uint32_t top = lane < 32 ? topA : topB;
  {
    const u32x4 c0 = LoadChunk(rs, top, j);
    const u32x4 c1 = LoadChunk(rs, top - 1, j);
    if (GPW == 2 || lane < 32) {  // one group per wave: the idle upper half shares the ring
      StoreChunk(ring_half, top, j, c0);
      StoreChunk(ring_half, top - 1, j, c1);
    }
  }
  u32x4 pend0 = LoadChunk(rs, top - 2, j);
  u32x4 pend1 = LoadChunk(rs, top - 3, j);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  int32_t sum = 0;
  // 64 steps fill one 64-byte sector of the lane's output; whole sectors leave in four
  // back-to-back 16-byte stores (smaller pieces of a lane's segment, 256 B apart
  // across lanes, are evicted half-written and written back several times).
  for (uint32_t sector = 0; sector < kSymsPerLane / 64; ++sector) {
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
        const uint64_t m = __builtin_amdgcn_uicmp(x, kStateL, 36 /* ULT */) & active_mask;
        const uint32_t mlo = static_cast<uint32_t>(m), mhi = static_cast<uint32_t>(m >> 32);
        const uint32_t tlo = __popc(mlo), thi = __popc(mhi);
        pA -= tlo;
        pB -= thi;
        if (active && x < kStateL) {
          // Word index p_half + rank: mbcnt over both halves counts the lower half's
          // bits for upper lanes, so upper lanes start from pB - tlo.
          const uint32_t b = pA + (upper & (pB - tlo - pA));
          const uint32_t widx = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, b));
          x = (x << 16) | ring_half[widx & (kRingWords - 1)];
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
      if (crossA || crossB) {
        if (lane < 32 ? crossA : crossB) {
          const uint32_t t = lane < 32 ? topA : topB;
          StoreChunk(ring_half, t - 2, j, pend0);
          pend0 = pend1;
          pend1 = LoadChunk(rs, t - 4, j);
        }
        topA -= crossA ? 1u : 0u;
        topB -= crossB ? 1u : 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    if (kSum) {
      // Byte sum of the sector (v_sad_u8 against zero), less the 64 biases.
      uint32_t bytes = 0;
#pragma unroll
      for (uint32_t i = 0; i < 16; ++i) bytes = __builtin_amdgcn_sad_u8(acc[i], 0u, bytes);
      sum += static_cast<int32_t>(bytes) - 64 * 128;
    }
    if (active) {
      uint4 *dst = reinterpret_cast<uint4 *>(outp + (kSymsPerLane - 64) - 64 * sector);
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i)
        dst[i] = make_uint4(acc[4 * i], acc[4 * i + 1], acc[4 * i + 2], acc[4 * i + 3]);
    }
  }
  const bool under = lane < 32 ? (pA < loA) : (pB < loB);
  if (active && under && j == 0) atomicOr(err, kErrStreamOffsets);
  if (kSum) {
    if (active && lane_sum) lane_sum[j] = sum;
    sum = WaveSumWidth32(sum);
    if (active && j == 0) *group_sum = sum;
  }
}

// readfirstlane returns a signed int: every use converts through uint32_t so no
// 64-bit widening can sign-extend a value with bit 31 set.
__device__ __forceinline__ uint32_t Uniform(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

__device__ __forceinline__ const uint8_t *Uniform(const uint8_t *p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint64_t lo = Uniform(static_cast<uint32_t>(v));
  const uint64_t hi = Uniform(static_cast<uint32_t>(v >> 32));
  return reinterpret_cast<const uint8_t *>(lo | (hi << 32));
}

// The work of BuildTableKernel for one table, straight into the rANS kernel's LDS
// arrays (256 <= THREADS): scan of the 256 frequencies by the first four waves, then
// every slot's symbol by the same binary lifting.  A table that does not sum to 2048
// becomes the identity table.  Returns whether the table is valid; ends with a
// work-group barrier.
template <uint32_t THREADS>
__device__ __forceinline__ bool BuildTableInLds(const uint16_t *fr, uint32_t *tab, uint32_t *cum, uint32_t *freq,
                                                uint32_t *wsum) {
  static_assert(THREADS >= 256, "one thread per frequency");
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  uint32_t f = 0, v = 0;
  if (t < 256) {
    f = fr[t];
    v = f;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t n = __shfl_up(v, off);
      if (lane >= static_cast<uint32_t>(off)) v += n;
    }
    if (lane == 63) wsum[wave] = v;
  }
  __syncthreads();
  if (t < 256) {
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += wsum[w];
    cum[t] = v + add - f;
    freq[t] = f;
  }
  const bool valid = wsum[0] + wsum[1] + wsum[2] + wsum[3] == kTableSize;
  __syncthreads();
#pragma unroll 1
  for (uint32_t slot = t; slot < kTableSize; slot += THREADS) {
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
      tab[slot] = freq[x] | ((slot - cum[x]) << 12) | (x << 24);
    } else {
      tab[slot] = kTableSize | (slot << 12);
    }
  }
  __syncthreads();
  return valid;
}

// kBuild: build the stream's table from its frequencies in LDS instead of loading
// the one BuildTableKernel wrote (the decode pipeline then skips that launch).
template <uint32_t WAVES, uint32_t GPW, bool kBuild>
// With 4 or 8 waves LDS allows 7 waves per SIMD; hold the VGPRs to that (at most 72).
__global__ __launch_bounds__(WAVES * 64, WAVES >= 4 ? 7 : 1)
void AnsDecodeKernel(KernelArgs a) {
  __shared__ uint32_t tab[kTableSize];
  __shared__ uint16_t ring[WAVES][2][kRingWords];
  __shared__ uint32_t build_cum[kBuild ? 256 : 1];
  __shared__ uint32_t build_freq[kBuild ? 256 : 1];
  __shared__ uint32_t build_wsum[4];
  constexpr uint32_t kGroupsPerUnit = WAVES * GPW;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t half = GPW == 2 ? (lane >> 5) : 0;
  const uint32_t j = lane & 31;
  const uint32_t total = StageTotal(a, kStageAns, kGroupsPerUnit);
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    const Located loc = Locate(a, kStageAns, u);
    const ImageDesc d = LoadDesc(a, loc.img);
    uint32_t k = 0, local = loc.local;
    for (; k < 3; ++k) {
      const uint32_t uk = (Pick4(d.groups, k) + kGroupsPerUnit - 1) / kGroupsPerUnit;
      if (local < uk) break;
      local -= uk;
    }
    k = Uniform(k);
    local = Uniform(local);
    // The group's offsets are requested first, so their latency overlaps the table's.
    const uint32_t G = Uniform(Pick4(d.groups, k));
    const uint32_t g = local * kGroupsPerUnit + wave * GPW + half;
    const uint32_t slen = Uniform(Pick4(d.stream_len, k));
    const __amdgpu_buffer_rsrc_t rs = MakeRsrc(Uniform(Pick4(d.streams, k)), slen);
    bool active = (GPW == 2 || lane < 32) && g < G;
    uint32_t end = 0, lo = 0;
    if (active) {
      end = __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(4 * g), 0, 0);
      lo = g == 0 ? 4 * G : __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(4 * (g - 1)), 0, 0);
    }
    __syncthreads();
    bool table_ok;
    if constexpr (kBuild) {
      const uint16_t *fr = reinterpret_cast<const uint16_t *>(Uniform(d.freqs + 512 * k));
      table_ok = BuildTableInLds<WAVES * 64>(fr, tab, build_cum, build_freq, build_wsum);
    } else {
      const uint32_t tix = Uniform(d.table_base + k);
      const uint32_t *src = a.tables + static_cast<size_t>(tix) * kTableSize;
      for (uint32_t i = threadIdx.x; i < kTableSize; i += WAVES * 64) tab[i] = src[i];
      table_ok = a.table_status[tix] != 0;
      __syncthreads();
    }
    if (!table_ok && threadIdx.x == 0) atomicOr(a.err, kErrFrequencies);

    if (active) {
      const bool ok = ((end | lo) & 3) == 0 && end <= slen && lo >= 4 * G && lo <= end &&
                      end - lo >= 4 * kLanesPerGroup;
      if (!ok) {
        if (j == 0) atomicOr(a.err, kErrStreamOffsets);
        active = false;
      }
    }
    const uint32_t x = active
        ? __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(end - 4 * kLanesPerGroup + 4 * j), 0, 0)
        : 0u;
    const uint32_t p = active ? (end - 4 * kLanesPerGroup) >> 1 : 0u;
    const uint32_t plo = active ? lo >> 1 : 0u;
    const uint32_t pA = __builtin_amdgcn_readlane(p, 0);
    const uint32_t loA = __builtin_amdgcn_readlane(plo, 0);
    const uint32_t pB = GPW == 2 ? __builtin_amdgcn_readlane(p, 32) : 0u;
    const uint32_t loB = GPW == 2 ? __builtin_amdgcn_readlane(plo, 32) : 0u;
    uint8_t *outp = d.symbols + SymbolOffset(d, k) + static_cast<size_t>(g) * kGroupSyms + j * kSymsPerLane;
    uint16_t *ring_half = ring[wave][half];
    if (k == 3)
      DecodeHalfGroups<true, GPW>(tab, ring_half, rs, lane, active, x, pA, pB, loA, loB, outp,
                                  d.group_sums + g,
                                  d.lane_sums ? d.lane_sums + g * kLanesPerGroup : nullptr, a.err);
    else
      DecodeHalfGroups<false, GPW>(tab, ring_half, rs, lane, active, x, pA, pB, loA, loB, outp, nullptr,
                                   nullptr, a.err);
  }
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
  const uint32_t total = StageTotal(a, kStageIndices, 1);
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    const Located loc = Locate(a, kStageIndices, u);
    const ImageDesc d = LoadDesc(a, loc.img);
    const uint32_t g = loc.local;
    if (wave == 0) {
      uint32_t s = 0;
      for (uint32_t i = lane; i < g; i += 64) s += static_cast<uint32_t>(d.group_sums[i]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) gprefix = s;
    }
    const uint8_t *diffs = d.symbols + SymbolOffset(d, 3) + static_cast<size_t>(g) * kGroupSyms + 32 * t;
    const uint4 q0 = reinterpret_cast<const uint4 *>(diffs)[0];
    const uint4 q1 = reinterpret_cast<const uint4 *>(diffs)[1];
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
    uint4 *dst = reinterpret_cast<uint4 *>(d.indices + static_cast<size_t>(g) * kGroupSyms + 32 * t);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      dst[i] = make_uint4(run[4 * i] + offset, run[4 * i + 1] + offset, run[4 * i + 2] + offset,
                          run[4 * i + 3] + offset);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Stage 4: inverse 5/3 wavelet (codec/inverse_wavelet.cl:98-276).
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

// C's truncating division of each half: negative values take divisor - 1 first.
__device__ __forceinline__ s16x2 TruncDiv4(s16x2 s) {
  const s16x2 neg = s >> static_cast<s16x2>(15);  // 0 or -1
  return (s + neg * static_cast<s16x2>(-3)) >> static_cast<s16x2>(2);
}
__device__ __forceinline__ s16x2 TruncDiv2(s16x2 s) {
  const u16x2 sign = __builtin_bit_cast(u16x2, s) >> static_cast<u16x2>(15);  // 0 or 1
  return (s + __builtin_bit_cast(s16x2, sign)) >> static_cast<s16x2>(1);
}

__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
  const uint8_t *src;  // tile-linear symbols (1024 bytes)
  int8_t *dst;         // top-left of the tile in the raster plane
  uint32_t pitch;      // plane row length (blocks)
};

__device__ __forceinline__ TileRef TileAt(const ImageDesc &d, uint32_t tile) {
  const uint32_t per_plane = d.nv / kTileSyms, tiles_x = d.bx / kTileDim;
  const uint32_t plane = tile / per_plane, t = tile - plane * per_plane;
  const uint32_t ty = t / tiles_x, tx = t - ty * tiles_x;
  const size_t base = static_cast<size_t>(plane) * d.nv;
  return {d.symbols + base + static_cast<size_t>(t) * kTileSyms,
          d.planes + base + static_cast<size_t>(ty) * kTileDim * d.bx + tx * kTileDim, d.bx};
}

// Tile `half` of wavelet unit `u` (uniform address math, one select per lane).
__device__ __forceinline__ TileRef WaveletTile(const KernelArgs &a, uint32_t u, uint32_t half) {
  const Located loc = Locate(a, kStageWavelet, u);
  const ImageDesc d = LoadDesc(a, loc.img);
  const TileRef t0 = TileAt(d, loc.local * kWaveletTilesPerUnit);
  const TileRef t1 = TileAt(d, loc.local * kWaveletTilesPerUnit + 1);
  return half ? t1 : t0;
}

template <uint32_t WAVES>
__global__ __launch_bounds__(WAVES * 64) void InvWaveletKernel(KernelArgs a) {
  __shared__ uint32_t lds[WAVES][2][kTileDim * kLdsStride];
  const uint32_t lane = threadIdx.x & 63, half = lane >> 5, r = lane & 31;
  const uint32_t wave = Uniform(threadIdx.x >> 6);
  const uint32_t total = StageTotal(a, kStageWavelet, 1);
  uint32_t *L = lds[wave][half];
  const uint32_t stride = gridDim.x * WAVES;
  // The rows of the next pair are requested before the current pair is transformed,
  // so a wave's loads overlap its own lifting when it takes more than one pair.
  uint32_t pair = blockIdx.x * WAVES + wave;
  TileRef nlo{}, nhi{};
  bool nhas_hi = false;
  uint4 na0{}, na1{}, nb0{}, nb1{};
  if (2 * pair < total) {
    nlo = WaveletTile(a, 2 * pair, half);
    nhas_hi = 2 * pair + 1 < total;
    nhi = nhas_hi ? WaveletTile(a, 2 * pair + 1, half) : nlo;
    na0 = reinterpret_cast<const uint4 *>(nlo.src + r * kTileDim)[0];
    na1 = reinterpret_cast<const uint4 *>(nlo.src + r * kTileDim)[1];
    if (nhas_hi) {
      nb0 = reinterpret_cast<const uint4 *>(nhi.src + r * kTileDim)[0];
      nb1 = reinterpret_cast<const uint4 *>(nhi.src + r * kTileDim)[1];
    }
  }
  for (; 2 * pair < total; pair += stride) {
    const TileRef lo = nlo, hi = nhi;
    const bool has_hi = nhas_hi;
    const uint4 a0 = na0, a1 = na1;
    const uint4 b0 = nhas_hi ? nb0 : make_uint4(0, 0, 0, 0), b1 = nhas_hi ? nb1 : make_uint4(0, 0, 0, 0);
    const uint32_t next = pair + stride;
    if (2 * next < total) {
      nlo = WaveletTile(a, 2 * next, half);
      nhas_hi = 2 * next + 1 < total;
      nhi = nhas_hi ? WaveletTile(a, 2 * next + 1, half) : nlo;
      na0 = reinterpret_cast<const uint4 *>(nlo.src + r * kTileDim)[0];
      na1 = reinterpret_cast<const uint4 *>(nlo.src + r * kTileDim)[1];
      if (nhas_hi) {
        nb0 = reinterpret_cast<const uint4 *>(nhi.src + r * kTileDim)[0];
        nb1 = reinterpret_cast<const uint4 *>(nhi.src + r * kTileDim)[1];
      }
    }
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
    uint32_t oa[8], ob[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      // Low bytes of four pairs: t01 = (A0, A1, B0, B1), t23 = (A2, A3, B2, B3).
      const uint32_t t01 = __builtin_amdgcn_perm(Bits(v[4 * w + 1]), Bits(v[4 * w]), 0x06020400u);
      const uint32_t t23 = __builtin_amdgcn_perm(Bits(v[4 * w + 3]), Bits(v[4 * w + 2]), 0x06020400u);
      oa[w] = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
      ob[w] = __builtin_amdgcn_perm(t23, t01, 0x07060302u);
    }
    uint4 *da = reinterpret_cast<uint4 *>(lo.dst + static_cast<size_t>(r) * lo.pitch);
    da[0] = make_uint4(oa[0], oa[1], oa[2], oa[3]);
    da[1] = make_uint4(oa[4], oa[5], oa[6], oa[7]);
    if (has_hi) {
      uint4 *db = reinterpret_cast<uint4 *>(hi.dst + static_cast<size_t>(r) * hi.pitch);
      db[0] = make_uint4(ob[0], ob[1], ob[2], ob[3]);
      db[1] = make_uint4(ob[4], ob[5], ob[6], ob[7]);
    }
  }
}

// ---------------------------------------------------------------------------
// Stage 5: DXT1 assembly (codec/assemble.cl:290-357) and RGB8 assembly (:359-406).
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
  const uint32_t total = StageTotal(a, kStageAssemble, 1);
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    const Located loc = Locate(a, kStageAssemble, u);
    const ImageDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const size_t b0 = static_cast<size_t>(loc.local) * kAssembleBlocksPerUnit + 4 * threadIdx.x;
    uint32_t pw[6];
#pragma unroll
    for (int p = 0; p < 6; ++p)
      pw[p] = *reinterpret_cast<const uint32_t *>(d.planes + static_cast<size_t>(p) * nv + b0);
    const int4 ix = *reinterpret_cast<const int4 *>(d.indices + b0);
    const uint32_t *pal = reinterpret_cast<const uint32_t *>(d.symbols + SymbolOffset(d, 2));
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
    uint4 *dst = reinterpret_cast<uint4 *>(d.out + 8 * b0);
    dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
    dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
    if (bad) atomicOr(a.err, kErrPaletteIndex);
  }
}

// Index scan fused into DXT1 assembly: one wave per 1024-block unit, which covers
// four 256-block segments of the index stream (one rANS lane each).  The value
// before the unit is the sum of the earlier groups' totals and the earlier lanes'
// totals of its group, both written by the rANS stage; a segment is a 4-per-lane
// register prefix plus a wave scan.  The palette indices never go to memory, so a
// block costs 1 diff byte instead of 1 + 4 + 4 bytes of index traffic.  Results are
// those of DecodeIndicesKernel followed by AssembleDxtKernel.
__device__ __forceinline__ int32_t WaveSum64(int32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <uint32_t WAVES>
__global__ __launch_bounds__(WAVES * 64) void AssembleFusedKernel(KernelArgs a) {
  constexpr uint32_t kSegments = kAssembleBlocksPerUnit / kSymsPerLane;  // 4
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = Uniform(threadIdx.x >> 6);
  const uint32_t total = StageTotal(a, kStageAssemble, 1);
  for (uint32_t u = blockIdx.x * WAVES + wave; u < total; u += gridDim.x * WAVES) {
    const Located loc = Locate(a, kStageAssemble, u);
    const ImageDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const uint32_t b0 = loc.local * kAssembleBlocksPerUnit;
    const uint32_t g = b0 / kGroupSyms;
    const uint32_t seg0 = (b0 % kGroupSyms) / kSymsPerLane;
    int32_t base = 0;
    for (uint32_t i = lane; i < g; i += 64) base += d.group_sums[i];
    if (lane < seg0) base += d.lane_sums[g * kLanesPerGroup + lane];
    base = WaveSum64(base);
    const uint8_t *diffs = d.symbols + SymbolOffset(d, 3);
    const uint32_t *pal = reinterpret_cast<const uint32_t *>(d.symbols + SymbolOffset(d, 2));
    const uint32_t words = d.palette_bytes / 4;
    bool bad = false;
#pragma unroll
    for (uint32_t seg = 0; seg < kSegments; ++seg) {
      const uint32_t blk = b0 + seg * kSymsPerLane + 4 * lane;
      const uint32_t dw = *reinterpret_cast<const uint32_t *>(diffs + blk);
      uint32_t pw[6];
#pragma unroll
      for (int p = 0; p < 6; ++p)
        pw[p] = *reinterpret_cast<const uint32_t *>(d.planes + static_cast<size_t>(p) * nv + blk);
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
      uint4 *dst = reinterpret_cast<uint4 *>(d.out + 8 * static_cast<size_t>(blk));
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
  const uint32_t total = StageTotal(a, kStageAssemble, 1);
  for (uint32_t u = blockIdx.x * WAVES + wave; u < total; u += gridDim.x * WAVES) {
    const Located loc = Locate(a, kStageAssemble, u);
    const ImageDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const uint32_t b0 = loc.local * kAssembleBlocksPerUnit;
    const uint32_t blk = b0 + kPerLane * lane;
    const uint4 dv = *reinterpret_cast<const uint4 *>(d.symbols + SymbolOffset(d, 3) + blk);
    uint4 pv[6];
#pragma unroll
    for (int p = 0; p < 6; ++p)
      pv[p] = *reinterpret_cast<const uint4 *>(d.planes + static_cast<size_t>(p) * nv + blk);
    const uint32_t g = b0 / kGroupSyms;
    const uint32_t seg0 = (b0 % kGroupSyms) / kSymsPerLane;
    int32_t base = 0;
    for (uint32_t i = lane; i < g; i += 64) base += d.group_sums[i];
    if (lane < seg0) base += d.lane_sums[g * kLanesPerGroup + lane];
    base = WaveSum64(base);
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
    const uint32_t *pal = reinterpret_cast<const uint32_t *>(d.symbols + SymbolOffset(d, 2));
    const uint32_t words = d.palette_bytes / 4;
    bool bad = false;
    uint4 *dst = reinterpret_cast<uint4 *>(d.out + 8 * static_cast<size_t>(blk));
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

__global__ __launch_bounds__(256) void AssembleRgbKernel(KernelArgs a) {
  const uint32_t total = StageTotal(a, kStageAssemble, 1);
  for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
    const Located loc = Locate(a, kStageAssemble, u);
    const ImageDesc d = LoadDesc(a, loc.img);
    const uint32_t nv = d.nv;
    const size_t b0 = static_cast<size_t>(loc.local) * kAssembleBlocksPerUnit + 4 * threadIdx.x;
    uint32_t pw[6];
#pragma unroll
    for (int p = 0; p < 6; ++p)
      pw[p] = *reinterpret_cast<const uint32_t *>(d.planes + static_cast<size_t>(p) * nv + b0);
    const int4 ix = *reinterpret_cast<const int4 *>(d.indices + b0);
    const uint32_t *pal = reinterpret_cast<const uint32_t *>(d.symbols + SymbolOffset(d, 2));
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
        uint32_t *dst = reinterpret_cast<uint32_t *>(
            d.out + 3 * (static_cast<size_t>(d.width) * (4 * byi + row) + 4 * bxi));
#pragma unroll
        for (int w = 0; w < 3; ++w)
          dst[w] = bytes[4 * w] | (bytes[4 * w + 1] << 8) | (bytes[4 * w + 2] << 16) |
                   (bytes[4 * w + 3] << 24);
      }
    }
    if (bad) atomicOr(a.err, kErrPaletteIndex);
  }
}

template <uint32_t WAVES, uint32_t GPW, bool kBuild = false>
hipError_t LaunchAns(const KernelArgs &args, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((AnsDecodeKernel<WAVES, GPW, kBuild>), dim3(grid), dim3(WAVES * 64), 0, stream, args);
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
      if (cfg.fused_waves) return hipSuccess;  // the assembly scans the indices itself
      hipLaunchKernelGGL(DecodeIndicesKernel, dim3(grid), dim3(256), 0, stream, args);
      return hipGetLastError();
    case kStageWavelet:
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
      if (cfg.rgb) {
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

}  // namespace gst
