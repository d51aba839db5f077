// gst_internal.h -- descriptors shared by the host orchestration and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gst {

constexpr uint32_t kTableSize = 2048;        // M (ans/ans_ocl_encode.cpp:10-18)
constexpr uint32_t kTableLog = 11;
constexpr uint32_t kStateL = 32768;          // k*M (ans/ans_decode.cl:6-8)
constexpr uint32_t kLanesPerGroup = 32;      // streams per rANS group (ans/ans.h:75)
constexpr uint32_t kSymsPerLane = 256;       // ans/ans.h:74
constexpr uint32_t kGroupSyms = kLanesPerGroup * kSymsPerLane;
constexpr uint32_t kTileDim = 32;            // codec/codec_base.h:22
constexpr uint32_t kTileSyms = kTileDim * kTileDim;
constexpr uint32_t kAssembleBlocksPerUnit = 1024;
constexpr uint32_t kWaveletTilesPerUnit = 2;  // a tile per half-wave; a wave takes two units

enum Stage : uint32_t {
  kStageTables = 0,
  kStageAns = 1,
  kStageIndices = 2,
  kStageWavelet = 3,
  kStageAssemble = 4,
  kNumStages = 5
};

enum DevErr : uint32_t {
  kErrStreamOffsets = 0x1,
  kErrFrequencies = 0x2,
  kErrPaletteIndex = 0x4,
};

// Per-image descriptor.  Stream k: 0 = Y, 1 = chroma, 2 = palette, 3 = indices
// (codec/encoder.cpp:122-144).  Symbol buffer layout (codec/decoder.cpp:441-461):
// [Y 2nv][chroma 4nv][palette P][indices nv]; planes p = Y1,Y2,Co1,Cg1,Co2,Cg2 are the
// first 6nv symbols (codec/assemble.cl:27-37).
struct ImageDesc {
  uint32_t width, height, nv, bx;
  uint32_t palette_bytes;
  uint32_t groups[4];
  uint32_t stream_len[4];
  uint32_t table_base;   // index of the image's first of 4 tables
  uint32_t rgb;          // 1: RGB8 output
  const uint8_t *freqs;  // 4 x 256 little-endian u16
  const uint8_t *streams[4];
  uint8_t *symbols;
  int8_t *planes;
  int32_t *indices;
  int32_t *group_sums;   // one per index-stream group
  int32_t *lane_sums;    // index-stream partial sums (layout at LaneSuffixBase), or null
  uint8_t *out;
};

// Partial sums of (diff - 128) over the index stream that the rANS stage leaves in
// lane_sums, so a palette index can start at any 32-block boundary: int32 entries
// [0, nv/256) hold, per lane (256 symbols), the inclusive prefix of lane totals over
// its group's lanes; entries [nv/256, nv/256 + nv/32) hold, per lane, R[k] for
// k = 0..7, the sum over the lane's positions >= 32k.
constexpr uint32_t kRowBlocks = 32;
__host__ __device__ inline uint32_t LaneSuffixBase(uint32_t nv) { return nv / kSymsPerLane; }

template <typename Desc>
__host__ __device__ inline uint32_t SymbolOffset(const Desc &d, uint32_t k) {
  return k == 0 ? 0u : k == 1 ? 2 * d.nv : k == 2 ? 6 * d.nv : 6 * d.nv + d.palette_bytes;
}

// rANS work units.  A unit is up to `per` consecutive groups of one stream (per = waves
// x groups per wave; the low 8 bits of the unit code), decoded by one work-group.  With
// kAnsPairShort in the code, a stream's remainder of at most per / 2 groups gets no
// work-group of its own: an image's short remainders go two to a work-group, one per
// half of its waves, each half with its own table.  Fewer, fuller work-groups then fill
// the chip in one pass.  An image's units are every stream's full units and long
// remainder, in stream order, then the pairs of short remainders.
constexpr uint32_t kAnsPairShort = 0x100;

__host__ __device__ inline bool AnsShortRemainder(uint32_t groups, uint32_t code) {
  const uint32_t per = code & 0xFFu, rem = groups % per;
  return (code & kAnsPairShort) != 0 && rem != 0 && rem <= per / 2;
}

// Units of stream k that are not short remainders.
__host__ __device__ inline uint32_t AnsStreamUnits(uint32_t groups, uint32_t code) {
  const uint32_t per = code & 0xFFu;
  return groups / per + (groups % per != 0 && !AnsShortRemainder(groups, code) ? 1u : 0u);
}

template <typename Desc>
__host__ __device__ inline uint32_t AnsUnitsOf(const Desc &d, uint32_t code) {
  uint32_t u = 0, shorts = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    u += AnsStreamUnits(d.groups[k], code);
    shorts += AnsShortRemainder(d.groups[k], code) ? 1u : 0u;
  }
  return u + (shorts + 1) / 2;
}

// Unit `local` of an image: one segment (k[0], first[0], count[0]) or, for a pair of
// short remainders, two (count[1] = 0 when the image's last pair has one).
struct AnsUnit {
  uint32_t k[2], first[2], count[2];
  bool pair;
};

template <typename Desc>
__host__ __device__ inline AnsUnit AnsUnitAt(const Desc &d, uint32_t code, uint32_t local) {
  const uint32_t per = code & 0xFFu;
  AnsUnit r = {{0u, 0u}, {0u, 0u}, {0u, 0u}, false};
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t g = d.groups[k], nk = AnsStreamUnits(g, code);
    if (local < nk) {
      r.k[0] = k;
      r.first[0] = local * per;
      r.count[0] = g - local * per < per ? g - local * per : per;
      return r;
    }
    local -= nk;
  }
  r.pair = true;
  uint32_t idx = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t g = d.groups[k];
    if (!AnsShortRemainder(g, code)) continue;
    // Constant subscripts only: a computed one would put the unit in scratch memory.
    if (idx == 2 * local) {
      r.k[0] = k;
      r.first[0] = g - g % per;
      r.count[0] = g % per;
    } else if (idx == 2 * local + 1) {
      r.k[1] = k;
      r.first[1] = g - g % per;
      r.count[1] = g % per;
    }
    ++idx;
  }
  return r;
}

// `ans_unit_code`: groups per rANS unit, with kAnsPairShort when short remainders pair up.
__host__ __device__ inline uint32_t StageUnitsOf(const ImageDesc &d, uint32_t stage,
                                                 uint32_t ans_unit_code) {
  switch (stage) {
    case kStageTables: return 4;
    case kStageAns: return AnsUnitsOf(d, ans_unit_code);
    case kStageIndices: return d.nv / kGroupSyms;
    case kStageWavelet: return 6 * (d.nv / kTileSyms) / kWaveletTilesPerUnit;
    default: return d.nv / kAssembleBlocksPerUnit;
  }
}

// Device plan for batched decodes of up to `cap` images, n of them bound: the header
// (padded to 64 bytes), then prefix[kNumStages][cap + 1] (work units per stage,
// exclusive prefix over the images; entries past index n hold 0xFFFFFFFF), then
// ImageDesc[cap], of which the first n are used.  Every offset depends on cap alone,
// which the kernels receive as an argument, so a kernel reads the header and the
// prefix in one round trip; the used bytes (everything up to the n-th descriptor)
// are one contiguous copy.
struct PlanHeader {
  uint32_t n_images;
  uint32_t units[kNumStages];
  uint32_t ans_unit_code;  // groups per rANS unit | kAnsPairShort
};
constexpr size_t kPlanHeaderBytes = 64;
static_assert(sizeof(PlanHeader) <= kPlanHeaderBytes, "plan header");
constexpr uint32_t kPrefixPad = 0xFFFFFFFFu;

__host__ __device__ inline size_t PlanPrefixOffset(uint32_t cap, uint32_t stage) {
  return kPlanHeaderBytes + sizeof(uint32_t) * static_cast<size_t>(stage) * (cap + 1);
}
__host__ __device__ inline size_t PlanImagesOffset(uint32_t cap) {
  return kPlanHeaderBytes + ((sizeof(uint32_t) * kNumStages * (cap + 1) + 63) & ~static_cast<size_t>(63));
}
__host__ __device__ inline size_t PlanUsedBytes(uint32_t cap, uint32_t n) {
  return PlanImagesOffset(cap) + sizeof(ImageDesc) * n;
}
__host__ __device__ inline size_t PlanBytesFor(uint32_t cap) { return PlanUsedBytes(cap, cap); }

// Kernel arguments (by value).  plan == nullptr selects the single image in `single`.
struct KernelArgs {
  const PlanHeader *plan;
  uint32_t plan_cap;       // the plan's image capacity (fixes its layout)
  uint32_t ans_unit_code;  // without a plan: the rANS unit code AnsAutoKernel decodes with
  uint32_t *err;
  uint32_t *tables;        // n_tables x 2048 packed entries
  uint32_t *table_status;  // n_tables: 1 = valid
  ImageDesc single;
};

// Launch shape of the rANS stage and the assembly flavour.
struct StageConfig {
  uint32_t ans_waves;            // waves per rANS work-group (1, 2, 4, 8)
  uint32_t ans_groups_per_wave;  // 1 or 2
  uint32_t ans_build_tables;     // 1: the rANS kernel builds its table in LDS (4 or 8 waves,
                                 // or 2 waves of two groups); the table stage launches nothing
  uint32_t ans_pair_short;       // 1: short stream remainders pair up in one work-group
                                 // (4 waves x 2 groups, tables built in LDS)
  uint32_t ans_auto;             // 1: AnsAutoKernel (4 waves x 2 groups, tables built in LDS):
                                 // ans_pair_short 1 runs the paired form, 0 the short-chain
                                 // form (8-byte table and ring entries, one LDS round trip a
                                 // step); the kernel takes the form from the unit code
  uint32_t wavelet_waves;        // waves per wavelet work-group (1, 2, 4); a wave takes 2 units
  uint32_t fused_waves;          // 0: separate index scan and assembly; else waves per
                                 // work-group of the fused scan + DXT1 assembly (1, 2, 4)
  uint32_t fused_wide;           // fused kernel layout: 0 = 4 blocks per lane x 4 segments,
                                 // 1 = 16 consecutive blocks per lane
  uint32_t tile_fused;           // 1: the wavelet and the index scan run inside the DXT1
                                 // assembly (two tile positions per 3-wave work-group);
                                 // the wavelet and index stages launch nothing
  uint32_t rgb;                  // 1: RGB8 assembly
};

// Work-groups that give every unit of `units` its own pass (no grid-stride repeat).
inline uint32_t SinglePassGrid(uint32_t stage, uint32_t units, const StageConfig &c) {
  if (c.tile_fused && (stage == kStageWavelet || stage == kStageIndices)) return 0;
  if (c.tile_fused && stage == kStageAssemble) return (units + 1) / 2;  // two units a group
  if (stage == kStageWavelet) {
    const uint32_t per = 2 * c.wavelet_waves;
    return (units + per - 1) / per;
  }
  if (c.ans_build_tables && stage == kStageTables) return 0;  // built by the rANS kernel
  if (c.fused_waves && stage == kStageIndices) return 0;  // folded into the assembly
  if (c.fused_waves && stage == kStageAssemble) return (units + c.fused_waves - 1) / c.fused_waves;
  return units;
}

// Whether LaunchStage puts a kernel on the stream for this stage: a grid of 0, tables
// built inside the rANS kernel, and an index scan or wavelet folded into the assembly
// launch nothing.
inline bool StageLaunches(uint32_t stage, uint32_t grid, const StageConfig &c) {
  if (grid == 0) return false;
  if (stage == kStageTables) return !c.ans_build_tables;
  if (stage == kStageIndices) return !c.fused_waves && !c.tile_fused;
  if (stage == kStageWavelet) return !c.tile_fused;
  return true;
}

// Launches one stage with `grid` work-groups (each loops over units grid-stride).
hipError_t LaunchStage(uint32_t stage, const KernelArgs &args, uint32_t grid,
                       const StageConfig &cfg, hipStream_t stream);

// Stand-alone rANS groups (`groups` end offsets at the head of `data`, n = 1..32
// interleaved streams each) against one packed table; out gets groups * n * 256 bytes.
hipError_t LaunchAnsGroups(const uint32_t *table, const uint8_t *data, uint32_t data_len, uint32_t groups,
                           uint32_t n, uint8_t *out, uint32_t *err, hipStream_t stream);

// Device-side comparison of decoded outputs with expected bytes: the server checks
// every completion this way under load.  Job i sets flags[job.flag] to 1 when its
// `bytes` (a multiple of 16, both pointers 16-byte aligned) differ anywhere.
struct CompareJob {
  const uint8_t *out;
  const uint8_t *expected;
  uint64_t bytes;
  uint64_t flag;
};
constexpr uint32_t kCompareJobs = 64;
struct CompareArgs {
  CompareJob job[kCompareJobs];
  uint32_t n;
  uint32_t *flags;
};
hipError_t LaunchCompare(const CompareArgs &a, hipStream_t stream);

// Host-to-device gather in one kernel: job i copies `bytes` from `src` (pinned host
// memory mapped for the device, 16-byte aligned) to `dst` (device memory, 16-byte
// aligned).  One launch replaces a DMA copy per job, whose fixed cost dominates copies
// of a few hundred KiB.
struct GatherJob {
  const uint8_t *src;
  uint8_t *dst;
  uint64_t bytes;
};
constexpr uint32_t kGatherJobs = 64;
struct GatherArgs {
  GatherJob job[kGatherJobs];
  uint32_t n;
};
hipError_t LaunchGather(const GatherArgs &a, hipStream_t stream);

// Static device scratch of gst_load_compressed_dxt_noalloc: one array per GPU, held
// until the process ends.  ResolveNoAllocScratch looks up the array's address on the
// current device (once per device; the caller makes `device` current);
// NoAllocScratch returns it, or null before it was resolved.
constexpr size_t kNoAllocScratchBytes = size_t{12} << 20;
hipError_t ResolveNoAllocScratch(int device);
uint8_t *NoAllocScratch(int device);

// The decoder library's one-time device set-up (gst_initialize_decoder(0) on first use).
int EnsureDecoderInit();

}  // namespace gst
