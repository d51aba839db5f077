// gst_decoder.cpp -- host orchestration and the C ABI of libgst_amd.so.
//
// Mirrors the reference decoder's host side (codec/decoder.cpp): header checks,
// offset computation, scratch carving, and the five-stage launch sequence
// build_table -> ans_decode -> decode_indices -> inv_wavelet -> assemble on one
// stream (stream order replaces the reference's OpenCL event DAG, :98-428).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gst_decoder.h"
#include "gst_internal.h"

using gst::ImageDesc;
using gst::KernelArgs;
using gst::PlanHeader;
using gst::StageConfig;

namespace {

constexpr size_t kAlign = 256;
size_t AlignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Globals {
  std::mutex mu;
  bool initialized = false;
  int device = 0;
  int cu_count = 0;
  // Preallocated scratch (PreallocateDecompressor): bump allocation in call order.
  uint8_t *pre_base = nullptr;
  size_t pre_size = 0;
  size_t pre_off = 0;
  // Accumulated device error bits per stream.
  std::map<hipStream_t, uint32_t *> err_words;
};

Globals &G() {
  static Globals g;
  return g;
}

int HipFail(hipError_t e) {
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? GST_ERR_OUT_OF_MEMORY : GST_ERR_HIP;
}

#define GST_HIP(call)                         \
  do {                                        \
    hipError_t _e = (call);                   \
    if (_e != hipSuccess) return HipFail(_e); \
  } while (0)

int EnsureInit() {
  Globals &g = G();
  if (g.initialized) return GST_OK;
  return gst_initialize_decoder(0) == GST_OK ? GST_OK : GST_ERR_NOT_INITIALIZED;
}

// Header rules: README.md:67-71 (w, h multiples of 128; w*h multiple of 16*256*32),
// codec/encoder.cpp:100-108 (palette padded to 8192 B), codec/entropy.cpp:202-262
// (stream = u32 end offsets + groups of >= 128 bytes, multiple of 4 bytes).
int CheckHeader(const gst_header &h) {
  const uint64_t w = h.width, hh = h.height;
  if (w == 0 || hh == 0 || (w % 128) || (hh % 128) || ((w * hh) % 131072)) return GST_ERR_BAD_HEADER;
  const uint64_t nv = w * hh / 16;
  if (nv > (1ull << 24)) return GST_ERR_BAD_HEADER;
  const uint64_t P = h.palette_bytes;
  if (P == 0 || (P % gst::kGroupSyms) || P > 4 * nv) return GST_ERR_BAD_HEADER;
  const uint64_t syms[4] = {2 * nv, 4 * nv, P, nv};
  const uint64_t sz[4] = {h.y_cmp_sz, h.chroma_cmp_sz, h.palette_sz, h.indices_sz};
  uint64_t total = 0;
  for (int k = 0; k < 4; ++k) {
    if (sz[k] % 4) return GST_ERR_BAD_HEADER;
    if (sz[k] < syms[k] / gst::kGroupSyms * (4 + 4 * gst::kLanesPerGroup)) return GST_ERR_BAD_HEADER;
    total += sz[k];
  }
  if (total >= (1ull << 31)) return GST_ERR_BAD_HEADER;
  return GST_OK;
}

uint64_t StreamBytes(const gst_header &h) {
  return static_cast<uint64_t>(h.y_cmp_sz) + h.chroma_cmp_sz + h.palette_sz + h.indices_sz;
}

uint32_t NumBlocks(const gst_header &h) { return h.width / 4 * (h.height / 4); }

// Per-image scratch carve-up (the call's tables and status words sit before it).
struct Carve {
  size_t symbols, planes, indices, group_sums, total;
};

Carve CarveFor(const gst_header &h) {
  const size_t nv = NumBlocks(h), P = h.palette_bytes;
  Carve c;
  size_t off = 0;
  c.symbols = off; off += AlignUp(7 * nv + P, kAlign);
  c.planes = off; off += AlignUp(6 * nv, kAlign);
  c.indices = off; off += AlignUp(4 * nv, kAlign);
  c.group_sums = off; off += AlignUp(4 * (nv / gst::kGroupSyms), kAlign);
  c.total = off;
  return c;
}

void FillDesc(const gst_header &h, const uint8_t *freqs, const uint8_t *streams, uint8_t *scratch,
              const Carve &c, uint32_t table_base, uint8_t *out, bool rgb, ImageDesc *d) {
  const uint32_t nv = NumBlocks(h);
  memset(d, 0, sizeof(*d));
  d->width = h.width;
  d->height = h.height;
  d->nv = nv;
  d->bx = h.width / 4;
  d->palette_bytes = h.palette_bytes;
  const uint32_t syms[4] = {2 * nv, 4 * nv, h.palette_bytes, nv};
  const uint32_t sz[4] = {h.y_cmp_sz, h.chroma_cmp_sz, h.palette_sz, h.indices_sz};
  size_t in_off = 0;
  for (int k = 0; k < 4; ++k) {
    d->groups[k] = syms[k] / gst::kGroupSyms;
    d->stream_len[k] = sz[k];
    d->streams[k] = streams + in_off;
    in_off += sz[k];
  }
  d->table_base = table_base;
  d->rgb = rgb ? 1 : 0;
  d->freqs = freqs;
  d->symbols = scratch + c.symbols;
  d->planes = reinterpret_cast<int8_t *>(scratch + c.planes);
  d->indices = reinterpret_cast<int32_t *>(scratch + c.indices);
  d->group_sums = reinterpret_cast<int32_t *>(scratch + c.group_sums);
  // The rANS stage leaves the index stream's partial sums at the head of the index
  // region; the separate index scan overwrites them afterwards, the fused assemblies
  // read them.
  d->lane_sums = d->indices;
  d->out = out;
}

StageConfig DefaultStageConfig(bool rgb) {
  StageConfig s{};
  s.ans_waves = 4;
  s.ans_groups_per_wave = 2;
  s.wavelet_waves = 1;
  s.fused_waves = rgb ? 0 : 4;
  s.tile_fused = rgb ? 0 : 1;
  s.ans_build_tables = 1;
  s.ans_auto = 1;  // GST_ANS_FORM_AUTO: the caller sets ans_pair_short from AutoPairs
  s.ans_pair_short = 0;
  s.rgb = rgb ? 1 : 0;
  return s;
}

// The rANS forms of the 4-wave x 2-group shape with tables built in LDS
// (GST_ANS_FORM_*): PLAIN is AnsDecodeKernel<4, 2, true> (units of 8 groups of one
// stream, 7 work-groups per CU); PAIRED and AUTO run AnsAutoKernel, which reads the
// unit code of its plan: AUTO pairs short remainders only when the 8-group units would
// not fit the device in one pass at 6 work-groups per CU.
void SetAnsForm(StageConfig *s, uint32_t form, bool pair = true) {
  s->ans_auto = (form == GST_ANS_FORM_PAIRED || form == GST_ANS_FORM_AUTO) ? 1 : 0;
  s->ans_pair_short = (form == GST_ANS_FORM_PAIRED || (form == GST_ANS_FORM_AUTO && pair)) ? 1 : 0;
}

uint32_t AnsFormOf(const StageConfig &s) {
  if (!s.ans_build_tables || s.ans_waves != 4 || s.ans_groups_per_wave != 2) return 0;
  if (!s.ans_auto) return GST_ANS_FORM_PLAIN;
  return s.ans_pair_short ? GST_ANS_FORM_PAIRED : GST_ANS_FORM_AUTO;
}

// Whether GST_ANS_FORM_AUTO pairs short remainders for these images: only when their
// units of 8 groups would take more than one pass of 6 work-groups per CU (with the
// 7th work-group of a CU a SIMD holds 7 decoding waves, and one pass is all that keeps
// the per-stream chains overlapped).
bool AutoPairs(const ImageDesc *descs, uint32_t n) {
  uint64_t units = 0;
  for (uint32_t i = 0; i < n; ++i) units += gst::AnsUnitsOf(descs[i], 8);
  const uint64_t cus = static_cast<uint64_t>(G().cu_count > 0 ? G().cu_count : 256);
  return units > 6 * cus;
}

// Groups per rANS unit, with gst::kAnsPairShort when short remainders pair up.
uint32_t AnsUnitCode(const StageConfig &s) {
  return s.ans_waves * s.ans_groups_per_wave | (s.ans_pair_short ? gst::kAnsPairShort : 0u);
}

uint32_t *ErrWord(hipStream_t s) {
  Globals &g = G();
  std::lock_guard<std::mutex> lk(g.mu);
  auto it = g.err_words.find(s);
  if (it != g.err_words.end()) return it->second;
  uint32_t *w = nullptr;
  if (hipMalloc(&w, sizeof(uint32_t)) != hipSuccess) return nullptr;
  if (hipMemset(w, 0, sizeof(uint32_t)) != hipSuccess) {
    (void)hipFree(w);
    return nullptr;
  }
  g.err_words[s] = w;
  return w;
}

// Pinned host staging for small descriptor uploads: slots recycled once the copy
// that last used them has completed.
class PinnedRing {
 public:
  static PinnedRing &Get() {
    static PinnedRing r;
    return r;
  }
  int Upload(void *d_dst, const void *src, size_t n, hipStream_t s) {
    std::lock_guard<std::mutex> lk(mu_);
    if (slots_.empty()) {
      for (int i = 0; i < kSlots; ++i) {
        Slot sl;
        // Coherent: the device reads a slot itself and must never see a stale cached copy.
        GST_HIP(hipHostMalloc(&sl.host, kSlotBytes, hipHostMallocCoherent));
        GST_HIP(hipHostGetDevicePointer(&sl.dev, sl.host, 0));
        GST_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
        sl.used = false;
        slots_.push_back(sl);
      }
    }
    if (n > kSlotBytes) {
      void *h = nullptr;
      GST_HIP(hipHostMalloc(&h, n, hipHostMallocDefault));
      memcpy(h, src, n);
      hipError_t e = hipMemcpyAsync(d_dst, h, n, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      (void)hipHostFree(h);
      GST_HIP(e);
      return GST_OK;
    }
    Slot &sl = slots_[next_];
    next_ = (next_ + 1) % kSlots;
    if (sl.used) GST_HIP(hipEventSynchronize(sl.ev));
    memcpy(sl.host, src, n);
    // One gather launch instead of a DMA copy, whose fixed cost (about 10 us) would sit
    // between the call and its first kernel.
    gst::GatherArgs ga;
    ga.n = 1;
    ga.job[0] = {static_cast<const uint8_t *>(sl.dev), static_cast<uint8_t *>(d_dst), n};
    GST_HIP(gst::LaunchGather(ga, s));
    GST_HIP(hipEventRecord(sl.ev, s));
    sl.used = true;
    return GST_OK;
  }

 private:
  static constexpr int kSlots = 16;
  static constexpr size_t kSlotBytes = 64 << 10;
  struct Slot {
    void *host;
    void *dev;  // the slot's address for device code
    hipEvent_t ev;
    bool used;
  };
  std::mutex mu_;
  std::vector<Slot> slots_;
  int next_ = 0;
};

// The used part of a plan of n <= cap images (layout in gst_internal.h): its size is
// PlanUsedBytes(cap, n).
void BuildPlan(const std::vector<ImageDesc> &descs, uint32_t cap, uint32_t ans_unit_code,
               std::vector<uint8_t> *blob, uint32_t units_out[gst::kNumStages]) {
  const uint32_t n = static_cast<uint32_t>(descs.size());
  blob->assign(gst::PlanUsedBytes(cap, n), 0);
  PlanHeader *ph = reinterpret_cast<PlanHeader *>(blob->data());
  ImageDesc *im = reinterpret_cast<ImageDesc *>(blob->data() + gst::PlanImagesOffset(cap));
  ph->n_images = n;
  ph->ans_unit_code = ans_unit_code;
  for (uint32_t st = 0; st < gst::kNumStages; ++st) {
    uint32_t *pre = reinterpret_cast<uint32_t *>(blob->data() + gst::PlanPrefixOffset(cap, st));
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
      pre[i] = acc;
      acc += gst::StageUnitsOf(descs[i], st, ans_unit_code);
    }
    pre[n] = acc;
    for (uint32_t i = n + 1; i <= cap; ++i) pre[i] = gst::kPrefixPad;
    ph->units[st] = acc;
    units_out[st] = acc;
  }
  memcpy(im, descs.data(), sizeof(ImageDesc) * n);
}

KernelArgs PlanArgs(uint8_t *d_blob, uint32_t cap, uint32_t *err, uint32_t *tables, uint32_t *status) {
  KernelArgs a;
  memset(&a, 0, sizeof(a));
  a.plan = reinterpret_cast<const PlanHeader *>(d_blob);
  a.plan_cap = cap;
  a.err = err;
  a.tables = tables;
  a.table_status = status;
  return a;
}

int LaunchAll(const KernelArgs &a, const uint32_t grids[gst::kNumStages], const StageConfig &cfg,
              hipStream_t s) {
  for (uint32_t st = 0; st < gst::kNumStages; ++st)
    GST_HIP(gst::LaunchStage(st, a, grids[st], cfg, s));
  return GST_OK;
}

// Scratch for one decode call: the preallocation when it has room, otherwise
// stream-ordered device memory released on the stream after the last kernel.
struct Scratch {
  uint8_t *ptr = nullptr;
  bool async = false;
};

int TakeScratch(size_t bytes, hipStream_t s, Scratch *out) {
  Globals &g = G();
  {
    std::lock_guard<std::mutex> lk(g.mu);
    if (g.pre_base) {
      const size_t off = AlignUp(g.pre_off, kAlign);
      if (off + bytes <= g.pre_size) {
        out->ptr = g.pre_base + off;
        out->async = false;
        g.pre_off = off + bytes;
        return GST_OK;
      }
    }
  }
  void *p = nullptr;
  GST_HIP(hipMallocAsync(&p, bytes, s));
  out->ptr = static_cast<uint8_t *>(p);
  out->async = true;
  return GST_OK;
}

// Scratch of one decode call: [tables 4n x 8 KiB][status 4n u32][plan blob if n > 1]
// [per-image carves].
struct CallLayout {
  size_t tables, status, plan, images, total;
  std::vector<size_t> carve_off;
  std::vector<Carve> carves;
};

CallLayout LayoutCall(const gst_header *hdrs, uint32_t n) {
  CallLayout L;
  size_t off = 0;
  L.tables = off; off += AlignUp(4ull * n * gst::kTableSize * 4, kAlign);
  L.status = off; off += AlignUp(4ull * n * 4, kAlign);
  L.plan = off; off += n > 1 ? AlignUp(gst::PlanBytesFor(n), kAlign) : 0;
  L.images = off;
  L.carves.resize(n);
  L.carve_off.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    L.carves[i] = CarveFor(hdrs[i]);
    L.carve_off[i] = off;
    off += L.carves[i].total;
  }
  L.total = off;
  return L;
}

// n images from a staging buffer in the reference layout (codec/decoder.cpp:98-428).
int DecodeStaging(const gst_header *hdrs, uint32_t n, hipStream_t s, const void *d_cmp, void *d_out,
                  bool rgb, uint32_t num_init, const gst_event *init, gst_event *done) {
  if (!hdrs || n == 0 || !d_cmp || !d_out || (num_init && !init)) return GST_ERR_INVALID_ARGUMENT;
  int rc = EnsureInit();
  if (rc != GST_OK) return rc;
  for (uint32_t i = 0; i < n; ++i)
    if ((rc = CheckHeader(hdrs[i])) != GST_OK) return rc;
  uint32_t *err = ErrWord(s);
  if (!err) return GST_ERR_HIP;
  for (uint32_t i = 0; i < num_init; ++i)
    GST_HIP(hipStreamWaitEvent(s, static_cast<hipEvent_t>(init[i]), 0));

  // codec/decoder.cpp:109-111,172-209: offsets block, n x 2048 B frequencies, streams.
  const size_t offsets_bytes = AlignUp(32ull * n, 512);
  const uint8_t *base = static_cast<const uint8_t *>(d_cmp);
  const uint8_t *streams = base + offsets_bytes + 2048ull * n;
  const CallLayout L = LayoutCall(hdrs, n);
  Scratch sc;
  if ((rc = TakeScratch(L.total, s, &sc)) != GST_OK) return rc;

  StageConfig cfg = DefaultStageConfig(rgb);
  std::vector<ImageDesc> descs(n);
  size_t in_off = 0, out_off = 0;
  for (uint32_t i = 0; i < n; ++i) {
    FillDesc(hdrs[i], base + offsets_bytes + 2048ull * i, streams + in_off, sc.ptr + L.carve_off[i],
             L.carves[i], 4 * i, static_cast<uint8_t *>(d_out) + out_off, rgb, &descs[i]);
    in_off += StreamBytes(hdrs[i]);
    out_off += rgb ? 3ull * hdrs[i].width * hdrs[i].height : hdrs[i].width * hdrs[i].height / 2ull;
  }
  SetAnsForm(&cfg, GST_ANS_FORM_AUTO, AutoPairs(descs.data(), n));
  uint32_t *tables = reinterpret_cast<uint32_t *>(sc.ptr + L.tables);
  uint32_t *status = reinterpret_cast<uint32_t *>(sc.ptr + L.status);
  uint32_t grids[gst::kNumStages];
  if (n == 1) {
    KernelArgs a;
    memset(&a, 0, sizeof(a));
    a.single = descs[0];
    a.ans_unit_code = AnsUnitCode(cfg);
    a.err = err;
    a.tables = tables;
    a.table_status = status;
    for (uint32_t st = 0; st < gst::kNumStages; ++st)
      grids[st] = gst::SinglePassGrid(st, gst::StageUnitsOf(descs[0], st, AnsUnitCode(cfg)), cfg);
    rc = LaunchAll(a, grids, cfg, s);
  } else {
    std::vector<uint8_t> blob;
    uint32_t units[gst::kNumStages];
    BuildPlan(descs, n, AnsUnitCode(cfg), &blob, units);
    uint8_t *d_blob = sc.ptr + L.plan;
    rc = PinnedRing::Get().Upload(d_blob, blob.data(), blob.size(), s);
    if (rc == GST_OK) {
      const KernelArgs a = PlanArgs(d_blob, n, err, tables, status);
      for (uint32_t st = 0; st < gst::kNumStages; ++st) grids[st] = gst::SinglePassGrid(st, units[st], cfg);
      rc = LaunchAll(a, grids, cfg, s);
    }
  }
  if (sc.async) {
    const hipError_t e = hipFreeAsync(sc.ptr, s);
    if (rc == GST_OK && e != hipSuccess) rc = HipFail(e);
  }
  if (rc != GST_OK) return rc;
  if (done) {
    hipEvent_t ev;
    GST_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    GST_HIP(hipEventRecord(ev, s));
    *done = ev;
  }
  return GST_OK;
}

}  // namespace

int gst::EnsureDecoderInit() { return EnsureInit(); }

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int gst_initialize_decoder(int device) {
  Globals &g = G();
  std::lock_guard<std::mutex> lk(g.mu);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    return GST_ERR_NOT_INITIALIZED;
  }
  GST_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  GST_HIP(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) return GST_ERR_NOT_INITIALIZED;
  // The kernels need 256-thread work-groups and 16 KiB + 8.4 KiB of LDS
  // (codec/decoder.cpp:558-591 checks the work-group sizes of the reference's).
  if (prop.maxThreadsPerBlock < 512 || prop.sharedMemPerBlock < 32768) return GST_ERR_NOT_INITIALIZED;
  // The no-allocation entry point's static scratch on this device (current here).
  GST_HIP(gst::ResolveNoAllocScratch(device));
  g.device = device;
  g.cu_count = prop.multiProcessorCount;
  g.initialized = true;
  return GST_OK;
}

int gst_parse_header(const uint8_t *file, size_t len, gst_header *hdr) {
  if (!file || !hdr) return GST_ERR_INVALID_ARGUMENT;
  if (len < sizeof(gst_header) + 2048) return GST_ERR_BAD_HEADER;
  memcpy(hdr, file, sizeof(gst_header));
  const int rc = CheckHeader(*hdr);
  if (rc != GST_OK) return rc;
  if (sizeof(gst_header) + 2048 + StreamBytes(*hdr) > len) return GST_ERR_BAD_HEADER;
  return GST_OK;
}

size_t gst_required_scratch_mem(const gst_header *hdr) {
  if (!hdr) return 0;
  // codec/decoder.cpp:41-47
  return 4 * 2048 * 6 + 17ull * hdr->width * hdr->height / 16 + hdr->palette_bytes;
}

int gst_preallocate_decompressor(size_t bytes) {
  int rc = EnsureInit();
  if (rc != GST_OK) return rc;
  Globals &g = G();
  std::lock_guard<std::mutex> lk(g.mu);
  if (g.pre_base) {
    GST_HIP(hipDeviceSynchronize());
    GST_HIP(hipFree(g.pre_base));
    g.pre_base = nullptr;
  }
  void *p = nullptr;
  GST_HIP(hipMalloc(&p, bytes));
  g.pre_base = static_cast<uint8_t *>(p);
  g.pre_size = bytes;
  g.pre_off = 0;
  return GST_OK;
}

int gst_free_decompressor(void) {
  Globals &g = G();
  std::lock_guard<std::mutex> lk(g.mu);
  if (g.pre_base) {
    GST_HIP(hipDeviceSynchronize());
    GST_HIP(hipFree(g.pre_base));
  }
  g.pre_base = nullptr;
  g.pre_size = g.pre_off = 0;
  return GST_OK;
}

int gst_read_error_bits(gst_stream stream, uint32_t *bits) {
  if (!bits) return GST_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t *w = ErrWord(s);
  if (!w) return GST_ERR_HIP;
  GST_HIP(hipStreamSynchronize(s));
  GST_HIP(hipMemcpy(bits, w, sizeof(uint32_t), hipMemcpyDeviceToHost));
  GST_HIP(hipMemset(w, 0, sizeof(uint32_t)));
  return GST_OK;
}

int gst_read_error_bits_async(gst_stream stream, uint32_t *host_bits) {
  if (!host_bits) return GST_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t *w = ErrWord(s);
  if (!w) return GST_ERR_HIP;
  GST_HIP(hipMemcpyAsync(host_bits, w, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  GST_HIP(hipMemsetAsync(w, 0, sizeof(uint32_t), s));
  return GST_OK;
}

int gst_staging_size(const gst_header *hdrs, uint32_t n, size_t *total, size_t *streams_offset) {
  if (!hdrs || n == 0 || !total) return GST_ERR_INVALID_ARGUMENT;
  size_t bytes = AlignUp(32ull * n, 512) + 2048ull * n;
  if (streams_offset) *streams_offset = bytes;
  for (uint32_t i = 0; i < n; ++i) bytes += StreamBytes(hdrs[i]);
  *total = bytes;
  return GST_OK;
}

int gst_upload_gather(const void *const *host_src, void *const *dev_dst, const uint64_t *bytes, uint32_t n,
                      gst_stream stream) {
  if (n && (!host_src || !dev_dst || !bytes)) return GST_ERR_INVALID_ARGUMENT;
  if (n == 0) return GST_OK;
  const int rc = EnsureInit();
  if (rc != GST_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  gst::GatherArgs ga;
  ga.n = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (bytes[i] == 0) continue;
    if (!host_src[i] || !dev_dst[i] ||
        ((reinterpret_cast<uintptr_t>(host_src[i]) | reinterpret_cast<uintptr_t>(dev_dst[i])) & 15))
      return GST_ERR_INVALID_ARGUMENT;
    // Only host memory the runtime pinned has a device address; a kernel must never read
    // anything else (pageable memory would fault the device).
    hipPointerAttribute_t attr{};
    void *src = nullptr;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipPointerGetAttributes(&attr, host_src[i]) != hipSuccess || attr.type != hipMemoryTypeHost ||
        hipHostGetDevicePointer(&src, const_cast<void *>(host_src[i]), 0) != hipSuccess || !src ||
        hipMemGetAddressRange(&base, &size, src) != hipSuccess ||
        static_cast<uint8_t *>(src) + bytes[i] > static_cast<uint8_t *>(base) + size) {
      (void)hipGetLastError();
      return GST_ERR_INVALID_ARGUMENT;  // not pinned, or the range runs past its allocation
    }
    ga.job[ga.n++] = {static_cast<const uint8_t *>(src), static_cast<uint8_t *>(dev_dst[i]), bytes[i]};
    if (ga.n == gst::kGatherJobs) {
      GST_HIP(gst::LaunchGather(ga, s));
      ga.n = 0;
    }
  }
  if (ga.n) GST_HIP(gst::LaunchGather(ga, s));
  return GST_OK;
}

int gst_pack_staging(const uint8_t *const *files, const size_t *lens, uint32_t n, uint8_t *out,
                     size_t cap) {
  if (!files || !lens || n == 0 || !out) return GST_ERR_INVALID_ARGUMENT;
  std::vector<gst_header> hdrs(n);
  for (uint32_t i = 0; i < n; ++i) {
    const int rc = gst_parse_header(files[i], lens[i], &hdrs[i]);
    if (rc != GST_OK) return rc;
  }
  size_t total = 0, soff = 0;
  gst_staging_size(hdrs.data(), n, &total, &soff);
  if (total > cap) return GST_ERR_INVALID_ARGUMENT;
  const size_t offsets_bytes = AlignUp(32ull * n, 512);
  memset(out, 0, offsets_bytes);
  uint32_t *out_off = reinterpret_cast<uint32_t *>(out);
  uint32_t *in_off = out_off + 4 * n;
  uint32_t oo = 0, io = 0;
  uint8_t *data = out + soff;
  for (uint32_t i = 0; i < n; ++i) {
    const gst_header &h = hdrs[i];
    const uint32_t nv = NumBlocks(h);
    // demo/photos_sf.cpp:778-789
    out_off[4 * i + 0] = oo; oo += 2 * nv;
    out_off[4 * i + 1] = oo; oo += 4 * nv;
    out_off[4 * i + 2] = oo; oo += h.palette_bytes;
    out_off[4 * i + 3] = oo; oo += nv;
    in_off[4 * i + 0] = io; io += h.y_cmp_sz;
    in_off[4 * i + 1] = io; io += h.chroma_cmp_sz;
    in_off[4 * i + 2] = io; io += h.palette_sz;
    in_off[4 * i + 3] = io; io += h.indices_sz;
    memcpy(out + offsets_bytes + 2048ull * i, files[i] + sizeof(gst_header), 2048);
    const size_t sb = StreamBytes(h);
    memcpy(data, files[i] + sizeof(gst_header) + 2048, sb);
    data += sb;
  }
  return GST_OK;
}

int gst_load_compressed_dxt(const gst_header *hdr, gst_stream stream, const void *d_cmp, void *d_out,
                            uint32_t num_init, const gst_event *init, gst_event *done) {
  return DecodeStaging(hdr, 1, static_cast<hipStream_t>(stream), d_cmp, d_out, false, num_init, init,
                       done);
}

int gst_load_compressed_dxt_noalloc(const gst_header *hdr, gst_stream stream, const void *d_cmp, void *d_out,
                                    uint32_t *d_err_bits) {
  if (!hdr || !d_cmp || !d_out || !d_err_bits) return GST_ERR_INVALID_ARGUMENT;
  int rc = EnsureInit();
  if (rc != GST_OK) return rc;
  if ((rc = CheckHeader(*hdr)) != GST_OK) return rc;
  int device = 0;
  GST_HIP(hipGetDevice(&device));
  uint8_t *base = gst::NoAllocScratch(device);
  if (!base) {
    // First call on a device the library was not initialized on.
    GST_HIP(gst::ResolveNoAllocScratch(device));
    base = gst::NoAllocScratch(device);
  }
  // The decode's two kernels (rANS with tables built in LDS, then wavelet + index scan +
  // DXT1 assembly) read and write only the symbols, the index stream's partial sums and
  // its group totals: [symbols 7nv+P][partial sums nv/64 + nv/8][group totals nv/2048].
  const size_t nv = NumBlocks(*hdr), P = hdr->palette_bytes;
  Carve c;
  size_t off = 0;
  c.symbols = off; off += AlignUp(7 * nv + P, kAlign);
  c.indices = off; off += AlignUp(4 * (nv / gst::kSymsPerLane + nv / gst::kRowBlocks), kAlign);
  c.group_sums = off; off += AlignUp(4 * (nv / gst::kGroupSyms), kAlign);
  c.planes = 0;  // not read: the wavelet runs inside the assembly
  c.total = off;
  if (c.total > gst::kNoAllocScratchBytes) return GST_ERR_TOO_LARGE;
  StageConfig cfg = DefaultStageConfig(false);
  const uint8_t *cmp = static_cast<const uint8_t *>(d_cmp);
  const size_t offsets_bytes = AlignUp(32, 512);  // codec/decoder.cpp:109-111 for one image
  KernelArgs a;
  memset(&a, 0, sizeof(a));
  FillDesc(*hdr, cmp + offsets_bytes, cmp + offsets_bytes + 2048, base, c, 0, static_cast<uint8_t *>(d_out),
           false, &a.single);
  SetAnsForm(&cfg, GST_ANS_FORM_AUTO, AutoPairs(&a.single, 1));
  a.ans_unit_code = AnsUnitCode(cfg);
  a.err = d_err_bits;
  uint32_t grids[gst::kNumStages];
  for (uint32_t st = 0; st < gst::kNumStages; ++st)
    grids[st] = gst::SinglePassGrid(st, gst::StageUnitsOf(a.single, st, AnsUnitCode(cfg)), cfg);
  return LaunchAll(a, grids, cfg, static_cast<hipStream_t>(stream));
}

int gst_load_compressed_dxts(const gst_header *hdrs, uint32_t n, gst_stream stream, const void *d_cmp,
                             void *d_out, uint32_t num_init, const gst_event *init, gst_event *done) {
  return DecodeStaging(hdrs, n, static_cast<hipStream_t>(stream), d_cmp, d_out, false, num_init, init,
                       done);
}

int gst_load_rgb(const gst_header *hdr, gst_stream stream, const void *d_cmp, void *d_out,
                 uint32_t num_init, const gst_event *init, gst_event *done) {
  return DecodeStaging(hdr, 1, static_cast<hipStream_t>(stream), d_cmp, d_out, true, num_init, init,
                       done);
}

int gst_load_rgbs(const gst_header *hdrs, uint32_t n, gst_stream stream, const void *d_cmp,
                  void *d_out, uint32_t num_init, const gst_event *init, gst_event *done) {
  return DecodeStaging(hdrs, n, static_cast<hipStream_t>(stream), d_cmp, d_out, true, num_init, init,
                       done);
}

int gst_decompress_dxt(const uint8_t *file, size_t len, uint8_t *out, size_t out_len) {
  if (!file || !out) return GST_ERR_INVALID_ARGUMENT;
  gst_header h;
  int rc = gst_parse_header(file, len, &h);
  if (rc != GST_OK) return rc;
  const size_t dxt = static_cast<size_t>(h.width) * h.height / 2;
  if (out_len < dxt) return GST_ERR_INVALID_ARGUMENT;
  if ((rc = EnsureInit()) != GST_OK) return rc;
  // codec/decoder.cpp:430-476: offsets block of 512 B, then the file minus its header.
  const size_t body = sizeof(gst_header) + 2048 + StreamBytes(h) - sizeof(gst_header);
  std::vector<uint8_t> staging(512 + body, 0);
  const uint8_t *const files[1] = {file};
  const size_t lens[1] = {len};
  if ((rc = gst_pack_staging(files, lens, 1, staging.data(), staging.size())) != GST_OK) return rc;
  hipStream_t s = nullptr;
  GST_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  void *d_cmp = nullptr, *d_out = nullptr;
  hipError_t e = hipMalloc(&d_cmp, staging.size());
  if (e == hipSuccess) e = hipMalloc(&d_out, dxt);
  if (e == hipSuccess) e = hipMemcpyAsync(d_cmp, staging.data(), staging.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    rc = DecodeStaging(&h, 1, s, d_cmp, d_out, false, 0, nullptr, nullptr);
    if (rc == GST_OK) e = hipMemcpyAsync(out, d_out, dxt, hipMemcpyDeviceToHost, s);
    if (rc == GST_OK && e == hipSuccess) e = hipStreamSynchronize(s);
  }
  uint32_t bits = 0;
  if (e == hipSuccess && rc == GST_OK) rc = gst_read_error_bits(s, &bits);
  if (d_cmp) (void)hipFree(d_cmp);
  if (d_out) (void)hipFree(d_out);
  {
    Globals &g = G();
    std::lock_guard<std::mutex> lk(g.mu);
    auto it = g.err_words.find(s);
    if (it != g.err_words.end()) {
      (void)hipFree(it->second);
      g.err_words.erase(it);
    }
  }
  (void)hipStreamDestroy(s);
  if (e != hipSuccess) return HipFail(e);
  if (rc != GST_OK) return rc;
  return bits ? GST_ERR_CORRUPT_STREAM : GST_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched decode of textures of uneven size (no reference counterpart; the
// reference's LoadCompressedDXTs needs equal sizes, codec/decoder.cpp:105-121).

struct gst_batch {
  gst_batch_config cfg;
  StageConfig stage;
  uint32_t grids[gst::kNumStages];
  uint8_t *dmem = nullptr;   // plan blob | tables | status | err | scratch
  size_t dbytes = 0;
  size_t plan_off = 0, tables_off = 0, status_off = 0, err_off = 0, scratch_off = 0;
  size_t scratch_cap = 0;
  void *h_blob = nullptr;    // pinned staging of the plan blob (coherent: never in a GPU cache)
  void *h_blob_dev = nullptr;  // its address for device code
  size_t blob_bytes = 0;
  hipEvent_t blob_ev = nullptr;
  bool blob_pending = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t n_bound = 0;
  std::vector<ImageDesc> bound;  // descriptors of the bound images (for re-planning)
  bool auto_form = false;        // GST_ANS_FORM_AUTO: the rANS form is chosen at each bind
};

namespace {

KernelArgs BatchArgs(gst_batch *b) {
  return PlanArgs(b->dmem + b->plan_off, b->cfg.capacity_images,
                  reinterpret_cast<uint32_t *>(b->dmem + b->err_off),
                  reinterpret_cast<uint32_t *>(b->dmem + b->tables_off),
                  reinterpret_cast<uint32_t *>(b->dmem + b->status_off));
}

// Builds the plan of the bound images for the batch's rANS work-unit size and
// copies it to the device on `s` through the pinned staging buffer.
int UploadPlan(gst_batch *b, hipStream_t s, uint32_t units[gst::kNumStages]) {
  std::vector<uint8_t> blob;
  BuildPlan(b->bound, b->cfg.capacity_images, AnsUnitCode(b->stage), &blob, units);
  if (b->blob_pending) GST_HIP(hipEventSynchronize(b->blob_ev));
  memcpy(b->h_blob, blob.data(), blob.size());
  // The device reads the blob from pinned memory itself (one gather launch): a DMA copy's
  // fixed cost (about 10 us) would sit between every bind and its decode.
  gst::GatherArgs ga;
  ga.n = 1;
  ga.job[0] = {static_cast<const uint8_t *>(b->h_blob_dev), b->dmem + b->plan_off, blob.size()};
  GST_HIP(gst::LaunchGather(ga, s));
  GST_HIP(hipEventRecord(b->blob_ev, s));
  b->blob_pending = true;
  return GST_OK;
}

void DropGraph(gst_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Captures the five stage launches on `s` into *exec.
int CaptureStages(gst_batch *b, hipStream_t s, hipGraph_t *graph, hipGraphExec_t *exec) {
  GST_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const KernelArgs a = BatchArgs(b);
  hipError_t e = hipSuccess;
  for (uint32_t st = 0; st < gst::kNumStages && e == hipSuccess; ++st)
    e = gst::LaunchStage(st, a, b->grids[st], b->stage, s);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &g);
  if (e == hipSuccess) e = end;
  if (e == hipSuccess) e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    *exec = nullptr;
    return HipFail(e);
  }
  *graph = g;
  return GST_OK;
}

// Events that only time: recording one skips the system-scope fence (an L2 write-back
// and invalidate), which would leave the next kernel with cold caches and make every
// timed kernel of a decode several microseconds slower than it runs without events.
hipError_t TimingEvent(hipEvent_t *e) { return hipEventCreateWithFlags(e, hipEventDisableSystemFence); }

// Median time (ms) of `reps` back-to-back launches of one stage, after one
// untimed launch.
int TimeStage(gst_batch *b, uint32_t st, uint32_t grid, const StageConfig &cfg, hipStream_t s,
              uint32_t reps, float *ms) {
  const KernelArgs a = BatchArgs(b);
  std::vector<hipEvent_t> ev(reps + 1, nullptr);
  int rc = GST_OK;
  hipError_t e = gst::LaunchStage(st, a, grid, cfg, s);
  for (size_t i = 0; i < ev.size() && e == hipSuccess; ++i) e = TimingEvent(&ev[i]);
  if (e == hipSuccess) e = hipEventRecord(ev[0], s);
  for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
    e = gst::LaunchStage(st, a, grid, cfg, s);
    if (e == hipSuccess) e = hipEventRecord(ev[r + 1], s);
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev[reps]);
  std::vector<float> t(reps, 0.f);
  for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) e = hipEventElapsedTime(&t[r], ev[r], ev[r + 1]);
  if (e != hipSuccess) {
    rc = HipFail(e);
  } else {
    std::sort(t.begin(), t.end());
    *ms = t[reps / 2];
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  return rc;
}

}  // namespace

extern "C" {

int gst_batch_create(const gst_batch_config *cfg, gst_batch **out) {
  if (!cfg || !out || cfg->capacity_images == 0 || cfg->capacity_blocks == 0)
    return GST_ERR_INVALID_ARGUMENT;
  int rc = EnsureInit();
  if (rc != GST_OK) return rc;
  gst_batch *b = new gst_batch();
  b->cfg = *cfg;
  b->stage.ans_groups_per_wave = cfg->ans_groups_per_wave ? cfg->ans_groups_per_wave : 2;
  b->stage.ans_waves = cfg->ans_waves_per_group ? cfg->ans_waves_per_group : 4;
  b->stage.wavelet_waves = cfg->wavelet_waves_per_group ? cfg->wavelet_waves_per_group : 1;
  b->stage.ans_build_tables = (cfg->separate_table_stage || b->stage.ans_waves < 4) ? 0 : 1;
  if (cfg->ans_form > GST_ANS_FORM_PLAIN) {
    delete b;
    return GST_ERR_INVALID_ARGUMENT;
  }
  const bool lds_4x2 = b->stage.ans_build_tables && b->stage.ans_waves == 4 && b->stage.ans_groups_per_wave == 2;
  SetAnsForm(&b->stage, lds_4x2 ? cfg->ans_form : GST_ANS_FORM_PLAIN);
  if (!lds_4x2) b->stage.ans_auto = b->stage.ans_pair_short = 0;
  b->auto_form = lds_4x2 && cfg->ans_form == GST_ANS_FORM_AUTO;
  b->stage.fused_waves = (cfg->rgb_output || cfg->separate_index_scan)
                             ? 0
                             : (cfg->assembly_waves_per_group ? cfg->assembly_waves_per_group : 4);
  b->stage.fused_wide = cfg->assembly_wide ? 1 : 0;
  b->stage.tile_fused = (cfg->rgb_output || cfg->separate_index_scan || cfg->separate_wavelet) ? 0 : 1;
  b->stage.rgb = cfg->rgb_output ? 1 : 0;
  const uint32_t cus = static_cast<uint32_t>(G().cu_count > 0 ? G().cu_count : 256);
  b->grids[gst::kStageTables] = 4 * cfg->capacity_images;
  b->grids[gst::kStageAns] = cfg->grid_ans ? cfg->grid_ans : cus * 8;
  b->grids[gst::kStageIndices] = cfg->grid_indices ? cfg->grid_indices : cus * 4;
  b->grids[gst::kStageWavelet] = cfg->grid_wavelet ? cfg->grid_wavelet : cus * 16;
  // 16 per CU: past the 6 fused work-groups a CU holds at once, so the hardware keeps
  // every CU fed while units remain (a persistent 8 per CU left a round-trip tail).
  b->grids[gst::kStageAssemble] = cfg->grid_assemble ? cfg->grid_assemble : cus * 16;
  const uint64_t cap = cfg->capacity_images, B = cfg->capacity_blocks;
  const uint64_t Pcap = cfg->capacity_palette ? cfg->capacity_palette : 4 * B;
  size_t off = 0;
  b->blob_bytes = gst::PlanBytesFor(cfg->capacity_images);
  b->plan_off = off; off += AlignUp(b->blob_bytes, kAlign);
  b->tables_off = off; off += AlignUp(4 * cap * gst::kTableSize * 4, kAlign);
  b->status_off = off; off += AlignUp(4 * cap * 4, kAlign);
  b->err_off = off; off += kAlign;
  b->scratch_off = off;
  // Per image: symbols 7nv+P, planes 6nv, indices 4nv, group sums nv/2048 bytes,
  // each region padded to 256 B.
  b->scratch_cap = AlignUp(17 * B + Pcap + B / 2048 + cap * 4 * kAlign, kAlign);
  off += b->scratch_cap;
  b->dbytes = off;
  hipError_t e = hipMalloc(&b->dmem, b->dbytes);
  if (e == hipSuccess) e = hipMemset(b->dmem, 0, b->scratch_off);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_blob, b->blob_bytes, hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer(&b->h_blob_dev, b->h_blob, 0);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->blob_ev, hipEventDisableTiming);
  if (e != hipSuccess) {
    gst_batch_destroy(b);
    return HipFail(e);
  }
  *out = b;
  return GST_OK;
}

int gst_batch_bind(gst_batch *b, const gst_header *hdrs, uint32_t n, const void *const *d_files,
                   void *const *d_outs, gst_stream stream) {
  if (!b || !hdrs || !d_files || !d_outs || n == 0 || n > b->cfg.capacity_images)
    return GST_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint64_t blocks = 0, pal = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const int rc = CheckHeader(hdrs[i]);
    if (rc != GST_OK) return rc;
    if (!d_files[i] || !d_outs[i]) return GST_ERR_INVALID_ARGUMENT;
    blocks += NumBlocks(hdrs[i]);
    pal += hdrs[i].palette_bytes;
  }
  const uint64_t Pcap = b->cfg.capacity_palette ? b->cfg.capacity_palette : 4 * b->cfg.capacity_blocks;
  if (blocks > b->cfg.capacity_blocks || pal > Pcap) return GST_ERR_OUT_OF_MEMORY;
  std::vector<ImageDesc> descs(n);
  uint8_t *scratch = b->dmem + b->scratch_off;
  size_t off = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const Carve c = CarveFor(hdrs[i]);
    if (off + c.total > b->scratch_cap) return GST_ERR_OUT_OF_MEMORY;
    const uint8_t *f = static_cast<const uint8_t *>(d_files[i]);
    FillDesc(hdrs[i], f + sizeof(gst_header), f + sizeof(gst_header) + 2048, scratch + off, c, 4 * i,
             static_cast<uint8_t *>(d_outs[i]), b->stage.rgb != 0, &descs[i]);
    off += c.total;
  }
  // The form travels in the plan, so a captured graph replays whichever one a bind picks.
  if (b->auto_form) SetAnsForm(&b->stage, GST_ANS_FORM_AUTO, AutoPairs(descs.data(), n));
  b->bound = std::move(descs);
  uint32_t units[gst::kNumStages];
  const int rc = UploadPlan(b, s, units);
  if (rc != GST_OK) return rc;
  b->n_bound = n;
  return GST_OK;
}

int gst_batch_autotune(gst_batch *b, gst_stream stream, uint32_t reps, gst_tune_result *out) {
  if (!b || b->n_bound == 0) return GST_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  reps = std::max<uint32_t>(reps ? reps : 5, 1);
  const uint32_t cus = static_cast<uint32_t>(G().cu_count > 0 ? G().cu_count : 256);
  gst_tune_result res;
  memset(&res, 0, sizeof(res));
  StageConfig best = b->stage;
  uint32_t grids[gst::kNumStages];
  memcpy(grids, b->grids, sizeof(grids));
  uint32_t units[gst::kNumStages];
  int rc;
  float ms = 0.f;

  // Tables as their own launch (one work-group per table); the rANS shapes that load
  // them are charged this time, the ones that build their table in LDS are not.
  best.ans_build_tables = 0;
  best.ans_auto = best.ans_pair_short = 0;
  grids[gst::kStageTables] = 4 * b->n_bound;
  float t_tables = 0.f;
  if ((rc = TimeStage(b, gst::kStageTables, grids[gst::kStageTables], best, s, reps, &t_tables)) != GST_OK)
    return rc;

  // rANS: streams per wavefront (32 = one group per wave, 64 = two) x waves per
  // work-group x table source (mode 0: loaded, 1: built in LDS; 2 and 3: AnsAutoKernel
  // with and without short remainders paired, 4 waves x 2 groups) x persistent grid;
  // each shape needs its own plan (units of W*GPW groups).
  static const uint32_t kAnsShapes[][2] = {{1, 2}, {1, 4}, {1, 8}, {2, 1}, {2, 2}, {2, 4}, {2, 8}};
  float best_ms = 1e30f, best_ans = 0.f;
  for (const auto &shape : kAnsShapes) {
    for (uint32_t mode = 0; mode < 4; ++mode) {
      const uint32_t build = mode >= 1 ? 1u : 0u;
      if (build && shape[1] < 4) continue;  // the LDS build needs 256 threads
      if (mode >= 2 && !(shape[0] == 2 && shape[1] == 4)) continue;
      StageConfig c = best;
      c.ans_groups_per_wave = shape[0];
      c.ans_waves = shape[1];
      c.ans_build_tables = build;
      c.ans_auto = mode >= 2 ? 1u : 0u;
      c.ans_pair_short = mode == 2 ? 1u : 0u;
      b->stage = c;
      if ((rc = UploadPlan(b, s, units)) != GST_OK) return rc;
      const uint32_t cand[] = {cus, 2 * cus, 4 * cus, 8 * cus, units[gst::kStageAns]};
      for (uint32_t g : cand) {
        if ((rc = TimeStage(b, gst::kStageAns, g, c, s, reps, &ms)) != GST_OK) return rc;
        ++res.candidates;
        const float total = ms + (build ? 0.f : t_tables);
        if (total < best_ms) {
          best_ms = total;
          best_ans = ms;
          best.ans_groups_per_wave = c.ans_groups_per_wave;
          best.ans_waves = c.ans_waves;
          best.ans_build_tables = build;
          best.ans_auto = c.ans_auto;
          best.ans_pair_short = c.ans_pair_short;
          grids[gst::kStageAns] = g;
        }
      }
    }
  }
  b->stage = best;
  if ((rc = UploadPlan(b, s, units)) != GST_OK) return rc;
  // Leave valid symbols behind for the stages that read them.
  if ((rc = TimeStage(b, gst::kStageAns, grids[gst::kStageAns], best, s, reps, &ms)) != GST_OK) return rc;
  res.stage_ms[gst::kStageTables] = best.ans_build_tables ? 0.f : t_tables;
  res.stage_ms[gst::kStageAns] = best_ans;
  if (best.ans_build_tables) grids[gst::kStageTables] = 0;

  // Wavelet inside the DXT1 assembly (grid), timed before anything overwrites the
  // rANS stage's partial sums; the separate shapes follow.
  const bool try_tile_fused = !best.rgb && !b->cfg.separate_index_scan && !b->cfg.separate_wavelet;
  float tile_ms = 1e30f;
  uint32_t tile_grid = 0;
  if (try_tile_fused) {
    StageConfig c = best;
    c.tile_fused = 1;
    const uint32_t cand[] = {2 * cus, 4 * cus, 6 * cus, 8 * cus,
                             gst::SinglePassGrid(gst::kStageAssemble, units[gst::kStageAssemble], c)};
    for (uint32_t g : cand) {
      if ((rc = TimeStage(b, gst::kStageAssemble, g, c, s, reps, &ms)) != GST_OK) return rc;
      ++res.candidates;
      if (ms < tile_ms) {
        tile_ms = ms;
        tile_grid = g;
      }
    }
  }
  best.tile_fused = 0;

  // Wavelet: waves per work-group (4 tiles per wave) x grid.
  best_ms = 1e30f;
  for (uint32_t w : {1u, 2u, 4u}) {
    StageConfig c = best;
    c.wavelet_waves = w;
    const uint32_t cand[] = {2 * cus, 4 * cus, 8 * cus, 16 * cus,
                             gst::SinglePassGrid(gst::kStageWavelet, units[gst::kStageWavelet], c)};
    for (uint32_t g : cand) {
      if ((rc = TimeStage(b, gst::kStageWavelet, g, c, s, reps, &ms)) != GST_OK) return rc;
      ++res.candidates;
      if (ms < best_ms) {
        best_ms = ms;
        best.wavelet_waves = w;
        grids[gst::kStageWavelet] = g;
      }
    }
  }
  res.stage_ms[gst::kStageWavelet] = best_ms;

  // Index scan + assembly, fused (waves per work-group x grid) or as two launches
  // (grid each).  The fused shapes go first: they read the lane totals the rANS stage
  // left at the head of the index region, which the separate scan overwrites.
  float fused_ms = 1e30f;
  uint32_t fused_w = 0, fused_grid = 0, fused_wide = 0;
  if (!best.rgb) {
    for (uint32_t wide = 0; wide < 2; ++wide) {
      for (uint32_t w : {1u, 2u, 4u}) {
        StageConfig c = best;
        c.fused_waves = w;
        c.fused_wide = wide;
        const uint32_t cand[] = {2 * cus, 4 * cus, 8 * cus, 16 * cus,
                                 gst::SinglePassGrid(gst::kStageAssemble, units[gst::kStageAssemble], c)};
        for (uint32_t g : cand) {
          if ((rc = TimeStage(b, gst::kStageAssemble, g, c, s, reps, &ms)) != GST_OK) return rc;
          ++res.candidates;
          if (ms < fused_ms) {
            fused_ms = ms;
            fused_w = w;
            fused_wide = wide;
            fused_grid = g;
          }
        }
      }
    }
  }
  float sep_ms[2] = {1e30f, 1e30f};
  uint32_t sep_grid[2] = {0, 0};
  {
    StageConfig c = best;
    c.fused_waves = 0;
    const uint32_t st2[2] = {gst::kStageIndices, gst::kStageAssemble};
    for (int i = 0; i < 2; ++i) {
      const uint32_t cand[] = {2 * cus, 4 * cus, 8 * cus, 16 * cus, units[st2[i]]};
      for (uint32_t g : cand) {
        if ((rc = TimeStage(b, st2[i], g, c, s, reps, &ms)) != GST_OK) return rc;
        ++res.candidates;
        if (ms < sep_ms[i]) {
          sep_ms[i] = ms;
          sep_grid[i] = g;
        }
      }
    }
  }
  if (fused_ms < sep_ms[0] + sep_ms[1]) {
    best.fused_waves = fused_w;
    best.fused_wide = fused_wide;
    grids[gst::kStageIndices] = 0;
    grids[gst::kStageAssemble] = fused_grid;
    res.stage_ms[gst::kStageIndices] = 0.f;
    res.stage_ms[gst::kStageAssemble] = fused_ms;
  } else {
    best.fused_waves = 0;
    grids[gst::kStageIndices] = sep_grid[0];
    grids[gst::kStageAssemble] = sep_grid[1];
    res.stage_ms[gst::kStageIndices] = sep_ms[0];
    res.stage_ms[gst::kStageAssemble] = sep_ms[1];
  }
  const float separate_ms = res.stage_ms[gst::kStageWavelet] + res.stage_ms[gst::kStageIndices] +
                            res.stage_ms[gst::kStageAssemble];
  if (try_tile_fused && tile_ms < separate_ms) {
    best.tile_fused = 1;
    grids[gst::kStageWavelet] = 0;
    grids[gst::kStageIndices] = 0;
    grids[gst::kStageAssemble] = tile_grid;
    res.stage_ms[gst::kStageWavelet] = 0.f;
    res.stage_ms[gst::kStageIndices] = 0.f;
    res.stage_ms[gst::kStageAssemble] = tile_ms;
  }

  b->stage = best;
  memcpy(b->grids, grids, sizeof(grids));
  DropGraph(b);  // a captured graph holds the old shapes
  res.ans_groups_per_wave = best.ans_groups_per_wave;
  res.ans_waves_per_group = best.ans_waves;
  res.wavelet_waves_per_group = best.wavelet_waves;
  res.assembly_waves_per_group = best.fused_waves;
  res.ans_builds_tables = best.ans_build_tables;
  res.ans_form = AnsFormOf(best);
  b->auto_form = false;  // the tuned form stays for later binds
  res.assembly_wide = best.fused_wide;
  res.wavelet_fused = best.tile_fused;
  res.grid_ans = grids[gst::kStageAns];
  res.grid_indices = grids[gst::kStageIndices];
  res.grid_wavelet = grids[gst::kStageWavelet];
  res.grid_assemble = grids[gst::kStageAssemble];
  for (uint32_t st = 0; st < gst::kNumStages; ++st) {
    if (gst::StageLaunches(st, grids[st], best))
      res.total_ms += res.stage_ms[st];
    else
      res.stage_ms[st] = std::numeric_limits<float>::quiet_NaN();
  }
  if (out) *out = res;
  return GST_OK;
}

int gst_batch_decode(gst_batch *b, gst_stream stream) {
  if (!b || b->n_bound == 0) return GST_ERR_INVALID_ARGUMENT;
  return LaunchAll(BatchArgs(b), b->grids, b->stage, static_cast<hipStream_t>(stream));
}

int gst_batch_capture_graph(gst_batch *b, gst_stream stream) {
  if (!b || b->n_bound == 0) return GST_ERR_INVALID_ARGUMENT;
  if (b->exec) return GST_OK;
  return CaptureStages(b, static_cast<hipStream_t>(stream), &b->graph, &b->exec);
}

int gst_batch_launch_graph(gst_batch *b, gst_stream stream) {
  if (!b || !b->exec || b->n_bound == 0) return GST_ERR_INVALID_ARGUMENT;
  GST_HIP(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)));
  return GST_OK;
}

int gst_batch_decode_timed_reps(gst_batch *b, gst_stream stream, uint32_t reps, float ms[5]) {
  if (!b || !ms || b->n_bound == 0 || reps == 0 || reps > 4096) return GST_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t launched = 0;
  for (uint32_t st = 0; st < gst::kNumStages; ++st) launched += gst::StageLaunches(st, b->grids[st], b->stage);
  // One event before the first decode and one after every launched stage; nothing waits
  // on the host until the last, so the decodes run back to back.
  std::vector<hipEvent_t> ev(1 + static_cast<size_t>(reps) * launched, nullptr);
  hipError_t e = hipSuccess;
  for (auto &x : ev)
    if (e == hipSuccess) e = TimingEvent(&x);
  const KernelArgs a = BatchArgs(b);
  if (e == hipSuccess) e = hipEventRecord(ev[0], s);
  size_t k = 0;
  for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
    for (uint32_t st = 0; st < gst::kNumStages && e == hipSuccess; ++st) {
      if (!gst::StageLaunches(st, b->grids[st], b->stage)) continue;
      e = gst::LaunchStage(st, a, b->grids[st], b->stage, s);
      if (e == hipSuccess) e = hipEventRecord(ev[++k], s);
    }
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev.back());
  double sum[gst::kNumStages] = {};
  k = 0;
  for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) {
    for (uint32_t st = 0; st < gst::kNumStages && e == hipSuccess; ++st) {
      if (!gst::StageLaunches(st, b->grids[st], b->stage)) continue;
      float t = 0.f;
      e = hipEventElapsedTime(&t, ev[k], ev[k + 1]);
      sum[st] += t;
      ++k;
    }
  }
  for (uint32_t st = 0; st < gst::kNumStages; ++st)
    ms[st] = gst::StageLaunches(st, b->grids[st], b->stage)
                 ? static_cast<float>(sum[st] / reps)
                 : std::numeric_limits<float>::quiet_NaN();  // nothing launched: no time to report
  for (auto x : ev)
    if (x) (void)hipEventDestroy(x);
  return e == hipSuccess ? GST_OK : HipFail(e);
}

int gst_batch_decode_timed(gst_batch *b, gst_stream stream, float ms[5]) {
  return gst_batch_decode_timed_reps(b, stream, 1, ms);
}

int gst_batch_error_bits(gst_batch *b, gst_stream stream, uint32_t *bits) {
  if (!b || !bits) return GST_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  GST_HIP(hipStreamSynchronize(s));
  GST_HIP(hipMemcpy(bits, b->dmem + b->err_off, sizeof(uint32_t), hipMemcpyDeviceToHost));
  GST_HIP(hipMemset(b->dmem + b->err_off, 0, sizeof(uint32_t)));
  return GST_OK;
}

int gst_batch_error_bits_async(gst_batch *b, gst_stream stream, uint32_t *host_bits) {
  if (!b || !host_bits) return GST_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  GST_HIP(hipMemcpyAsync(host_bits, b->dmem + b->err_off, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  GST_HIP(hipMemsetAsync(b->dmem + b->err_off, 0, sizeof(uint32_t), s));
  return GST_OK;
}

size_t gst_batch_device_bytes(const gst_batch *b) { return b ? b->dbytes : 0; }

int gst_batch_destroy(gst_batch *b) {
  if (!b) return GST_ERR_INVALID_ARGUMENT;
  if (b->dmem) (void)hipDeviceSynchronize();
  DropGraph(b);
  if (b->blob_ev) (void)hipEventDestroy(b->blob_ev);
  if (b->h_blob) (void)hipHostFree(b->h_blob);
  if (b->dmem) (void)hipFree(b->dmem);
  delete b;
  return GST_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// One launch per stage, single image.

namespace {

int StageLaunch(uint32_t stage, const gst_header *hdr, const uint8_t *d_freqs, const uint8_t *d_streams,
                const uint32_t *d_tables, const uint8_t *d_symbols, const int32_t *d_group_sums,
                const int8_t *d_planes, const int32_t *d_indices, void *d_out, uint32_t *d_err,
                bool rgb, hipStream_t s) {
  int rc = EnsureInit();
  if (rc != GST_OK) return rc;
  gst_header h;
  if (hdr) {
    h = *hdr;
    if ((rc = CheckHeader(h)) != GST_OK) return rc;
  } else {
    // build_table needs no geometry; any legal header keeps the descriptor sane.
    h = gst_header{512, 256, 8192, 4 * 4 + 4 * 128, 8 * 4 + 8 * 128, 4 + 128, 4 + 128};
  }
  Carve c = CarveFor(h);
  ImageDesc d;
  FillDesc(h, d_freqs, d_streams ? d_streams : d_freqs, const_cast<uint8_t *>(d_symbols) - c.symbols, c,
           0, static_cast<uint8_t *>(d_out), rgb, &d);
  d.symbols = const_cast<uint8_t *>(d_symbols);
  d.planes = const_cast<int8_t *>(d_planes);
  d.indices = const_cast<int32_t *>(d_indices);
  d.group_sums = const_cast<int32_t *>(d_group_sums);
  d.lane_sums = nullptr;  // the per-stage buffers hold no room for lane totals
  KernelArgs a;
  memset(&a, 0, sizeof(a));
  a.single = d;
  a.err = d_err;
  a.tables = const_cast<uint32_t *>(d_tables);
  a.table_status = const_cast<uint32_t *>(d_tables) + 4 * gst::kTableSize;
  StageConfig cfg = DefaultStageConfig(rgb);
  cfg.fused_waves = 0;  // each stage is its own launch here
  cfg.tile_fused = 0;
  cfg.ans_build_tables = 0;
  cfg.ans_auto = 0;
  cfg.ans_pair_short = 0;
  const uint32_t grid = gst::SinglePassGrid(stage, gst::StageUnitsOf(d, stage, AnsUnitCode(cfg)), cfg);
  GST_HIP(gst::LaunchStage(stage, a, grid, cfg, s));
  return GST_OK;
}

}  // namespace

extern "C" {

int gst_stage_build_table(const void *d_freqs, uint32_t *d_tables, uint32_t *d_err, gst_stream stream) {
  if (!d_freqs || !d_tables || !d_err) return GST_ERR_INVALID_ARGUMENT;
  return StageLaunch(gst::kStageTables, nullptr, static_cast<const uint8_t *>(d_freqs), nullptr,
                     d_tables, nullptr, nullptr, nullptr, nullptr, nullptr, d_err, false,
                     static_cast<hipStream_t>(stream));
}

int gst_stage_ans_decode(const gst_header *hdr, const void *d_streams, const uint32_t *d_tables,
                         uint8_t *d_symbols, int32_t *d_group_sums, uint32_t *d_err, gst_stream stream) {
  if (!hdr || !d_streams || !d_tables || !d_symbols || !d_group_sums || !d_err)
    return GST_ERR_INVALID_ARGUMENT;
  return StageLaunch(gst::kStageAns, hdr, nullptr, static_cast<const uint8_t *>(d_streams), d_tables,
                     d_symbols, d_group_sums, nullptr, nullptr, nullptr, d_err, false,
                     static_cast<hipStream_t>(stream));
}

int gst_stage_decode_indices(const gst_header *hdr, const uint8_t *d_symbols, const int32_t *d_group_sums,
                             int32_t *d_indices, gst_stream stream) {
  if (!hdr || !d_symbols || !d_group_sums || !d_indices) return GST_ERR_INVALID_ARGUMENT;
  return StageLaunch(gst::kStageIndices, hdr, nullptr, nullptr, nullptr, d_symbols, d_group_sums,
                     nullptr, d_indices, nullptr, nullptr, false, static_cast<hipStream_t>(stream));
}

int gst_stage_inv_wavelet(const gst_header *hdr, const uint8_t *d_symbols, int8_t *d_planes,
                          gst_stream stream) {
  if (!hdr || !d_symbols || !d_planes) return GST_ERR_INVALID_ARGUMENT;
  return StageLaunch(gst::kStageWavelet, hdr, nullptr, nullptr, nullptr, d_symbols, nullptr, d_planes,
                     nullptr, nullptr, nullptr, false, static_cast<hipStream_t>(stream));
}

int gst_stage_assemble_dxt(const gst_header *hdr, const uint8_t *d_symbols, const int8_t *d_planes,
                           const int32_t *d_indices, void *d_out, uint32_t *d_err, gst_stream stream) {
  if (!hdr || !d_symbols || !d_planes || !d_indices || !d_out || !d_err) return GST_ERR_INVALID_ARGUMENT;
  return StageLaunch(gst::kStageAssemble, hdr, nullptr, nullptr, nullptr, d_symbols, nullptr, d_planes,
                     d_indices, d_out, d_err, false, static_cast<hipStream_t>(stream));
}

int gst_stage_assemble_rgb(const gst_header *hdr, const uint8_t *d_symbols, const int8_t *d_planes,
                           const int32_t *d_indices, void *d_out, uint32_t *d_err, gst_stream stream) {
  if (!hdr || !d_symbols || !d_planes || !d_indices || !d_out || !d_err) return GST_ERR_INVALID_ARGUMENT;
  return StageLaunch(gst::kStageAssemble, hdr, nullptr, nullptr, nullptr, d_symbols, nullptr, d_planes,
                     d_indices, d_out, d_err, true, static_cast<hipStream_t>(stream));
}

}  // extern "C"
