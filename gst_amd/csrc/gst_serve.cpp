// gst_serve.cpp -- continuous-batching decode server and load generator (gst_serve.h).
#include "gst_serve.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gst_decoder.h"
#include "gst_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double Micros(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::micro>(b - a).count();
}

constexpr uint64_t kAlign = 256;
uint64_t AlignUp(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Range allocator with coalescing (offsets into an arena).  Inputs are taken first-fit
// from the bottom and outputs last-fit from the top, so the inputs of requests admitted
// one after another tend to sit back to back (one copy for several of them) while
// outputs, which live longer, gather at the other end.
class RangeAllocator {
 public:
  void Reset(uint64_t size) {
    free_.clear();
    if (size) free_[0] = size;
    used_ = peak_ = 0;
  }
  bool Alloc(uint64_t n, uint64_t *off) {
    n = AlignUp(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const uint64_t rest = it->second - n, start = it->first + n;
      free_.erase(it);
      if (rest) free_[start] = rest;
      Take(n);
      return true;
    }
    return false;
  }
  bool AllocHigh(uint64_t n, uint64_t *off) {
    n = AlignUp(n);
    for (auto it = free_.rbegin(); it != free_.rend(); ++it) {
      if (it->second < n) continue;
      it->second -= n;
      *off = it->first + it->second;
      if (it->second == 0) free_.erase(std::next(it).base());
      Take(n);
      return true;
    }
    return false;
  }
  void Free(uint64_t off, uint64_t n) {
    n = AlignUp(n);
    if (n == 0) return;
    used_ -= n;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free_[off] = n;
  }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }

 private:
  void Take(uint64_t n) {
    used_ += n;
    peak_ = std::max(peak_, used_);
  }
  std::map<uint64_t, uint64_t> free_;
  uint64_t used_ = 0, peak_ = 0;
};

struct Req {
  uint64_t ticket = 0, tag = 0;
  gst_header hdr{};
  size_t len = 0;
  const uint8_t *d_prestaged = nullptr;   // device input given by the caller, or null
  const uint8_t *d_expected = nullptr;
  uint64_t host_in_off = 0, host_in_bytes = 0;    // pinned input staging (host inputs)
  uint64_t host_out_off = 0, host_out_bytes = 0;  // pinned output copy (copy_to_host)
  uint64_t dev_in_off = 0, dev_in_bytes = 0;      // arena: input (host inputs only)
  uint64_t dev_out_off = 0, out_bytes = 0;        // arena: output
  uint64_t in_bytes = 0;  // device input layout: the file, or the reference staging layout (plain)
  uint64_t copy_seq = 0;  // the copy-stream H2D copy that brings the input (0: none)
  bool copy_pending = false;  // admitted, input not copied yet: the launch copies it
  bool in_batch = false;      // taken into a batch (scheduler thread)
  bool admitted = false;
  Clock::time_point t_arrival, t_launch, t_dev, t_host;
  int status = GST_OK;
  uint32_t verified = 0;
  uint32_t batch_images = 0;
};

struct Slot {
  gst_batch *batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_dev = nullptr, ev_end = nullptr;
  uint32_t *err_host = nullptr;    // pinned
  uint32_t *flags_host = nullptr;  // pinned, one per batch position: the outputs' comparison
  uint32_t *flags_dev = nullptr;
  bool graph_ready = false;
  bool busy = false, dev_seen = false;
  std::vector<std::shared_ptr<Req>> reqs;
};

// Algorithmic bytes of one image through the batch's default two kernels: the rANS
// stage reads the streams and the four 512-byte frequency tables (tables are built in
// LDS) and writes the symbols and the index stream's partial sums; the fused wavelet +
// assembly reads the symbols and partial sums and writes the DXT1 blocks.
uint64_t PipelineBytes(const gst_header &h) {
  const uint64_t nv = static_cast<uint64_t>(h.width) * h.height / 16;
  const uint64_t S = static_cast<uint64_t>(h.y_cmp_sz) + h.chroma_cmp_sz + h.palette_sz + h.indices_sz;
  const uint64_t partial = nv / 64 + nv / 8;
  return (S + 2048) + (7 * nv + h.palette_bytes + partial) + (7 * nv + h.palette_bytes + partial) + 8 * nv;
}

// H2D copies in flight are marked by events in a ring, one per copy group; batches
// launch in admission order, so a batch waits for the group of its last input only.
// A batch that waits on an event re-recorded since by a later group waits for that
// later copy: longer than needed, never too short.
constexpr uint32_t kCopyEvents = 256;
// Inputs waiting for a slot go ahead on the copy stream once this many bytes of them
// have gathered: a host-to-device copy of a few hundred KiB runs far below the link's
// rate (fixed cost per copy), one of several MiB near it.
constexpr uint64_t kCopyGroupBytes = 8ull << 20;

}  // namespace

struct gst_server {
  gst_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_space;  // host staging freed
  std::condition_variable cv_work;   // new arrivals (scheduler wake-up)
  std::condition_variable cv_done;   // completions
  std::deque<std::shared_ptr<Req>> arrivals;  // FIFO: waiting for admission
  std::deque<std::shared_ptr<Req>> admitted;  // admitted, not yet launched
  std::map<uint64_t, std::shared_ptr<Req>> held;  // completed, not released
  std::vector<gst_completion> done;
  RangeAllocator dev_alloc, host_alloc;
  uint8_t *d_arena = nullptr, *h_staging = nullptr;
  uint8_t *h_staging_dev = nullptr;  // the staging's address for device code
  std::unique_ptr<uint8_t[]> dry_staging;  // a dry run's staging: ordinary memory, same copies
  std::vector<Slot> slots;
  hipStream_t copy_stream = nullptr;
  std::vector<hipEvent_t> copy_ev;  // kCopyEvents
  uint64_t copy_seq = 0;            // copy groups issued (scheduler thread)
  uint64_t launches = 0;            // batch launches (fault injection counts them)
  std::thread sched;
  std::atomic<bool> stop{false};
  uint64_t next_ticket = 1;
  uint64_t in_flight = 0;  // submitted, not completed
  uint32_t staging_copies = 0;  // submits copying into the staging without the lock
  hipEvent_t ev_ref = nullptr;
  std::vector<std::pair<double, double>> busy;  // GPU-timeline intervals (ms)
  gst_server_stats st{};
  double sum_batch = 0;
};

namespace {

int HipErr(hipError_t e) {
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? GST_ERR_OUT_OF_MEMORY : GST_ERR_HIP;
}

// Device input layout: batched decodes read the whole file; plain decodes read the
// reference staging layout (a 512-byte offsets block, then the file minus its 28-byte
// header).  The pinned staging holds the same layout, so a copy moves it unchanged.
uint64_t InputBytes(const gst_server &s, size_t len) {
  return s.cfg.plain ? 512 + len - sizeof(gst_header) : len;
}

uint8_t *InputPtr(const gst_server &s, const Req &r) {
  return r.d_prestaged ? const_cast<uint8_t *>(r.d_prestaged) : s.d_arena + r.dev_in_off;
}

uint8_t *OutputPtr(const gst_server &s, const Req &r) { return s.d_arena + r.dev_out_off; }

// Returns the request's arena ranges (a request never admitted holds none).
void FreeDeviceLocked(gst_server &s, Req &r) {
  if (r.dev_in_bytes) s.dev_alloc.Free(r.dev_in_off, r.dev_in_bytes);
  if (r.admitted) s.dev_alloc.Free(r.dev_out_off, r.out_bytes);
  r.dev_in_bytes = 0;
  r.admitted = false;
}

void PublishLocked(gst_server &s, const std::shared_ptr<Req> &r) {
  gst_completion c{};
  c.tag = r->tag;
  c.ticket = r->ticket;
  c.status = r->status;
  c.batch_images = r->batch_images;
  c.latency_device_us = Micros(r->t_arrival, r->t_dev);
  c.latency_host_us = s.cfg.copy_to_host ? Micros(r->t_arrival, r->t_host) : c.latency_device_us;
  c.queue_us = Micros(r->t_arrival, r->t_launch);
  c.output_bytes = r->out_bytes;
  c.d_output = s.d_arena ? OutputPtr(s, *r) : nullptr;
  c.verified = r->verified;
  s.done.push_back(c);
  s.st.completed++;
  if (r->status != GST_OK) s.st.failed++;
  if (r->verified) s.st.verified++;
  if (r->verified == 2) s.st.mismatches++;
  s.st.texels += static_cast<uint64_t>(r->hdr.width) * r->hdr.height;
  s.st.algorithmic_bytes += PipelineBytes(r->hdr);
  s.in_flight--;
  // The input is no longer needed once the decode has completed.
  if (r->host_in_bytes) s.host_alloc.Free(r->host_in_off, r->host_in_bytes);
  if (r->host_out_bytes) s.host_alloc.Free(r->host_out_off, r->host_out_bytes);
  r->host_in_bytes = r->host_out_bytes = 0;
  if (r->dev_in_bytes) s.dev_alloc.Free(r->dev_in_off, r->dev_in_bytes);
  r->dev_in_bytes = 0;
  if (s.cfg.auto_release || !r->admitted) {
    FreeDeviceLocked(s, *r);
  } else {
    s.held[r->ticket] = r;
  }
}

// Copies the inputs of newly admitted host requests to the arena on the copy stream,
// merging requests whose staging and arena ranges are both contiguous, and marks the
// group with one event (scheduler thread only).
// Coalesced H2D copies of the inputs of `reqs` (adjacent in both the pinned staging and
// the arena: one copy) on `stream`.  On the copy stream (`event`) the group gets an event
// that the launching slot's stream waits for; on a slot's own stream the decode that
// follows is ordered after them anyway.
int IssueCopies(gst_server &s, const std::vector<std::shared_ptr<Req>> &reqs, hipStream_t stream, bool event) {
  if (reqs.empty() || s.cfg.dry_run) return GST_OK;
  size_t i = 0;
  uint64_t bytes = 0, copies = 0;
  gst::GatherArgs ga;
  ga.n = 0;
  while (i < reqs.size()) {
    const Req &a = *reqs[i];
    uint64_t h0 = a.host_in_off, d0 = a.dev_in_off, n = a.in_bytes;
    size_t j = i + 1;
    for (; j < reqs.size(); ++j) {
      const Req &b = *reqs[j];
      const uint64_t gap = AlignUp(h0 + n) - (h0 + n);
      if (b.host_in_off != h0 + n + gap || b.dev_in_off != d0 + n + gap) break;
      n += gap + b.in_bytes;
    }
    hipError_t e = hipSuccess;
    if (s.cfg.copy_kernel) {
      ga.job[ga.n++] = {s.h_staging_dev + h0, s.d_arena + d0, n};
      if (ga.n == gst::kGatherJobs) {
        e = gst::LaunchGather(ga, stream);
        ga.n = 0;
      }
    } else {
      e = hipMemcpyAsync(s.d_arena + d0, s.h_staging + h0, n, hipMemcpyHostToDevice, stream);
    }
    if (e != hipSuccess) return HipErr(e);
    bytes += n;
    ++copies;
    i = j;
  }
  if (ga.n) {
    const hipError_t e = gst::LaunchGather(ga, stream);
    if (e != hipSuccess) return HipErr(e);
  }
  for (auto &r : reqs) r->copy_pending = false;
  if (event) {
    const uint64_t seq = ++s.copy_seq;
    const hipError_t e = hipEventRecord(s.copy_ev[seq % kCopyEvents], stream);
    if (e != hipSuccess) return HipErr(e);
    for (auto &r : reqs) r->copy_seq = seq;
  }
  std::lock_guard<std::mutex> lk(s.mu);
  s.st.h2d_bytes += bytes;
  s.st.h2d_copies += copies;
  return GST_OK;
}

// Enqueues one batch on a slot (scheduler thread only).
int LaunchBatch(gst_server &s, Slot &slot, std::vector<std::shared_ptr<Req>> &&reqs) {
  const auto now = Clock::now();
  uint64_t wait_seq = 0;
  std::vector<std::shared_ptr<Req>> to_copy;  // admitted in this round: copied on the slot's stream
  for (auto &r : reqs) {
    r->t_launch = now;
    r->batch_images = static_cast<uint32_t>(reqs.size());
    wait_seq = std::max(wait_seq, r->copy_seq);
    if (r->copy_pending) to_copy.push_back(r);
  }
  slot.reqs = std::move(reqs);
  slot.dev_seen = false;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    s.st.batches++;
    s.sum_batch += static_cast<double>(slot.reqs.size());
    s.st.max_batch_images = std::max<double>(s.st.max_batch_images, static_cast<double>(slot.reqs.size()));
  }
  s.launches++;
  if (s.cfg.dry_run) {
    if (s.cfg.fail_launch_every && s.launches % s.cfg.fail_launch_every == 0) return GST_ERR_HIP;
    slot.busy = true;
    return GST_OK;
  }
  hipStream_t st = slot.stream;
  hipError_t e = hipSuccess;
  if (wait_seq) {
    if ((e = hipStreamWaitEvent(st, s.copy_ev[wait_seq % kCopyEvents], 0)) != hipSuccess) return HipErr(e);
  }
  if (!to_copy.empty()) {
    const int rc = IssueCopies(s, to_copy, st, false);
    if (rc != GST_OK) return rc;
  }
  if (s.cfg.poison_outputs) {
    for (auto &r : slot.reqs)
      if ((e = hipMemsetAsync(OutputPtr(s, *r), 0xA5, r->out_bytes, st)) != hipSuccess) return HipErr(e);
  }
  if ((e = hipEventRecord(slot.ev_start, st)) != hipSuccess) return HipErr(e);
  int rc = GST_OK;
  if (s.cfg.plain) {
    // One request per batch; its own error bits follow it to pinned memory (after the
    // output-ready event: they are read at the batch's end).
    for (auto &r : slot.reqs) {
      rc = gst_load_compressed_dxt(&r->hdr, st, InputPtr(s, *r), OutputPtr(s, *r), 0, nullptr, nullptr);
      if (rc != GST_OK) return rc;
    }
    if ((e = hipEventRecord(slot.ev_dev, st)) != hipSuccess) return HipErr(e);
    if ((rc = gst_read_error_bits_async(st, slot.err_host)) != GST_OK) return rc;
  } else {
    std::vector<gst_header> hdrs;
    std::vector<const void *> files;
    std::vector<void *> outs;
    for (auto &r : slot.reqs) {
      hdrs.push_back(r->hdr);
      files.push_back(InputPtr(s, *r));
      outs.push_back(OutputPtr(s, *r));
    }
    rc = gst_batch_bind(slot.batch, hdrs.data(), static_cast<uint32_t>(hdrs.size()), files.data(),
                        outs.data(), st);
    if (s.cfg.graph_launch) {
      if (rc == GST_OK && !slot.graph_ready) {
        rc = gst_batch_capture_graph(slot.batch, st);
        slot.graph_ready = rc == GST_OK;
      }
      if (rc == GST_OK) rc = gst_batch_launch_graph(slot.batch, st);
    } else if (rc == GST_OK) {
      rc = gst_batch_decode(slot.batch, st);
    }
    if (rc != GST_OK) return rc;
    // Output ready in HBM here; the error bits' copy back (read at the batch's end) and
    // the optional comparison come after.
    if ((e = hipEventRecord(slot.ev_dev, st)) != hipSuccess) return HipErr(e);
    if ((rc = gst_batch_error_bits_async(slot.batch, st, slot.err_host)) != GST_OK) return rc;
  }
  // Comparison with the expected outputs, after the point the device latency stops.
  gst::CompareArgs ca;
  memset(&ca, 0, sizeof(ca));
  ca.flags = slot.flags_dev;
  bool compared = false;
  for (size_t i = 0; i < slot.reqs.size(); ++i) {
    const Req &r = *slot.reqs[i];
    if (!r.d_expected) continue;
    ca.job[ca.n++] = {OutputPtr(s, r), r.d_expected, r.out_bytes, i};
    compared = true;
    if (ca.n == gst::kCompareJobs) {
      if ((e = gst::LaunchCompare(ca, st)) != hipSuccess) return HipErr(e);
      ca.n = 0;
    }
  }
  if (ca.n && (e = gst::LaunchCompare(ca, st)) != hipSuccess) return HipErr(e);
  if (compared) {
    const size_t fb = sizeof(uint32_t) * slot.reqs.size();
    if ((e = hipMemcpyAsync(slot.flags_host, slot.flags_dev, fb, hipMemcpyDeviceToHost, st)) != hipSuccess)
      return HipErr(e);
    if ((e = hipMemsetAsync(slot.flags_dev, 0, fb, st)) != hipSuccess) return HipErr(e);
  }
  if (s.cfg.copy_to_host) {
    for (auto &r : slot.reqs) {
      e = hipMemcpyAsync(s.h_staging + r->host_out_off, OutputPtr(s, *r), r->out_bytes, hipMemcpyDeviceToHost,
                         st);
      if (e != hipSuccess) return HipErr(e);
    }
  }
  if ((e = hipEventRecord(slot.ev_end, st)) != hipSuccess) return HipErr(e);
  slot.busy = true;
  return GST_OK;
}

// Advances a busy slot; returns true when it published its batch.
bool PollSlot(gst_server &s, Slot &slot) {
  if (!slot.busy) return false;
  if (!s.cfg.dry_run) {
    if (!slot.dev_seen) {
      if (hipEventQuery(slot.ev_dev) != hipSuccess) return false;
      slot.dev_seen = true;
      const auto now = Clock::now();
      for (auto &r : slot.reqs) r->t_dev = now;
      float t0 = 0, t1 = 0;
      if (hipEventElapsedTime(&t0, s.ev_ref, slot.ev_start) == hipSuccess &&
          hipEventElapsedTime(&t1, s.ev_ref, slot.ev_dev) == hipSuccess) {
        std::lock_guard<std::mutex> lk(s.mu);
        s.busy.emplace_back(t0, t1);
      }
    }
    if (hipEventQuery(slot.ev_end) != hipSuccess) return false;
    const auto now = Clock::now();
    for (auto &r : slot.reqs) r->t_host = now;
    if (*slot.err_host) {
      for (auto &r : slot.reqs) r->status = GST_ERR_CORRUPT_STREAM;
      *slot.err_host = 0;
    }
    for (size_t i = 0; i < slot.reqs.size(); ++i)
      if (slot.reqs[i]->d_expected) slot.reqs[i]->verified = slot.flags_host[i] ? 2u : 1u;
  } else {
    const auto now = Clock::now();
    for (auto &r : slot.reqs) r->t_dev = r->t_host = now;
  }
  {
    std::lock_guard<std::mutex> lk(s.mu);
    for (auto &r : slot.reqs) PublishLocked(s, r);
  }
  slot.reqs.clear();
  slot.busy = false;
  s.cv_space.notify_all();
  s.cv_done.notify_all();
  return true;
}

// Fails every request of a batch whose launch failed part-way.  Whatever was already
// enqueued on the slot's stream (copies, kernels) may still read or write the batch's
// memory, so the stream is drained before any of it can be handed out again.
void FailBatch(gst_server &s, Slot &slot, int rc) {
  if (!s.cfg.dry_run) {
    (void)hipStreamSynchronize(s.copy_stream);
    (void)hipStreamSynchronize(slot.stream);
  }
  (void)hipGetLastError();
  std::lock_guard<std::mutex> lk(s.mu);
  const auto now = Clock::now();
  for (auto &r : slot.reqs) {
    r->status = rc;
    r->t_launch = r->t_dev = r->t_host = now;
    PublishLocked(s, r);
  }
  slot.reqs.clear();
  slot.busy = false;
  s.cv_space.notify_all();
  s.cv_done.notify_all();
}

void SchedulerLoop(gst_server *sp) {
  gst_server &s = *sp;
  if (!s.cfg.dry_run) (void)hipSetDevice(0);
  const uint64_t max_blocks = s.cfg.max_batch_blocks;
  auto last_work = Clock::now();
  std::vector<std::shared_ptr<Req>> backlog;  // admitted, in no batch, input not copied yet
  uint64_t backlog_bytes = 0;
  for (;;) {
    bool progress = false;
    for (auto &slot : s.slots) progress |= PollSlot(s, slot);
    // Admission (FIFO), then batch formation for every free slot.
    std::vector<std::vector<std::shared_ptr<Req>>> batches;
    std::vector<Slot *> targets;
    std::vector<std::shared_ptr<Req>> fresh;  // admitted now, input to copy
    {
      std::lock_guard<std::mutex> lk(s.mu);
      while (!s.arrivals.empty()) {
        auto &r = s.arrivals.front();
        uint64_t in_off = 0, out_off = 0;
        const bool copy = r->d_prestaged == nullptr;
        if (!s.dev_alloc.AllocHigh(r->out_bytes, &out_off)) {
          s.st.admission_waits++;
          break;
        }
        if (copy && !s.dev_alloc.Alloc(r->in_bytes, &in_off)) {
          s.dev_alloc.Free(out_off, r->out_bytes);
          s.st.admission_waits++;
          break;
        }
        r->dev_out_off = out_off;
        if (copy) {
          r->dev_in_off = in_off;
          r->dev_in_bytes = r->in_bytes;
          r->copy_pending = true;
          fresh.push_back(r);
        }
        r->admitted = true;
        s.admitted.push_back(r);
        s.arrivals.pop_front();
        s.st.peak_arena_bytes = std::max<uint64_t>(s.st.peak_arena_bytes, s.dev_alloc.used());
      }
    }
    {
      std::lock_guard<std::mutex> lk(s.mu);
      for (auto &slot : s.slots) {
        if (slot.busy || s.admitted.empty()) continue;
        std::vector<std::shared_ptr<Req>> b;
        uint64_t blocks = 0;
        const uint32_t max_images = s.cfg.plain ? 1u : s.cfg.max_batch_images;
        while (!s.admitted.empty() && b.size() < max_images) {
          auto &r = s.admitted.front();
          const uint64_t nv = static_cast<uint64_t>(r->hdr.width) * r->hdr.height / 16;
          if (!b.empty() && blocks + nv > max_blocks) break;
          blocks += nv;
          r->in_batch = true;
          b.push_back(r);
          s.admitted.pop_front();
        }
        batches.push_back(std::move(b));
        targets.push_back(&slot);
        slot.busy = true;  // reserved until LaunchBatch runs below
      }
    }
    // Inputs that no free slot took join the backlog, which goes ahead on the copy stream
    // (overlapping the decodes in flight) once it holds kCopyGroupBytes; a request a slot
    // takes first is copied on that slot's stream right before its decode (no
    // cross-stream wait on the launch path).
    for (auto &r : fresh) {
      if (r->in_batch) continue;
      backlog.push_back(r);
      backlog_bytes += r->in_bytes;
    }
    if (!backlog.empty() && !batches.empty()) {  // drop what a slot took (copied at launch)
      backlog.erase(std::remove_if(backlog.begin(), backlog.end(),
                                   [](const std::shared_ptr<Req> &r) { return r->in_batch; }),
                    backlog.end());
      backlog_bytes = 0;
      for (auto &r : backlog) backlog_bytes += r->in_bytes;
    }
    if (backlog_bytes >= kCopyGroupBytes) {
      const int rc = IssueCopies(s, backlog, s.copy_stream, true);
      std::vector<std::shared_ptr<Req>> copied;
      copied.swap(backlog);
      backlog_bytes = 0;
      if (rc != GST_OK) {
        // The copies may be partly enqueued: drain the copy stream, then fail the
        // requests that were to use them.
        (void)hipStreamSynchronize(s.copy_stream);
        std::lock_guard<std::mutex> lk(s.mu);
        const auto now = Clock::now();
        for (auto &r : copied) {
          auto it = std::find(s.admitted.begin(), s.admitted.end(), r);
          if (it != s.admitted.end()) s.admitted.erase(it);
          r->status = rc;
          r->t_launch = r->t_dev = r->t_host = now;
          PublishLocked(s, r);
        }
        s.cv_space.notify_all();
        s.cv_done.notify_all();
      }
      progress = true;
    }
    for (size_t i = 0; i < batches.size(); ++i) {
      Slot &slot = *targets[i];
      const auto t0 = Clock::now();
      const int rc = LaunchBatch(s, slot, std::move(batches[i]));
      const double us = Micros(t0, Clock::now());
      {
        std::lock_guard<std::mutex> lk(s.mu);
        s.st.launch_host_ms += us * 1e-3;
      }
      if (rc != GST_OK) FailBatch(s, slot, rc);
      progress = true;
    }
    bool busy = false;
    for (auto &slot : s.slots) busy |= slot.busy;
    if (progress || busy) last_work = Clock::now();
    if (!progress) {
      std::unique_lock<std::mutex> lk(s.mu);
      if (s.stop && !busy && s.admitted.empty()) {
        // Shutting down: requests still waiting for memory that only a release could
        // free are completed with GST_ERR_BUSY.
        const auto now = Clock::now();
        for (auto &r : s.arrivals) {
          r->status = GST_ERR_BUSY;
          r->t_launch = r->t_dev = r->t_host = now;
          PublishLocked(s, r);
        }
        s.arrivals.clear();
        s.cv_space.notify_all();
        s.cv_done.notify_all();
        break;
      }
      if (!busy && s.arrivals.empty() && s.admitted.empty()) {
        if (Micros(last_work, Clock::now()) < s.cfg.idle_spin_us) {
          lk.unlock();
          std::this_thread::yield();  // still polling: an arrival starts without a wake-up
        } else {
          s.cv_work.wait_for(lk, std::chrono::milliseconds(2));
        }
      } else if (!busy) {
        // Waiting for memory that only a release can free.
        s.cv_work.wait_for(lk, std::chrono::microseconds(200));
      } else {
        lk.unlock();
        std::this_thread::yield();
      }
    }
  }
}

int Submit(gst_server *s, const gst_request &q, uint64_t *ticket) {
  if (!s || !q.file || (q.d_expected && (reinterpret_cast<uintptr_t>(q.d_expected) & 15)))
    return GST_ERR_INVALID_ARGUMENT;
  auto r = std::make_shared<Req>();
  r->t_arrival = Clock::now();
  // gst_parse_header reads only the 28-byte header and checks the stream sizes against
  // the stated length, which for a pre-staged file is the device copy's.
  const int rc = gst_parse_header(q.file, q.len, &r->hdr);
  if (q.d_file && s->cfg.plain) return GST_ERR_INVALID_ARGUMENT;  // plain decodes read the staging layout
  if (rc != GST_OK) return rc;
  r->tag = q.tag;
  r->len = q.len;
  r->d_prestaged = static_cast<const uint8_t *>(q.d_file);
  r->d_expected = static_cast<const uint8_t *>(q.d_expected);
  r->in_bytes = InputBytes(*s, q.len);
  r->out_bytes = static_cast<uint64_t>(r->hdr.width) * r->hdr.height / 2;
  r->host_in_bytes = q.d_file ? 0 : r->in_bytes;
  r->host_out_bytes = s->cfg.copy_to_host ? r->out_bytes : 0;
  const uint64_t dev = (q.d_file ? 0 : AlignUp(r->in_bytes)) + AlignUp(r->out_bytes);
  const uint64_t host = AlignUp(r->host_in_bytes) + AlignUp(r->host_out_bytes);
  const uint64_t nv = static_cast<uint64_t>(r->hdr.width) * r->hdr.height / 16;
  if (dev > s->cfg.hbm_budget_bytes || host > s->cfg.host_staging_bytes || nv > s->cfg.max_batch_blocks)
    return GST_ERR_OUT_OF_MEMORY;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool ok_in = true, ok_out = true;
    if (r->host_in_bytes) ok_in = s->host_alloc.Alloc(r->host_in_bytes, &r->host_in_off);
    if (ok_in && r->host_out_bytes) {
      ok_out = s->host_alloc.AllocHigh(r->host_out_bytes, &r->host_out_off);
      if (!ok_out && r->host_in_bytes) s->host_alloc.Free(r->host_in_off, r->host_in_bytes);
    }
    if (ok_in && ok_out) break;
    if (s->stop) return GST_ERR_BUSY;
    s->cv_space.wait(lk);
  }
  if (s->h_staging && r->host_in_bytes) {
    // The staging range is this request's alone until it is queued: the copy runs
    // without the lock, so submitters copy in parallel and the scheduler never waits.
    // Destroy waits for such copies before it frees the staging.
    s->staging_copies++;
    lk.unlock();
    uint8_t *dst = s->h_staging + r->host_in_off;
    if (s->cfg.plain) {
      // Offsets block content is not read: the decoder derives offsets from the header.
      memcpy(dst + 512, q.file + sizeof(gst_header), q.len - sizeof(gst_header));
    } else {
      memcpy(dst, q.file, q.len);
    }
    lk.lock();
    s->staging_copies--;
    if (s->stop) {  // shutting down: nothing would complete a request queued now
      s->host_alloc.Free(r->host_in_off, r->host_in_bytes);
      if (r->host_out_bytes) s->host_alloc.Free(r->host_out_off, r->host_out_bytes);
      s->cv_space.notify_all();
      return GST_ERR_BUSY;
    }
  }
  r->ticket = s->next_ticket++;
  if (ticket) *ticket = r->ticket;
  s->arrivals.push_back(r);
  s->in_flight++;
  s->st.submitted++;
  if (q.d_file) s->st.prestaged++;
  lk.unlock();
  s->cv_work.notify_one();
  return GST_OK;
}

}  // namespace

extern "C" {

int gst_server_create(const gst_server_config *cfg, gst_server **out) {
  if (!cfg || !out || cfg->hbm_budget_bytes == 0) return GST_ERR_INVALID_ARGUMENT;
  std::unique_ptr<gst_server> s(new gst_server());
  s->cfg = *cfg;
  // Four batches in flight keep the device busier than three under a saturating load
  // (size mix, pre-staged, 512 outstanding, one box: 1.32 against 1.23 M Mtexel/s; six no
  // better, eight slower).
  if (!s->cfg.slots) s->cfg.slots = s->cfg.plain ? 1 : 4;
  if (!s->cfg.max_batch_images) s->cfg.max_batch_images = 256;
  if (!s->cfg.max_batch_blocks) s->cfg.max_batch_blocks = 16ull << 20;
  if (!s->cfg.host_staging_bytes) s->cfg.host_staging_bytes = 1ull << 30;
  s->dev_alloc.Reset(s->cfg.hbm_budget_bytes);
  s->host_alloc.Reset(s->cfg.host_staging_bytes);
  s->slots.resize(s->cfg.slots);
  if (s->cfg.dry_run) {
    s->dry_staging.reset(new uint8_t[s->cfg.host_staging_bytes]);  // pages touched only where copied
    s->h_staging = s->dry_staging.get();
  } else {
    int rc = gst_initialize_decoder(0);
    if (rc != GST_OK) return rc;
    const uint32_t max_images = s->cfg.plain ? 1u : s->cfg.max_batch_images;
    hipError_t e = hipMalloc(&s->d_arena, s->cfg.hbm_budget_bytes);
    // The gather kernel reads the staging directly: coherent memory is never held in a
    // GPU cache, so a staging range refilled by a later submit is always read afresh.
    const unsigned staging_flags = s->cfg.copy_kernel ? hipHostMallocCoherent : hipHostMallocDefault;
    if (e == hipSuccess) e = hipHostMalloc(&s->h_staging, s->cfg.host_staging_bytes, staging_flags);
    if (e == hipSuccess && s->cfg.copy_kernel)
      e = hipHostGetDevicePointer(reinterpret_cast<void **>(&s->h_staging_dev), s->h_staging, 0);
    if (e == hipSuccess) e = hipEventCreate(&s->ev_ref);
    if (e == hipSuccess) e = hipEventRecord(s->ev_ref, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(s->ev_ref);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking);
    s->copy_ev.assign(kCopyEvents, nullptr);
    for (auto &ev : s->copy_ev)
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    for (auto &slot : s->slots) {
      if (e != hipSuccess) break;
      e = hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreate(&slot.ev_start);
      if (e == hipSuccess) e = hipEventCreate(&slot.ev_dev);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&slot.ev_end, hipEventDisableTiming);
      if (e == hipSuccess) e = hipHostMalloc(&slot.err_host, sizeof(uint32_t), hipHostMallocDefault);
      if (e == hipSuccess) *slot.err_host = 0;
      if (e == hipSuccess) e = hipHostMalloc(&slot.flags_host, sizeof(uint32_t) * max_images, hipHostMallocDefault);
      if (e == hipSuccess) e = hipMalloc(&slot.flags_dev, sizeof(uint32_t) * max_images);
      if (e == hipSuccess) e = hipMemset(slot.flags_dev, 0, sizeof(uint32_t) * max_images);
      if (e == hipSuccess && !s->cfg.plain) {
        gst_batch_config bc{};
        bc.capacity_images = s->cfg.max_batch_images;
        bc.capacity_blocks = s->cfg.max_batch_blocks;
        bc.ans_groups_per_wave = s->cfg.ans_groups_per_wave;
        bc.ans_waves_per_group = s->cfg.ans_waves_per_group;
        rc = gst_batch_create(&bc, &slot.batch);
        if (rc != GST_OK) {
          gst_server_destroy(s.release());
          return rc;
        }
      }
    }
    if (e != hipSuccess) {
      gst_server_destroy(s.release());
      return HipErr(e);
    }
  }
  gst_server *raw = s.release();
  raw->sched = std::thread(SchedulerLoop, raw);
  *out = raw;
  return GST_OK;
}

int gst_server_submit(gst_server *s, const uint8_t *file, size_t len, uint64_t tag, uint64_t *ticket) {
  if (!s || !file) return GST_ERR_INVALID_ARGUMENT;
  gst_request q{};
  q.file = file;
  q.len = len;
  q.tag = tag;
  return Submit(s, q, ticket);
}

int gst_server_submit_request(gst_server *s, const gst_request *req, uint64_t *ticket) {
  if (!s || !req) return GST_ERR_INVALID_ARGUMENT;
  return Submit(s, *req, ticket);
}

int gst_server_poll(gst_server *s, gst_completion *out, uint32_t max, uint32_t *n) {
  if (!s || !n || (max && !out)) return GST_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  const uint32_t k = static_cast<uint32_t>(std::min<size_t>(max, s->done.size()));
  std::copy(s->done.begin(), s->done.begin() + k, out);
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return GST_OK;
}

int gst_server_drain(gst_server *s) {
  if (!s) return GST_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [s] { return s->in_flight == 0; });
  return GST_OK;
}

int gst_server_release(gst_server *s, uint64_t ticket) {
  if (!s) return GST_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->held.find(ticket);
  if (it == s->held.end()) return GST_ERR_INVALID_ARGUMENT;
  FreeDeviceLocked(*s, *it->second);
  s->held.erase(it);
  s->cv_work.notify_one();
  return GST_OK;
}

int gst_server_stats_get(gst_server *s, gst_server_stats *st) {
  if (!s || !st) return GST_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->st;
  st->mean_batch_images = s->st.batches ? s->sum_batch / static_cast<double>(s->st.batches) : 0.0;
  // Union of the GPU-timeline batch intervals.
  std::vector<std::pair<double, double>> iv = s->busy;
  std::sort(iv.begin(), iv.end());
  double total = 0, cur_s = -1, cur_e = -1;
  for (auto &p : iv) {
    if (p.first > cur_e) {
      if (cur_e > cur_s) total += cur_e - cur_s;
      cur_s = p.first;
      cur_e = p.second;
    } else {
      cur_e = std::max(cur_e, p.second);
    }
  }
  if (cur_e > cur_s) total += cur_e - cur_s;
  st->device_busy_ms = total;
  return GST_OK;
}

int gst_server_destroy(gst_server *s) {
  if (!s) return GST_ERR_INVALID_ARGUMENT;
  if (s->sched.joinable()) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->stop = true;
      s->cv_space.notify_all();
      s->cv_space.wait(lk, [s] { return s->staging_copies == 0; });  // no submit still copying
    }
    s->cv_work.notify_all();
    s->cv_space.notify_all();
    s->sched.join();
  }
  if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
  for (auto &slot : s->slots) {
    if (slot.stream) (void)hipStreamSynchronize(slot.stream);
    if (slot.batch) gst_batch_destroy(slot.batch);
    if (slot.ev_start) (void)hipEventDestroy(slot.ev_start);
    if (slot.ev_dev) (void)hipEventDestroy(slot.ev_dev);
    if (slot.ev_end) (void)hipEventDestroy(slot.ev_end);
    if (slot.err_host) (void)hipHostFree(slot.err_host);
    if (slot.flags_host) (void)hipHostFree(slot.flags_host);
    if (slot.flags_dev) (void)hipFree(slot.flags_dev);
    if (slot.stream) (void)hipStreamDestroy(slot.stream);
  }
  for (auto ev : s->copy_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  if (s->ev_ref) (void)hipEventDestroy(s->ev_ref);
  if (s->d_arena) (void)hipFree(s->d_arena);
  if (s->h_staging && !s->dry_staging) (void)hipHostFree(s->h_staging);
  delete s;
  return GST_OK;
}

int gst_server_run_load(gst_server *s, const gst_load_spec *spec, double *lat_dev, double *lat_host,
                        uint64_t lat_cap, gst_load_result *res) {
  // Requires auto_release: a submit blocked on staging space must not wait on device
  // memory that only this thread's releases could free.
  if (!s || !spec || !spec->pool || !spec->pool_lens || spec->pool_size == 0 || spec->duration_s <= 0 ||
      (!spec->closed_loop && spec->qps <= 0) || !s->cfg.auto_release)
    return GST_ERR_INVALID_ARGUMENT;
  const uint32_t pool_size = spec->pool_size;
  // SplitMix64 draws: file choice and exponential inter-arrival gaps.
  auto next_u64 = [](uint64_t &state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto uniform = [&](uint64_t &state) {
    return (static_cast<double>(next_u64(state) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  uint64_t state = spec->seed;
  std::vector<double> cdf(pool_size);
  double acc = 0;
  for (uint32_t i = 0; i < pool_size; ++i) {
    acc += spec->pool_weights ? spec->pool_weights[i] : 1.0;
    cdf[i] = acc;
  }
  auto pick = [&](uint64_t &st) {
    const double u = uniform(st) * acc;
    return static_cast<uint32_t>(std::min<size_t>(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin(),
                                                  pool_size - 1));
  };
  gst_load_result out{};
  std::vector<gst_completion> buf(256);
  auto collect = [&]() {
    uint32_t n = 0;
    for (;;) {
      gst_server_poll(s, buf.data(), static_cast<uint32_t>(buf.size()), &n);
      for (uint32_t i = 0; i < n; ++i) {
        if (out.requests < lat_cap) {
          if (lat_dev) lat_dev[out.requests] = buf[i].latency_device_us;
          if (lat_host) lat_host[out.requests] = buf[i].latency_host_us;
        }
        out.requests++;
        if (buf[i].status != GST_OK) out.failed++;
        if (buf[i].verified) out.verified++;
        if (buf[i].verified == 2) out.mismatches++;
      }
      if (n < buf.size()) return n;
    }
  };
  auto submit = [&](uint64_t tag, uint64_t &st) {
    const uint32_t i = pick(st);
    gst_request q{};
    q.file = spec->pool[i];
    q.len = spec->pool_lens[i];
    q.tag = tag;
    q.d_file = spec->pool_device ? spec->pool_device[i] : nullptr;
    q.d_expected = spec->pool_expected ? spec->pool_expected[i] : nullptr;
    return Submit(s, q, nullptr);
  };
  const auto t0 = Clock::now();
  const auto t_end =
      t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(spec->duration_s));
  int rc = GST_OK;
  uint64_t submitted = 0;
  if (!spec->closed_loop) {
    auto next = t0;
    while (next < t_end && rc == GST_OK) {
      while (Clock::now() < next) {
        collect();
        if (next - Clock::now() > std::chrono::microseconds(300))
          std::this_thread::sleep_for(std::chrono::microseconds(100));
      }
      rc = submit(submitted++, state);
      const double gap = -std::log(uniform(state)) / spec->qps;
      next += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(gap));
    }
  } else if (spec->submit_threads <= 1) {
    for (uint32_t k = 0; k < spec->closed_loop && rc == GST_OK; ++k) rc = submit(submitted++, state);
    while (Clock::now() < t_end && rc == GST_OK) {
      const uint64_t before = out.requests;
      collect();
      for (uint64_t k = before; k < out.requests && rc == GST_OK; ++k) rc = submit(submitted++, state);
      if (out.requests == before) std::this_thread::yield();
    }
  } else {
    // This thread collects completions and hands each one back as a submission credit;
    // the submitting threads (draws seeded seed + thread index) take credits and submit.
    std::atomic<int64_t> credits{static_cast<int64_t>(spec->closed_loop)};
    std::atomic<uint64_t> next_tag{0};
    std::atomic<bool> finish{false};
    std::atomic<int> first_err{GST_OK};
    std::vector<std::thread> workers;
    for (uint32_t t = 0; t < spec->submit_threads; ++t) {
      workers.emplace_back([&, t]() {
        uint64_t st = spec->seed + 0x632BE59BD9B4E019ull * (t + 1);
        while (!finish.load(std::memory_order_acquire)) {
          int64_t c = credits.load(std::memory_order_relaxed);
          if (c <= 0 || !credits.compare_exchange_weak(c, c - 1, std::memory_order_acq_rel)) {
            std::this_thread::yield();
            continue;
          }
          const int e = submit(next_tag.fetch_add(1, std::memory_order_relaxed), st);
          if (e != GST_OK) {
            int ok = GST_OK;
            first_err.compare_exchange_strong(ok, e);
            finish.store(true, std::memory_order_release);
          }
        }
      });
    }
    while (Clock::now() < t_end && !finish.load(std::memory_order_acquire)) {
      const uint64_t before = out.requests;
      collect();
      if (out.requests > before)
        credits.fetch_add(static_cast<int64_t>(out.requests - before), std::memory_order_acq_rel);
      else
        std::this_thread::yield();
    }
    finish.store(true, std::memory_order_release);
    for (auto &w : workers) w.join();
    rc = first_err.load();
  }
  gst_server_drain(s);
  collect();
  out.wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
  if (res) *res = out;
  return rc;
}

}  // extern "C"
