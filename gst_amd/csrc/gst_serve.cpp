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

namespace {

using Clock = std::chrono::steady_clock;

double Micros(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::micro>(b - a).count();
}

constexpr uint64_t kAlign = 256;
uint64_t AlignUp(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit range allocator with coalescing (offsets into an arena).
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
      used_ += n;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void Free(uint64_t off, uint64_t n) {
    n = AlignUp(n);
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
  std::map<uint64_t, uint64_t> free_;
  uint64_t used_ = 0, peak_ = 0;
};

struct Req {
  uint64_t ticket = 0, tag = 0;
  gst_header hdr{};
  size_t len = 0;
  uint64_t host_off = 0, host_bytes = 0;   // pinned input staging (+ output if copied back)
  uint64_t dev_off = 0, dev_bytes = 0;     // arena: input then output
  uint64_t in_bytes = 0, out_bytes = 0;
  bool admitted = false;
  Clock::time_point t_arrival, t_launch, t_dev, t_host;
  int status = GST_OK;
  uint32_t batch_images = 0;
};

enum SlotState { kIdle, kDevPending, kHostPending };

struct Slot {
  gst_batch *batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_dev = nullptr, ev_host = nullptr;
  uint32_t *err_host = nullptr;  // pinned
  bool graph_ready = false;
  SlotState state = kIdle;
  std::vector<std::shared_ptr<Req>> reqs;
};

// Algorithmic bytes of one image through the batch's default two kernels: the rANS
// stage reads the streams and the four 8 KiB tables' frequencies and writes the
// symbols and the index stream's partial sums; the fused wavelet + assembly reads the
// symbols and partial sums and writes the DXT1 blocks.
uint64_t PipelineBytes(const gst_header &h) {
  const uint64_t nv = static_cast<uint64_t>(h.width) * h.height / 16;
  const uint64_t S = static_cast<uint64_t>(h.y_cmp_sz) + h.chroma_cmp_sz + h.palette_sz + h.indices_sz;
  const uint64_t partial = nv / 64 + nv / 8;
  return (S + 32768) + (7 * nv + h.palette_bytes + partial) + (7 * nv + h.palette_bytes + partial) + 8 * nv;
}

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
  std::vector<Slot> slots;
  std::thread sched;
  std::atomic<bool> stop{false};
  uint64_t next_ticket = 1;
  uint64_t in_flight = 0;  // submitted, not completed
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

uint64_t DevBytes(const gst_server &s, const gst_header &h, size_t len, uint64_t *in_bytes,
                  uint64_t *out_bytes) {
  // Batched: the whole file.  Plain: the reference staging layout (offsets block of
  // 512 B, then the file minus its 28-byte header).
  *in_bytes = AlignUp(s.cfg.plain ? 512 + len - sizeof(gst_header) : len);
  *out_bytes = static_cast<uint64_t>(h.width) * h.height / 2;
  return *in_bytes + AlignUp(*out_bytes);
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
  c.d_output = s.d_arena ? s.d_arena + r->dev_off + r->in_bytes : nullptr;
  s.done.push_back(c);
  s.st.completed++;
  if (r->status != GST_OK) s.st.failed++;
  s.st.texels += static_cast<uint64_t>(r->hdr.width) * r->hdr.height;
  s.st.algorithmic_bytes += PipelineBytes(r->hdr);
  s.in_flight--;
  if (s.cfg.auto_release) {
    s.dev_alloc.Free(r->dev_off, r->dev_bytes);
  } else {
    s.held[r->ticket] = r;
  }
}

// Enqueues one batch on a slot (scheduler thread only).
int LaunchBatch(gst_server &s, Slot &slot, std::vector<std::shared_ptr<Req>> &&reqs) {
  const auto now = Clock::now();
  for (auto &r : reqs) {
    r->t_launch = now;
    r->batch_images = static_cast<uint32_t>(reqs.size());
  }
  slot.reqs = std::move(reqs);
  {
    std::lock_guard<std::mutex> lk(s.mu);
    s.st.batches++;
    s.sum_batch += static_cast<double>(slot.reqs.size());
    s.st.max_batch_images = std::max<double>(s.st.max_batch_images, static_cast<double>(slot.reqs.size()));
  }
  if (s.cfg.dry_run) {
    slot.state = kDevPending;
    return GST_OK;
  }
  hipStream_t st = slot.stream;
  hipError_t e = hipSuccess;
  for (auto &r : slot.reqs) {
    uint8_t *dst = s.d_arena + r->dev_off;
    const uint8_t *src = s.h_staging + r->host_off;
    if (s.cfg.plain) {
      // Offsets block content is not read: the decoder derives offsets from the header.
      e = hipMemcpyAsync(dst + 512, src + sizeof(gst_header), r->len - sizeof(gst_header),
                         hipMemcpyHostToDevice, st);
    } else {
      e = hipMemcpyAsync(dst, src, r->len, hipMemcpyHostToDevice, st);
    }
    if (e != hipSuccess) return HipErr(e);
  }
  if ((e = hipEventRecord(slot.ev_start, st)) != hipSuccess) return HipErr(e);
  int rc = GST_OK;
  if (s.cfg.plain) {
    for (auto &r : slot.reqs) {
      uint8_t *base = s.d_arena + r->dev_off;
      rc = gst_load_compressed_dxt(&r->hdr, st, base, base + r->in_bytes, 0, nullptr, nullptr);
      if (rc != GST_OK) return rc;
    }
  } else {
    std::vector<gst_header> hdrs;
    std::vector<const void *> files;
    std::vector<void *> outs;
    for (auto &r : slot.reqs) {
      hdrs.push_back(r->hdr);
      files.push_back(s.d_arena + r->dev_off);
      outs.push_back(s.d_arena + r->dev_off + r->in_bytes);
    }
    rc = gst_batch_bind(slot.batch, hdrs.data(), static_cast<uint32_t>(hdrs.size()), files.data(),
                        outs.data(), st);
    if (!s.cfg.direct_launch) {
      if (rc == GST_OK && !slot.graph_ready) {
        rc = gst_batch_capture_graph(slot.batch, st);
        slot.graph_ready = rc == GST_OK;
      }
      if (rc == GST_OK) rc = gst_batch_launch_graph(slot.batch, st);
    } else if (rc == GST_OK) {
      rc = gst_batch_decode(slot.batch, st);
    }
    if (rc == GST_OK) rc = gst_batch_error_bits_async(slot.batch, st, slot.err_host);
    if (rc != GST_OK) return rc;
  }
  if ((e = hipEventRecord(slot.ev_dev, st)) != hipSuccess) return HipErr(e);
  if (s.cfg.copy_to_host) {
    for (auto &r : slot.reqs) {
      e = hipMemcpyAsync(s.h_staging + r->host_off + AlignUp(r->len), s.d_arena + r->dev_off + r->in_bytes,
                         r->out_bytes, hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) return HipErr(e);
    }
    if ((e = hipEventRecord(slot.ev_host, st)) != hipSuccess) return HipErr(e);
  }
  slot.state = kDevPending;
  return GST_OK;
}

// Advances a slot whose events completed; returns true if it changed state.
bool PollSlot(gst_server &s, Slot &slot) {
  if (slot.state == kIdle) return false;
  const auto now = Clock::now();
  if (slot.state == kDevPending) {
    if (!s.cfg.dry_run && hipEventQuery(slot.ev_dev) != hipSuccess) return false;
    for (auto &r : slot.reqs) r->t_dev = now;
    if (!s.cfg.dry_run) {
      float t0 = 0, t1 = 0;
      if (hipEventElapsedTime(&t0, s.ev_ref, slot.ev_start) == hipSuccess &&
          hipEventElapsedTime(&t1, s.ev_ref, slot.ev_dev) == hipSuccess) {
        std::lock_guard<std::mutex> lk(s.mu);
        s.busy.emplace_back(t0, t1);
      }
      const uint32_t bits = s.cfg.plain ? 0u : *slot.err_host;
      if (bits)
        for (auto &r : slot.reqs) r->status = GST_ERR_CORRUPT_STREAM;
    }
    slot.state = s.cfg.copy_to_host && !s.cfg.dry_run ? kHostPending : kIdle;
    if (slot.state == kHostPending) return true;
  } else if (slot.state == kHostPending) {
    if (hipEventQuery(slot.ev_host) != hipSuccess) return false;
    slot.state = kIdle;
  }
  for (auto &r : slot.reqs) r->t_host = now;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    for (auto &r : slot.reqs) {
      s.host_alloc.Free(r->host_off, r->host_bytes);
      PublishLocked(s, r);
    }
  }
  slot.reqs.clear();
  s.cv_space.notify_all();
  s.cv_done.notify_all();
  return true;
}

void SchedulerLoop(gst_server *sp) {
  gst_server &s = *sp;
  if (!s.cfg.dry_run) (void)hipSetDevice(0);
  const uint64_t max_blocks = s.cfg.max_batch_blocks;
  for (;;) {
    bool progress = false;
    for (auto &slot : s.slots) progress |= PollSlot(s, slot);
    // Admission (FIFO), then batch formation for every free slot.
    std::vector<std::vector<std::shared_ptr<Req>>> batches;
    std::vector<Slot *> targets;
    {
      std::lock_guard<std::mutex> lk(s.mu);
      while (!s.arrivals.empty()) {
        auto &r = s.arrivals.front();
        uint64_t off = 0;
        if (!s.dev_alloc.Alloc(r->dev_bytes, &off)) {
          s.st.admission_waits++;
          break;
        }
        r->dev_off = off;
        r->admitted = true;
        s.admitted.push_back(r);
        s.arrivals.pop_front();
        s.st.peak_arena_bytes = std::max<uint64_t>(s.st.peak_arena_bytes, s.dev_alloc.used());
      }
      for (auto &slot : s.slots) {
        if (slot.state != kIdle || s.admitted.empty()) continue;
        std::vector<std::shared_ptr<Req>> b;
        uint64_t blocks = 0;
        const uint32_t max_images = s.cfg.plain ? 1u : s.cfg.max_batch_images;
        while (!s.admitted.empty() && b.size() < max_images) {
          auto &r = s.admitted.front();
          const uint64_t nv = static_cast<uint64_t>(r->hdr.width) * r->hdr.height / 16;
          if (!b.empty() && blocks + nv > max_blocks) break;
          blocks += nv;
          b.push_back(r);
          s.admitted.pop_front();
        }
        batches.push_back(std::move(b));
        targets.push_back(&slot);
      }
    }
    for (size_t i = 0; i < batches.size(); ++i) {
      Slot &slot = *targets[i];
      const int rc = LaunchBatch(s, slot, std::move(batches[i]));
      if (rc != GST_OK) {
        // Fail the batch's requests; the slot stays usable.
        std::lock_guard<std::mutex> lk(s.mu);
        for (auto &r : slot.reqs) {
          r->status = rc;
          r->t_launch = r->t_dev = r->t_host = Clock::now();
          s.host_alloc.Free(r->host_off, r->host_bytes);
          PublishLocked(s, r);
        }
        slot.reqs.clear();
        slot.state = kIdle;
        s.cv_space.notify_all();
        s.cv_done.notify_all();
      }
      progress = true;
    }
    bool busy = false;
    for (auto &slot : s.slots) busy |= slot.state != kIdle;
    if (!progress) {
      std::unique_lock<std::mutex> lk(s.mu);
      if (s.stop && s.arrivals.empty() && s.admitted.empty() && !busy) break;
      if (!busy && s.arrivals.empty() && s.admitted.empty()) {
        s.cv_work.wait_for(lk, std::chrono::milliseconds(2));
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

}  // namespace

extern "C" {

int gst_server_create(const gst_server_config *cfg, gst_server **out) {
  if (!cfg || !out || cfg->hbm_budget_bytes == 0) return GST_ERR_INVALID_ARGUMENT;
  std::unique_ptr<gst_server> s(new gst_server());
  s->cfg = *cfg;
  if (!s->cfg.slots) s->cfg.slots = s->cfg.plain ? 1 : 3;
  if (!s->cfg.max_batch_images) s->cfg.max_batch_images = 256;
  if (!s->cfg.max_batch_blocks) s->cfg.max_batch_blocks = 16ull << 20;
  if (!s->cfg.host_staging_bytes) s->cfg.host_staging_bytes = 1ull << 30;
  s->dev_alloc.Reset(s->cfg.hbm_budget_bytes);
  s->host_alloc.Reset(s->cfg.host_staging_bytes);
  s->slots.resize(s->cfg.slots);
  if (!s->cfg.dry_run) {
    int rc = gst_initialize_decoder(0);
    if (rc != GST_OK) return rc;
    hipError_t e = hipMalloc(&s->d_arena, s->cfg.hbm_budget_bytes);
    if (e == hipSuccess) e = hipHostMalloc(&s->h_staging, s->cfg.host_staging_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreate(&s->ev_ref);
    if (e == hipSuccess) e = hipEventRecord(s->ev_ref, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(s->ev_ref);
    for (auto &slot : s->slots) {
      if (e != hipSuccess) break;
      e = hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreate(&slot.ev_start);
      if (e == hipSuccess) e = hipEventCreate(&slot.ev_dev);
      if (e == hipSuccess) e = hipEventCreate(&slot.ev_host);
      if (e == hipSuccess) e = hipHostMalloc(&slot.err_host, sizeof(uint32_t), hipHostMallocDefault);
      if (e == hipSuccess) *slot.err_host = 0;
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
  auto r = std::make_shared<Req>();
  r->t_arrival = Clock::now();
  int rc = gst_parse_header(file, len, &r->hdr);
  if (rc != GST_OK) return rc;
  r->tag = tag;
  r->len = len;
  r->dev_bytes = DevBytes(*s, r->hdr, len, &r->in_bytes, &r->out_bytes);
  r->host_bytes = AlignUp(len) + (s->cfg.copy_to_host ? AlignUp(r->out_bytes) : 0);
  const uint64_t nv = static_cast<uint64_t>(r->hdr.width) * r->hdr.height / 16;
  if (r->dev_bytes > s->cfg.hbm_budget_bytes || r->host_bytes > s->cfg.host_staging_bytes ||
      nv > s->cfg.max_batch_blocks)
    return GST_ERR_OUT_OF_MEMORY;
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->host_alloc.Alloc(r->host_bytes, &r->host_off)) {
    if (s->stop) return GST_ERR_BUSY;
    s->cv_space.wait(lk);
  }
  if (s->h_staging) memcpy(s->h_staging + r->host_off, file, len);
  r->ticket = s->next_ticket++;
  if (ticket) *ticket = r->ticket;
  s->arrivals.push_back(r);
  s->in_flight++;
  s->st.submitted++;
  lk.unlock();
  s->cv_work.notify_one();
  return GST_OK;
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
  s->dev_alloc.Free(it->second->dev_off, it->second->dev_bytes);
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
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_space.notify_all();
    s->sched.join();
  }
  for (auto &slot : s->slots) {
    if (slot.stream) (void)hipStreamSynchronize(slot.stream);
    if (slot.batch) gst_batch_destroy(slot.batch);
    if (slot.ev_start) (void)hipEventDestroy(slot.ev_start);
    if (slot.ev_dev) (void)hipEventDestroy(slot.ev_dev);
    if (slot.ev_host) (void)hipEventDestroy(slot.ev_host);
    if (slot.err_host) (void)hipHostFree(slot.err_host);
    if (slot.stream) (void)hipStreamDestroy(slot.stream);
  }
  if (s->ev_ref) (void)hipEventDestroy(s->ev_ref);
  if (s->d_arena) (void)hipFree(s->d_arena);
  if (s->h_staging) (void)hipHostFree(s->h_staging);
  delete s;
  return GST_OK;
}

int gst_server_run_load(gst_server *s, const uint8_t *const *pool, const size_t *pool_lens,
                        const double *pool_weights, uint32_t pool_size, double qps, double duration_s,
                        uint32_t closed_loop, uint64_t seed, double *lat_dev, double *lat_host,
                        uint64_t lat_cap, uint64_t *n_lat, double *wall_s) {
  // Requires auto_release: a submit blocked on staging space must not wait on device
  // memory that only this thread's releases could free.
  if (!s || !pool || !pool_lens || pool_size == 0 || duration_s <= 0 || (!closed_loop && qps <= 0) ||
      !s->cfg.auto_release)
    return GST_ERR_INVALID_ARGUMENT;
  // SplitMix64 draws: file choice and exponential inter-arrival gaps.
  uint64_t state = seed;
  auto next_u64 = [&state]() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto uniform = [&]() { return (static_cast<double>(next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  std::vector<double> cdf(pool_size);
  double acc = 0;
  for (uint32_t i = 0; i < pool_size; ++i) {
    acc += pool_weights ? pool_weights[i] : 1.0;
    cdf[i] = acc;
  }
  auto pick = [&]() {
    const double u = uniform() * acc;
    return static_cast<uint32_t>(std::min<size_t>(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin(),
                                                  pool_size - 1));
  };
  uint64_t count = 0;
  std::vector<gst_completion> buf(256);
  auto collect = [&]() {
    uint32_t n = 0;
    for (;;) {
      gst_server_poll(s, buf.data(), static_cast<uint32_t>(buf.size()), &n);
      for (uint32_t i = 0; i < n; ++i) {
        if (count < lat_cap) {
          if (lat_dev) lat_dev[count] = buf[i].latency_device_us;
          if (lat_host) lat_host[count] = buf[i].latency_host_us;
        }
        count++;
      }
      if (n < buf.size()) return n;
    }
  };
  const auto t0 = Clock::now();
  const auto t_end = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(duration_s));
  int rc = GST_OK;
  if (!closed_loop) {
    auto next = t0;
    while (next < t_end && rc == GST_OK) {
      while (Clock::now() < next) {
        collect();
        if (next - Clock::now() > std::chrono::microseconds(300))
          std::this_thread::sleep_for(std::chrono::microseconds(100));
      }
      const uint32_t i = pick();
      rc = gst_server_submit(s, pool[i], pool_lens[i], count, nullptr);
      const double gap = -std::log(uniform()) / qps;
      next += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(gap));
    }
  } else {
    uint64_t submitted = 0;
    for (uint32_t k = 0; k < closed_loop && rc == GST_OK; ++k, ++submitted) {
      const uint32_t i = pick();
      rc = gst_server_submit(s, pool[i], pool_lens[i], submitted, nullptr);
    }
    while (Clock::now() < t_end && rc == GST_OK) {
      const uint64_t before = count;
      collect();
      for (uint64_t k = before; k < count && rc == GST_OK; ++k, ++submitted) {
        const uint32_t i = pick();
        rc = gst_server_submit(s, pool[i], pool_lens[i], submitted, nullptr);
      }
      if (count == before) std::this_thread::yield();
    }
  }
  gst_server_drain(s);
  collect();
  if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
  if (n_lat) *n_lat = count;
  return rc;
}

}  // extern "C"
