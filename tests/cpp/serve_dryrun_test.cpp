// serve_dryrun_test.cpp -- the server's scheduler, admission, release and failure paths
// driven from several host threads at once, in dry-run mode (no GPU work).  Built with a
// host sanitizer (ThreadSanitizer or AddressSanitizer, on the host code only) by
// gst_amd/build.py; tests/test_sanitizers.py runs it.  Exit status 0 when every check
// holds; the sanitizer reports races and memory errors itself (non-zero exit).
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "gst_decoder.h"
#include "gst_serve.h"

namespace {

std::atomic<int> g_failures{0};

#define CHECK(cond)                                                  \
  do {                                                               \
    if (!(cond)) {                                                   \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      ++g_failures;                                                  \
    }                                                                \
  } while (0)

// A file a dry run accepts: a valid header and zero-filled frequencies and streams of
// the smallest legal sizes (a dry run reads the header only).
std::vector<uint8_t> FakeFile(uint32_t w, uint32_t h) {
  const uint32_t nv = w * h / 16, per_group = 4 + 4 * 32;
  gst_header hdr{};
  hdr.width = w;
  hdr.height = h;
  hdr.palette_bytes = 8192;
  hdr.y_cmp_sz = 2 * nv / 8192 * per_group;
  hdr.chroma_cmp_sz = 4 * nv / 8192 * per_group;
  hdr.palette_sz = per_group;
  hdr.indices_sz = nv / 8192 * per_group;
  std::vector<uint8_t> f(sizeof(hdr) + 2048 + hdr.y_cmp_sz + hdr.chroma_cmp_sz + hdr.palette_sz + hdr.indices_sz, 0);
  memcpy(f.data(), &hdr, sizeof(hdr));
  return f;
}

uint64_t DevBytes(const std::vector<uint8_t> &f) {
  gst_header h;
  memcpy(&h, f.data(), sizeof(h));
  auto a = [](uint64_t v) { return (v + 255) / 256 * 256; };
  return a(f.size()) + a(static_cast<uint64_t>(h.width) * h.height / 2);
}

// Several submitting threads against one server; every ticket completes exactly once.
void ManySubmitters(uint32_t fail_every) {
  const std::vector<std::vector<uint8_t>> files = {FakeFile(512, 512), FakeFile(1024, 1024), FakeFile(512, 256),
                                                   FakeFile(2048, 2048)};
  gst_server_config cfg{};
  cfg.hbm_budget_bytes = 64ull << 20;
  cfg.dry_run = 1;
  cfg.auto_release = 1;
  cfg.slots = 3;
  cfg.max_batch_images = 8;
  cfg.fail_launch_every = fail_every;
  gst_server *s = nullptr;
  CHECK(gst_server_create(&cfg, &s) == GST_OK);
  constexpr int kThreads = 6, kEach = 150;
  std::mutex mu;
  std::vector<uint64_t> tickets;
  std::vector<std::thread> th;
  for (int t = 0; t < kThreads; ++t) {
    th.emplace_back([&, t] {
      for (int i = 0; i < kEach; ++i) {
        const auto &f = files[(t + i) % files.size()];
        uint64_t ticket = 0;
        const int rc = gst_server_submit(s, f.data(), f.size(), static_cast<uint64_t>(t) * 1000 + i, &ticket);
        CHECK(rc == GST_OK);
        std::lock_guard<std::mutex> lk(mu);
        tickets.push_back(ticket);
      }
    });
  }
  // A poller thread collects completions while the submitters run.
  std::atomic<bool> done{false};
  std::map<uint64_t, int> seen;
  uint64_t failed = 0;
  std::thread poller([&] {
    std::vector<gst_completion> buf(64);
    while (!done.load()) {
      uint32_t n = 0;
      CHECK(gst_server_poll(s, buf.data(), static_cast<uint32_t>(buf.size()), &n) == GST_OK);
      for (uint32_t i = 0; i < n; ++i) {
        seen[buf[i].ticket]++;
        if (buf[i].status != GST_OK) ++failed;
      }
      if (!n) std::this_thread::yield();
    }
  });
  for (auto &x : th) x.join();
  CHECK(gst_server_drain(s) == GST_OK);
  done = true;
  poller.join();
  std::vector<gst_completion> buf(1024);
  uint32_t n = 0;
  CHECK(gst_server_poll(s, buf.data(), static_cast<uint32_t>(buf.size()), &n) == GST_OK);
  for (uint32_t i = 0; i < n; ++i) {
    seen[buf[i].ticket]++;
    if (buf[i].status != GST_OK) ++failed;
  }
  CHECK(tickets.size() == static_cast<size_t>(kThreads) * kEach);
  CHECK(seen.size() == tickets.size());
  for (uint64_t t : tickets) CHECK(seen[t] == 1);
  gst_server_stats st{};
  CHECK(gst_server_stats_get(s, &st) == GST_OK);
  CHECK(st.completed == tickets.size() && st.submitted == tickets.size());
  CHECK(st.failed == failed);
  if (fail_every) CHECK(failed > 0);
  else CHECK(failed == 0);
  CHECK(gst_server_destroy(s) == GST_OK);
}

// A budget that holds a few requests, completions held until another thread releases
// them: admission waits on releases made concurrently with submits and polls.
void ReleasesFromAnotherThread() {
  const std::vector<uint8_t> f = FakeFile(1024, 1024);
  gst_server_config cfg{};
  cfg.hbm_budget_bytes = 3 * DevBytes(f);
  cfg.dry_run = 1;
  cfg.auto_release = 0;
  gst_server *s = nullptr;
  CHECK(gst_server_create(&cfg, &s) == GST_OK);
  constexpr int kTotal = 300;
  std::atomic<int> released{0};
  std::thread releaser([&] {
    std::vector<gst_completion> buf(16);
    while (released.load() < kTotal) {
      uint32_t n = 0;
      CHECK(gst_server_poll(s, buf.data(), static_cast<uint32_t>(buf.size()), &n) == GST_OK);
      for (uint32_t i = 0; i < n; ++i) {
        CHECK(buf[i].status == GST_OK);
        CHECK(gst_server_release(s, buf[i].ticket) == GST_OK);
        released++;
      }
      if (!n) std::this_thread::yield();
    }
  });
  std::vector<std::thread> th;
  for (int t = 0; t < 3; ++t)
    th.emplace_back([&] {
      for (int i = 0; i < kTotal / 3; ++i) CHECK(gst_server_submit(s, f.data(), f.size(), i, nullptr) == GST_OK);
    });
  for (auto &x : th) x.join();
  releaser.join();
  gst_server_stats st{};
  CHECK(gst_server_stats_get(s, &st) == GST_OK);
  CHECK(st.completed == kTotal && st.peak_arena_bytes <= cfg.hbm_budget_bytes);
  CHECK(st.admission_waits > 0);
  CHECK(gst_server_release(s, 123456789) == GST_ERR_INVALID_ARGUMENT);
  CHECK(gst_server_destroy(s) == GST_OK);
}

// Destroying a server while requests wait for memory that only a release could free:
// they complete with GST_ERR_BUSY instead of blocking the destroy.
void DestroyWithWaitingRequests() {
  const std::vector<uint8_t> f = FakeFile(512, 512);
  gst_server_config cfg{};
  cfg.hbm_budget_bytes = DevBytes(f);
  cfg.dry_run = 1;
  cfg.auto_release = 0;
  gst_server *s = nullptr;
  CHECK(gst_server_create(&cfg, &s) == GST_OK);
  for (int i = 0; i < 4; ++i) CHECK(gst_server_submit(s, f.data(), f.size(), i, nullptr) == GST_OK);
  CHECK(gst_server_destroy(s) == GST_OK);
}

// Pre-staged requests (device pointers never dereferenced in a dry run) and the load
// generator's closed loop.
void PrestagedAndLoad() {
  const std::vector<uint8_t> a = FakeFile(512, 512), b = FakeFile(2048, 2048);
  gst_server_config cfg{};
  cfg.hbm_budget_bytes = 256ull << 20;
  cfg.dry_run = 1;
  cfg.auto_release = 1;
  gst_server *s = nullptr;
  CHECK(gst_server_create(&cfg, &s) == GST_OK);
  const uint8_t *pool[2] = {a.data(), b.data()};
  const size_t lens[2] = {a.size(), b.size()};
  const void *dev[2] = {reinterpret_cast<const void *>(0x100000), reinterpret_cast<const void *>(0x200000)};
  gst_load_spec spec{};
  spec.pool = pool;
  spec.pool_lens = lens;
  spec.pool_size = 2;
  spec.pool_device = dev;
  spec.duration_s = 0.2;
  spec.closed_loop = 16;
  spec.seed = 3;
  std::vector<double> lat(1 << 16);
  gst_load_result res{};
  CHECK(gst_server_run_load(s, &spec, lat.data(), nullptr, lat.size(), &res) == GST_OK);
  CHECK(res.requests >= 16 && res.failed == 0);
  gst_server_stats st{};
  CHECK(gst_server_stats_get(s, &st) == GST_OK);
  CHECK(st.prestaged == st.submitted && st.h2d_bytes == 0);
  CHECK(gst_server_destroy(s) == GST_OK);
}

// The load generator's closed loop with several submitting threads copying inputs into
// the staging in parallel (a dry run's staging is ordinary memory, so the copies run).
void LoadWithSubmitThreads() {
  const std::vector<uint8_t> a = FakeFile(512, 512), b = FakeFile(1024, 1024);
  gst_server_config cfg{};
  cfg.hbm_budget_bytes = 256ull << 20;
  cfg.host_staging_bytes = 8ull << 20;  // small: submitters also wait for space
  cfg.dry_run = 1;
  cfg.auto_release = 1;
  gst_server *s = nullptr;
  CHECK(gst_server_create(&cfg, &s) == GST_OK);
  const uint8_t *pool[2] = {a.data(), b.data()};
  const size_t lens[2] = {a.size(), b.size()};
  gst_load_spec spec{};
  spec.pool = pool;
  spec.pool_lens = lens;
  spec.pool_size = 2;
  spec.duration_s = 0.2;
  spec.closed_loop = 64;
  spec.seed = 5;
  spec.submit_threads = 4;
  gst_load_result res{};
  CHECK(gst_server_run_load(s, &spec, nullptr, nullptr, 0, &res) == GST_OK);
  gst_server_stats st{};
  CHECK(gst_server_stats_get(s, &st) == GST_OK);
  CHECK(res.requests >= 64 && res.failed == 0);
  CHECK(st.submitted == st.completed && st.completed == res.requests);
  CHECK(gst_server_destroy(s) == GST_OK);
}

}  // namespace

int main() {
  fprintf(stderr, "many submitters\n");
  ManySubmitters(0);
  fprintf(stderr, "many submitters, failing launches\n");
  ManySubmitters(4);
  fprintf(stderr, "releases from another thread\n");
  ReleasesFromAnotherThread();
  fprintf(stderr, "destroy with waiting requests\n");
  DestroyWithWaitingRequests();
  fprintf(stderr, "pre-staged requests and the load generator\n");
  PrestagedAndLoad();
  fprintf(stderr, "load generator with several submitting threads\n");
  LoadWithSubmitThreads();
  if (g_failures.load()) {
    fprintf(stderr, "%d checks failed\n", g_failures.load());
    return 1;
  }
  printf("ok\n");
  return 0;
}
