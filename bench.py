#!/usr/bin/env python3
"""GST decode benchmark on one MI355X.

Headline workload (printed as one JSON line): 128 seeded textures of uneven size drawn
uniformly from {512x512, 1024x1024, 512x1024, 2048x2048}, decoded as one batch (two kernels
after autotuning: rANS with tables built in LDS, then wavelet + index scan + DXT1 assembly).
A "step" is one decode of the whole batch with inputs and outputs resident in HBM.  Steps
are timed three ways -- hipGraph replays, stream launches, and two plans in flight on two
streams (step k on plan k % 2) -- and the fastest is the step time; all three are reported.
Beside it the same corpus is decoded
image by image through the plain LoadCompressedDXT path, single 512x512 and 4096x4096
textures are decoded through the plain path, and the CPU restatement of the decoder
(oracle/) is timed on the host cores as the CPU baseline.  Every texture is encoded as the
reference encoder gentenc does, byte for byte (stb_dxt blocks, gentenc's palette reuse and
streams; gste_compress_dxt), so the files are the ones the reference would decode.  Every
GPU result is checked bit-exact against the encoder's blocks (which the oracle reproduces)
before it is timed.

Roofline: the dominant kernel's algorithmic bytes per launch (per-stage read+write of
the decode dataflow) over its average launch time, measured with hipEvents on the
decode stream, against 8 TB/s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "decoded Mtexel/s & HBM roofline %, batched vs per-image; p50/p99 latency @QPS"
HBM_PEAK_GBS = 8000.0
SIZES_C3 = [(512, 512), (1024, 1024), (512, 1024), (2048, 2048)]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def corpus_c3(n=128, size_seed=42):
    rng = np.random.default_rng(size_seed)
    picks = rng.integers(0, len(SIZES_C3), n)
    return [(1000 + i, SIZES_C3[p][0], SIZES_C3[p][1]) for i, p in enumerate(picks)]


def stage_bytes(hdrs, mode="tile_fused", tables_in_lds=True):
    """Algorithmic read+write bytes per stage for one decode of `hdrs`; 0 for a stage
    with no launch of its own.  `mode` is how the last stages run: "tile_fused"
    (wavelet, index scan and DXT1 assembly in one kernel: the int8 planes stay in LDS),
    "fused" (index scan inside the assembly: it reads the nv diff bytes instead of nv
    int32 indices) or "separate".  With `tables_in_lds` the rANS kernel builds its
    tables from the four 512-byte frequency tables of each image (2048 bytes read, no
    table traffic); otherwise build_table writes 4 x 8 KiB tables that the rANS stage
    reads.  The rANS stage also writes the index stream's partial sums (nv/64 + nv/8
    bytes), which the fused assemblies read."""
    nv = sum(h.num_blocks for h in hdrs)
    P = sum(h.palette_bytes for h in hdrs)
    S = sum(h.stream_bytes for h in hdrs)
    n = len(hdrs)
    partial = nv // 64 + nv // 8
    scan_folded = mode in ("tile_fused", "fused")
    table_read = 2048 * n if tables_in_lds else 32768 * n
    return {
        "build_table": 0 if tables_in_lds else 2048 * n + 32768 * n,
        "ans_decode": (S + table_read) + (7 * nv + P) + partial,
        "decode_indices": 0 if scan_folded else nv + 4 * nv,
        "inv_wavelet": 0 if mode == "tile_fused" else 6 * nv + 6 * nv,
        "assemble": ((6 * nv + nv + partial + P) if scan_folded else (6 * nv + 4 * nv + P)) + 8 * nv,
    }


def measured_traffic(stage, n_images, requested):
    """HBM bytes per launch of `stage` (read + write) from the committed rocprofv3
    summary of this workload (L2-to-memory requests by size: 128 x RDREQ_128B + 64 x
    RDREQ_64B + 32 x RDREQ_32B read, 64 x WRREQ_64B + 32 x the other write requests);
    None for another corpus size or a stage the summary lacks."""
    path = os.path.join(ROOT, "profiles", "c3_batch_traffic.json")
    if requested != 128 or n_images != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        per = json.load(f)["bytes_per_launch"].get(stage)
    return None if per is None else per["read"] + per["write"]


def pipeline_bytes(hdrs, mode="tile_fused", tables_in_lds=True):
    """B_pipe: the algorithmic bytes of the kernels the decode launches in `mode`;
    B_min: inputs read once (streams, frequencies) and the DXT1 output written once."""
    b_pipe = sum(stage_bytes(hdrs, mode, tables_in_lds).values())
    b_min = sum((h.stream_bytes + 2048) + 8 * h.num_blocks for h in hdrs)
    return b_pipe, b_min


class Gpu:
    def __init__(self):
        import torch
        self.torch = torch
        self.dev = torch.device("cuda:0")
        # A created stream: graph capture is not possible on the null stream.
        self.main = torch.cuda.Stream(device=self.dev)
        torch.cuda.set_stream(self.main)

    @property
    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def upload_many(self, blobs, align=256):
        offs, total = [], 0
        for b in blobs:
            offs.append(total)
            total += (len(b) + align - 1) // align * align
        host = np.zeros(total, np.uint8)
        for o, b in zip(offs, blobs):
            host[o:o + len(b)] = np.frombuffer(b, np.uint8)
        dev = self.torch.from_numpy(host).to(self.dev)
        return dev, [dev.data_ptr() + o for o in offs]

    def empty(self, n):
        return self.torch.empty(int(n), dtype=self.torch.uint8, device=self.dev)

    def timed_two(self, s0, s1, fn, reps):
        """Like timed, for work on two streams: s1 starts after the start event on s0,
        and the end event on s0 waits for s1."""
        torch = self.torch
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(s0)
        s1.wait_event(start)
        fn()
        s0.wait_stream(s1)
        end.record(s0)
        torch.cuda.synchronize()
        return start.elapsed_time(end) / reps

    def timed(self, fn, reps):
        torch = self.torch
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        for _ in range(reps):
            fn()
        end.record()
        torch.cuda.synchronize()
        return start.elapsed_time(end) / reps


def check_outputs(gpu, d_out, offsets, dxts, what):
    """Every image's bytes of `d_out` (poisoned before the decode) against its blocks."""
    gpu.torch.cuda.synchronize()
    host = d_out.cpu().numpy()
    for i, (o, d) in enumerate(zip(offsets, dxts)):
        if host[o:o + len(d)].tobytes() != d:
            raise SystemExit("%s: image %d differs from the expected DXT1 blocks" % (what, i))


LAUNCH_MODES = ("graph", "stream", "two_plans", "two_plan_graphs")


def run_batched(gpu, files, dxts, steps, warmup, tune=True, modes=LAUNCH_MODES):
    """The batch decoded K steps after W warm-up steps in each launch mode of `modes`:
    graph (one hipGraph replay a step), stream (the stage kernels launched on the
    stream), two_plans (a second plan bound to the same inputs with its own scratch and
    outputs, on a second stream; step k runs on plan k % 2, so one plan's rANS stage
    fills the other's assembly tail, as the server's slots run batches), two_plan_graphs
    (the same with each plan's captured graph replayed on its stream).  The step time is
    the fastest mode.
    Outputs are poisoned and checked bit-exact after every mode."""
    import gst_amd
    hdrs = [gst_amd.ParseHeader(f) for f in files]
    d_files, file_ptrs = gpu.upload_many(files)
    out_sizes = [h.dxt_bytes for h in hdrs]
    out_offs = list(np.cumsum([0] + out_sizes[:-1]))
    d_out = gpu.empty(sum(out_sizes))
    nv, pal = sum(h.num_blocks for h in hdrs), sum(h.palette_bytes for h in hdrs)
    b = gst_amd.Batch(len(hdrs), nv, pal)
    b.bind(hdrs, file_ptrs, [d_out.data_ptr() + int(o) for o in out_offs], gpu.stream)
    tuned = None
    if tune:
        t0 = time.perf_counter()
        tuned = b.autotune(gpu.stream, reps=5)
        tuned["seconds"] = round(time.perf_counter() - t0, 2)
        log("autotune: %s" % tuned)
    d_out.fill_(0xA5)
    b.decode(gpu.stream)
    check_outputs(gpu, d_out, out_offs, dxts, "batched eager")
    if b.error_bits(gpu.stream):
        raise SystemExit("batched decode raised device error bits")
    launch = {}
    if "graph" in modes:
        b.capture_graph(gpu.stream)
        for _ in range(warmup):
            b.launch_graph(gpu.stream)
        launch["graph"] = gpu.timed(lambda: b.launch_graph(gpu.stream), steps)
        d_out.fill_(0xA5)
        b.launch_graph(gpu.stream)
        check_outputs(gpu, d_out, out_offs, dxts, "batched graph")
    if "stream" in modes:
        for _ in range(warmup):
            b.decode(gpu.stream)
        launch["stream"] = gpu.timed(lambda: b.decode(gpu.stream), steps)
        d_out.fill_(0xA5)
        b.decode(gpu.stream)
        check_outputs(gpu, d_out, out_offs, dxts, "batched stream launches")
    if "two_plans" in modes or "two_plan_graphs" in modes:
        s0 = gpu.torch.cuda.current_stream()
        s1 = gpu.torch.cuda.Stream(device=gpu.dev)
        d_out2 = gpu.empty(sum(out_sizes))
        b2 = gst_amd.Batch(len(hdrs), nv, pal)
        b2.bind(hdrs, file_ptrs, [d_out2.data_ptr() + int(o) for o in out_offs], s1.cuda_stream)
        if tune:
            b2.autotune(s1.cuda_stream, reps=5)
        plans = ((b, s0), (b2, s1))

        def two_plans(k_steps, graphs=False):
            for k in range(k_steps):
                plan, st = plans[k % 2]
                if graphs:
                    plan.launch_graph(st.cuda_stream)
                else:
                    plan.decode(st.cuda_stream)

        def check_both(what):
            gpu.torch.cuda.synchronize()
            check_outputs(gpu, d_out, out_offs, dxts, what + ", first")
            check_outputs(gpu, d_out2, out_offs, dxts, what + ", second")

        if "two_plans" in modes:
            two_plans(2 * warmup)
            launch["two_plans"] = gpu.timed_two(s0, s1, lambda: two_plans(steps), steps)
            d_out.fill_(0xA5)
            d_out2.fill_(0xA5)
            two_plans(2)
            check_both("two plans")
        if "two_plan_graphs" in modes:
            for plan, st in plans:
                plan.capture_graph(st.cuda_stream)
            two_plans(2 * warmup, graphs=True)
            launch["two_plan_graphs"] = gpu.timed_two(s0, s1, lambda: two_plans(steps, graphs=True), steps)
            d_out.fill_(0xA5)
            d_out2.fill_(0xA5)
            two_plans(2, graphs=True)
            check_both("two plans, graph replays")
        b2.close()
        del d_out2
    ms = min(launch.values())
    # Per-stage kernel time in the pipeline's order and cache state: 200 decodes back to
    # back with a timing-only hipEvent (no cache write-back) recorded on the decode stream
    # after every launched stage, after `warmup` untimed ones; mean per stage.
    d_out.fill_(0xA5)
    for _ in range(warmup):
        b.decode(gpu.stream)
    stage = b.decode_timed(gpu.stream, reps=200)
    check_outputs(gpu, d_out, out_offs, dxts, "batched timed")
    dev_bytes = b.device_bytes
    b.close()
    del d_files
    return hdrs, ms, stage, dev_bytes, tuned, launch


def run_plain(gpu, files, dxts, passes, warmup):
    """Image by image through LoadCompressedDXT (the reference's per-image drop-in)."""
    import gst_amd
    stagings, hdrs = [], []
    for f in files:
        s, (h,) = gst_amd.PackStaging([f])
        stagings.append(s)
        hdrs.append(h)
    d_in, in_ptrs = gpu.upload_many(stagings)
    out_sizes = [h.dxt_bytes for h in hdrs]
    out_offs = list(np.cumsum([0] + out_sizes[:-1]))
    d_out = gpu.empty(sum(out_sizes))
    scratch = sum(gst_amd.RequiredScratchMem(h) for h in hdrs)

    def one_pass():
        for h, p, o in zip(hdrs, in_ptrs, out_offs):
            gst_amd.LoadCompressedDXT(h, gpu.stream, p, d_out.data_ptr() + int(o))

    times = []
    for i in range(warmup + passes):
        gst_amd.PreallocateDecompressor(scratch)
        t = gpu.timed(one_pass, 1)
        gst_amd.FreeDecompressor()
        if i == 0:
            check_outputs(gpu, d_out, out_offs, dxts, "plain")
        if i >= warmup:
            times.append(t)
    return hdrs, float(np.median(times))


def cgroup_cpu_quota():
    """CPUs granted by the cgroup's CPU quota (cgroup v2 cpu.max), or None."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        return None if quota == "max" else max(1, int(int(quota) / int(period)))
    except (OSError, ValueError):
        return None


def host_cores():
    """(os.cpu_count(), the cores this process may use: its CPU affinity, capped by the
    cgroup CPU quota when there is one)."""
    usable = len(os.sched_getaffinity(0))
    quota = cgroup_cpu_quota()
    return os.cpu_count() or 1, min(usable, quota) if quota else usable


def run_cpu_baseline(files, min_seconds=10.0, max_passes=50):
    """The oracle (the C restatement of the five kernels) decoding the whole corpus on
    every core this process may use, until at least `min_seconds` have passed."""
    from oracle import gst_oracle as go
    threads = host_cores()[1]
    texels = 0
    for f in files:
        rc, h = go.parse_header(f)
        texels += h.width * h.height
    rates = []
    t_total = 0.0
    while (t_total < min_seconds or len(rates) < 2) and len(rates) < max_passes:
        t0 = time.perf_counter()
        for f in files:
            rc, _, _ = go.decode_dxt(f, nthreads=threads)
            assert rc == go.OK
        dt = time.perf_counter() - t0
        t_total += dt
        rates.append(texels / dt / 1e6)
    return float(np.median(rates)), threads, len(rates), t_total


def run_cpu_single(f, expected, runs=7):
    """One texture decoded by the oracle on every core: a warm-up, then the median of
    `runs` timed decodes, the output checked against the expected blocks."""
    from oracle import gst_oracle as go
    threads = host_cores()[1]
    rc, out, _ = go.decode_dxt(f, nthreads=threads)
    if rc != go.OK or out != expected:
        raise SystemExit("CPU baseline: the oracle's decode differs from the expected blocks")
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        go.decode_dxt(f, nthreads=threads)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)) * 1e3, threads


SERVED_MIX = [((512, 512), 0.50, 16), ((1024, 1024), 0.30, 12), ((2048, 2048), 0.15, 6),
              ((4096, 4096), 0.05, 2)]


def run_served(gpu, qps_list, duration_s, budget_gib=32, closed_loop=512, tuned=None):
    """Poisson arrivals of the size mix through the continuous-batching server and
    through the plain per-request server, then saturating closed loops.  Latency is
    arrival -> output in HBM (and -> copied back to pinned host memory).  Every
    completion is compared with its expected blocks on the device (the pool's expected
    outputs stay in HBM); a failure or a difference ends the benchmark.  The batched
    server takes the rANS launch shape the batched leg's autotune chose."""
    shape = {}
    if tuned:
        shape = {"groups_per_wave": tuned["ans_groups_per_wave"], "waves_per_group": tuned["ans_waves_per_group"]}
    from gst_amd import encode
    from gst_amd.serve import Server
    specs, weights = [], []
    seed = 5000
    for (w, h), p, count in SERVED_MIX:
        for _ in range(count):
            specs.append((seed, w, h))
            weights.append(p / count)
            seed += 1
    pool, dxts = encode.make_textures(specs, nthreads=host_cores()[1], encoder="gentenc")
    d_pool, pool_dev = gpu.upload_many(pool)
    d_exp, pool_exp = gpu.upload_many(dxts)
    gpu.torch.cuda.synchronize()
    out = {"pool": "%d pre-encoded textures (gentenc bytes), size mix {512^2: 50%%, 1024^2: 30%%, 2048^2: 15%%, "
                   "4096^2: 5%%}" % len(pool),
           "hbm_budget_GiB": budget_gib, "duration_s_per_point": duration_s,
           "checked": "every completion compared with its expected DXT1 blocks on the device, except in "
                      "closed_loop.batched_prestaged_unchecked (the same requests without the comparison)"}

    def summarize(lat_dev, lat_host, res, st, checked=True):
        if res["failed"] or res["mismatches"] or (checked and res["verified"] != res["requests"]):
            raise SystemExit("served: %s" % res)
        busy = st["device_busy_ms"] * 1e-3
        wall = res["wall_s"]
        return {
            "requests": int(res["requests"]),
            "outputs_checked": int(res["verified"]),
            "p50_us": round(float(np.percentile(lat_dev, 50)), 1),
            "p99_us": round(float(np.percentile(lat_dev, 99)), 1),
            "p50_with_d2h_us": round(float(np.percentile(lat_host, 50)), 1),
            "p99_with_d2h_us": round(float(np.percentile(lat_host, 99)), 1),
            "Mtexel_s": round(st["texels"] / wall / 1e6, 1),
            "Mtexel_s_busy": round(st["texels"] / busy / 1e6, 1) if busy > 0 else None,
            "busy_frac": round(busy / wall, 4),
            "B_pipe_roofline_busy": round(st["algorithmic_bytes"] / busy / (HBM_PEAK_GBS * 1e9), 4)
            if busy > 0 else None,
            "B_pipe_roofline_wall": round(st["algorithmic_bytes"] / wall / (HBM_PEAK_GBS * 1e9), 4),
            "mean_batch_images": round(st["mean_batch_images"], 2),
            "max_batch_images": int(st["max_batch_images"]),
            "h2d_GB": round(st["h2d_bytes"] / 1e9, 3),
            "h2d_copies": int(st["h2d_copies"]),
            "scheduler_launch_ms": round(st["launch_host_ms"], 1),
        }

    def one(label, server_kw, load_kw, checked=True):
        with Server(budget_gib << 30, **server_kw) as s:
            s.run_load(pool[:4], qps=200.0, duration_s=0.2, seed=1, pool_expected=pool_exp[:4])  # warm-up
            base = s.stats()
            lat_dev, lat_host, res = s.run_load(pool, weights, pool_expected=pool_exp if checked else None,
                                                duration_s=duration_s, **load_kw)
            st = s.stats()
        for k in ("texels", "algorithmic_bytes", "h2d_bytes", "h2d_copies"):
            st[k] -= base[k]
        st["device_busy_ms"] -= base["device_busy_ms"]
        st["launch_host_ms"] -= base["launch_host_ms"]
        r = summarize(lat_dev, lat_host, res, st, checked)
        log("served %s: %s" % (label, r))
        return r

    for mode in ("batched", "plain"):
        res = {}
        for qps in qps_list:
            kw = dict(plain=(mode == "plain"), copy_to_host=True, **(shape if mode == "batched" else {}))
            res[str(int(qps))] = one("%s @%d req/s" % (mode, qps), kw, dict(qps=qps, seed=int(qps)))
        out[mode] = res
    # Saturating closed loops: host-submitted inputs (four submitting threads copy them
    # into pinned staging, the gather kernel reads them into HBM: the ingest path), the
    # same with a DMA copy per run of adjacent inputs and one submitting thread, the pool
    # pre-staged in HBM (decode and scheduling only), and the plain per-request server.
    out["closed_loop"] = {
        "depth": closed_loop,
        "batched_ingest": one("closed loop, ingest", dict(shape, copy_kernel=True),
                              dict(closed_loop=closed_loop, seed=7, submit_threads=4)),
        "batched_ingest_dma": one("closed loop, ingest by DMA copies", dict(shape),
                                  dict(closed_loop=closed_loop, seed=7)),
        "batched_prestaged": one("closed loop, pre-staged", dict(shape),
                                 dict(closed_loop=closed_loop, seed=7, pool_device=pool_dev)),
        # The same requests with no expected outputs: the server alone, without the
        # on-device comparison (two more 8-byte-a-block reads per output) that every other
        # point runs; its outputs are the ones the point above checked.
        "batched_prestaged_unchecked": one("closed loop, pre-staged, no comparison", dict(shape),
                                           dict(closed_loop=closed_loop, seed=7, pool_device=pool_dev),
                                           checked=False),
        "plain_ingest": one("closed loop, plain", dict(plain=True), dict(closed_loop=64, seed=7)),
    }
    cl = out["closed_loop"]
    cl["speedup_prestaged_vs_plain"] = round(cl["batched_prestaged"]["Mtexel_s"] / cl["plain_ingest"]["Mtexel_s"], 2)
    cl["speedup_ingest_vs_plain"] = round(cl["batched_ingest"]["Mtexel_s"] / cl["plain_ingest"]["Mtexel_s"], 2)
    del d_pool, d_exp
    return out


def run_h2d_rates(gpu, total=1 << 30):
    """GB/s of uploading `total` bytes of pinned host memory to the device in ranges of one
    size: one DMA copy per range (torch's copy of a pinned tensor slice) against the gather
    kernel (gst_upload_gather, 64 ranges a launch).  Best of 3 passes each; the last
    pass's bytes are checked."""
    import gst_amd
    torch = gpu.torch
    host = torch.empty(total, dtype=torch.uint8).pin_memory()
    host[::4096] = torch.arange(total // 4096, dtype=torch.int64).to(torch.uint8)
    dev = gpu.empty(total)
    out = {}
    for chunk in (64 << 10, 256 << 10, 1 << 20, 4 << 20, 16 << 20):
        n = total // chunk
        hp, dp = host.data_ptr(), dev.data_ptr()

        def dma():
            for i in range(n):
                dev[i * chunk:(i + 1) * chunk].copy_(host[i * chunk:(i + 1) * chunk], non_blocking=True)

        def gather():
            gst_amd.UploadGather([hp + i * chunk for i in range(n)], [dp + i * chunk for i in range(n)],
                                 [chunk] * n, gpu.stream)

        row = {}
        for name, fn in (("dma", dma), ("gather", gather)):
            ms = min(gpu.timed(fn, 1) for _ in range(3))
            dev.fill_(0xA5)
            fn()
            torch.cuda.synchronize()
            if not torch.equal(dev[::4096].cpu(), host[::4096]):
                raise SystemExit("h2d %s: bytes differ at chunk %d" % (name, chunk))
            row[name] = round(total / (ms * 1e-3) / 1e9, 1)
        out["%d KiB" % (chunk >> 10)] = row
        log("h2d %d KiB: %s" % (chunk >> 10, row))
    return out


def run_stage_kernels(gpu, reps=20):
    """The reference's five stage launches one at a time (codec/decoder.cpp:165-417,
    gst_stage_* here) on the single-image configurations: C1 (the 512x512 seed-0
    texture) and C2 (the 4096x4096 seed-1 texture, noise 0.3 of range).  Each stage is
    timed with hipEvents on the decode stream (median of `reps`), its read+write bytes
    from the per-stage table, and the final DXT1 blocks are checked bit-exact."""
    import gst_amd
    from gst_amd import encode
    torch = gpu.torch
    out = {}
    for name, (seed, w, h) in (("C1_512x512", (0, 512, 512)), ("C2_4096x4096", (1, 4096, 4096))):
        (f,), (d,) = encode.make_textures([(seed, w, h)], encoder="gentenc")
        hdr = gst_amd.ParseHeader(f)
        nv, P, S = hdr.num_blocks, hdr.palette_bytes, hdr.stream_bytes
        d_file, (base,) = gpu.upload_many([f])
        t_tab, t_err, t_sym, t_gs, t_idx, t_pl, t_out = (gpu.empty(n) for n in (
            4 * 2048 * 4 + 16, 256, 7 * nv + P, 4 * (nv // 8192), 4 * nv, 6 * nv, 8 * nv))
        t_err.zero_()
        st = gpu.stream
        calls = {
            "build_table": (lambda: gst_amd.StageBuildTable(base + 28, t_tab.data_ptr(), t_err.data_ptr(), st),
                            2048 + 32768),
            "ans_decode": (lambda: gst_amd.StageAnsDecode(hdr, base + 28 + 2048, t_tab.data_ptr(), t_sym.data_ptr(),
                                                          t_gs.data_ptr(), t_err.data_ptr(), st),
                           S + 32768 + 7 * nv + P + 4 * (nv // 8192)),
            "decode_indices": (lambda: gst_amd.StageDecodeIndices(hdr, t_sym.data_ptr(), t_gs.data_ptr(),
                                                                  t_idx.data_ptr(), st), nv + 4 * nv),
            "inv_wavelet": (lambda: gst_amd.StageInvWavelet(hdr, t_sym.data_ptr(), t_pl.data_ptr(), st),
                            6 * nv + 6 * nv),
            "assemble": (lambda: gst_amd.StageAssembleDXT(hdr, t_sym.data_ptr(), t_pl.data_ptr(), t_idx.data_ptr(),
                                                          t_out.data_ptr(), t_err.data_ptr(), st),
                         6 * nv + 4 * nv + P + 8 * nv),
        }
        for fn, _ in calls.values():  # one pass in pipeline order leaves every stage's inputs
            fn()
        torch.cuda.synchronize()
        if t_out.cpu().numpy().tobytes() != d or t_err.cpu().numpy()[:4].any():
            raise SystemExit("stage kernels, %s: output differs from the expected DXT1 blocks" % name)
        res = {}
        for k, (fn, nbytes) in calls.items():
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
            fn()
            ev[0].record()
            for i in range(reps):
                fn()
                ev[i + 1].record()
            torch.cuda.synchronize()
            ms = float(np.median([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))
            res[k] = {"us": round(ms * 1e3, 2), "bytes": int(nbytes), "GB_s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                      "hbm_frac": round(nbytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
        out[name] = res
        log("stage kernels %s: %s" % (name, res))
    return out


def run_ans_api(files, reps=3):
    """The stand-alone rANS API (gst_ans_*, the OpenCLDecoder counterpart) over the
    corpus' luma streams: per image, RebuildTable from its Y frequencies, then one Decode
    of all its Y groups (32 interleaved streams each).  End to end through the blocking
    API: host packing, upload, the groups kernel, read-back.  Every stream must decode back
    to the encoder's initial state (the API reports a corrupt stream otherwise)."""
    import gst_amd
    prep = []
    for f in files:
        h = gst_amd.ParseHeader(f)
        freqs = [int(v) for v in np.frombuffer(f[28:28 + 512], "<u2")]
        y = f[28 + 2048:28 + 2048 + h.y_cmp_sz]
        groups = 2 * h.num_blocks // 8192
        ends = np.frombuffer(y[:4 * groups], "<u4")
        blobs, states, lo = [], [], 4 * groups
        for g in range(groups):
            grp = y[lo:int(ends[g])]
            blobs.append(grp[:-128])
            states += [int(v) for v in np.frombuffer(grp[-128:], "<u4")]
            lo = int(ends[g])
        prep.append((freqs, blobs, states))
    dec = gst_amd.AnsDecoder(prep[0][0], 32)
    symbols = sum(len(p[2]) for p in prep) * 256
    best = None
    for _ in range(reps + 1):  # the first pass warms up
        t0 = time.perf_counter()
        for freqs, blobs, states in prep:
            dec.rebuild_table(freqs)
            dec.decode(states, blobs)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    dec.close()
    out = {"images": len(prep), "groups": symbols // 8192, "symbols": symbols, "seconds": round(best, 4),
           "Msym_s": round(symbols / best / 1e6, 1),
           "path": "per image: RebuildTable + one Decode of all luma groups, blocking, host to host"}
    log("ans api: %s" % out)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="timed steps per launch mode")
    ap.add_argument("--warmup", type=int, default=300,
                    help="untimed steps before each launch mode's timed steps: the GPU's clocks take "
                         "about 25 ms of back-to-back decodes to settle after the autotuner's "
                         "stop-and-go, and short timed windows inside that ramp read up to 13%% slow")
    ap.add_argument("--images", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-image comparison")
    ap.add_argument("--no-served", action="store_true", help="skip the served (Poisson arrivals) leg")
    ap.add_argument("--served-seconds", type=float, default=3.0)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotune)")
    ap.add_argument("--no-ans-api", action="store_true", help="skip the stand-alone rANS API leg")
    ap.add_argument("--launch-modes", default=",".join(LAUNCH_MODES),
                    help="comma list of the batch launch modes to time (default: all): %s" % ",".join(LAUNCH_MODES))
    ap.add_argument("--stage-kernels", action="store_true",
                    help="only time the five per-stage kernels (gst_stage_*) on the 512x512 and 4096x4096 "
                         "single-image configurations and print their GB/s")
    ap.add_argument("--h2d-rates", action="store_true",
                    help="only time host-to-device uploads of 1 GiB of pinned memory in ranges of one "
                         "size (64 KiB .. 16 MiB): a DMA copy per range against gst_upload_gather")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of 512 MiB (5x): a kernel of known traffic "
                         "against which HBM byte counters are checked")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    gpu = Gpu()
    if args.copy_calibration:
        n = 512 << 20
        src = gpu.torch.ones(n, dtype=gpu.torch.uint8, device=gpu.dev)
        dst = gpu.empty(n)
        ms = gpu.timed(lambda: dst.copy_(src), 5)
        print(json.dumps({"copy_bytes_each_way": n, "ms": round(ms, 4),
                          "GBs_read_plus_write": round(2 * n / (ms * 1e-3) / 1e9, 1)}), flush=True)
        return

    import gst_amd
    from gst_amd import encode
    if not gst_amd.InitializeDecoder(0):
        raise SystemExit("cuda:0 is not a usable gfx950")
    if args.h2d_rates:
        print(json.dumps({"h2d_GBs": run_h2d_rates(gpu)}), flush=True)
        return
    if args.stage_kernels:
        print(json.dumps({"stage_kernels": run_stage_kernels(gpu)}), flush=True)
        return

    t0 = time.perf_counter()
    specs = corpus_c3(args.images)
    enc_threads = host_cores()[1]
    files, dxts = encode.make_textures(specs, nthreads=enc_threads, encoder="gentenc")
    enc_seconds = time.perf_counter() - t0
    log("corpus: %d textures encoded in %.1fs" % (len(files), enc_seconds))

    modes = tuple(m for m in args.launch_modes.split(",") if m)
    if not modes or any(m not in LAUNCH_MODES for m in modes):
        raise SystemExit("--launch-modes takes a comma list of %s" % ",".join(LAUNCH_MODES))
    hdrs, ms, stage, dev_bytes, tuned, launch = run_batched(gpu, files, dxts, args.steps, args.warmup,
                                                            tune=not args.no_tune, modes=modes)
    texels = sum(h.width * h.height for h in hdrs)
    value = texels / (ms * 1e-3) / 1e6
    log("batched: %.3f ms/step  %.0f Mtexel/s  stages %s" % (ms, value, stage))

    if tuned is None or tuned["wavelet_fused"]:  # the batch default runs the wavelet in the assembly
        mode = "tile_fused"
    else:
        mode = "fused" if tuned["assembly_waves_per_group"] > 0 else "separate"
    lds_tables = tuned is None or tuned["ans_builds_tables"] == 1
    sb = stage_bytes(hdrs, mode, lds_tables)
    launched = {k: v for k, v in stage.items() if v is not None}
    dom = max(launched, key=launched.get)
    achieved = sb[dom] / (stage[dom] * 1e-3) / 1e9
    b_pipe, b_min = pipeline_bytes(hdrs, mode, tables_in_lds=lds_tables)
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mtexel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic: seeded gradient + value-noise + rectangle textures, encoded "
                "byte-identically to the reference encoder (gentenc)",
        "config": {
            "workload": "C3 batched: %d seeded textures {512x512,1024x1024,512x1024,2048x2048}, "
                        "one pass of the persistent stage kernels per step, %s"
                        % (len(hdrs), {"graph": "hipGraph replays", "stream": "stream launches",
                                       "two_plans": "two plans in flight on two streams",
                                       "two_plan_graphs": "two plans in flight on two streams, hipGraph replays"}
                           [min(launch, key=launch.get)]),
            "images": len(hdrs),
            "texels": texels,
            "dxt1_blocks": texels // 16,
            "compressed_bytes": sum(len(f) for f in files),
            "plan_device_bytes": dev_bytes,
            "parallelism": "single GPU",
            "corpus_encode": {"encoder": "gste_compress_dxt (byte-identical to gentenc), CPU",
                              "threads": enc_threads, "seconds": round(enc_seconds, 2)},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom, len(hdrs), args.images),
        },
        "roofline_by_kernel": {
            k: {"achieved": round(sb[k] / (v * 1e-3) / 1e9, 1),
                "frac": round(sb[k] / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "algorithmic_bytes": sb[k],
                "traffic": measured_traffic(k, len(hdrs), args.images)}
            for k, v in launched.items() if sb[k]},
        "pipeline_roofline": {
            "B_pipe_bytes": b_pipe,
            "B_pipe_frac": round(b_pipe / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "B_min_bytes": b_min,
            "B_min_frac": round(b_min / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
        },
        "autotune": tuned,
        "launch_ms_per_step": {k: round(v, 4) for k, v in launch.items()},
        "stage_ms": {k: (round(v, 4) if v is not None else None) for k, v in stage.items()},
        "stage_gbs": {k: (round(sb[k] / (v * 1e-3) / 1e9, 1) if v is not None and sb[k] else None)
                      for k, v in stage.items()},
        "stage_note": "null: the stage has no launch of its own in the tuned shapes",
        "last_stages": {"tile_fused": "wavelet, index scan and DXT1 assembly in one kernel",
                        "fused": "wavelet launch, then index scan inside the assembly",
                        "separate": "wavelet, index scan and assembly launches"}[mode],
    }

    c2 = None
    if not args.no_plain:
        _, ms_plain = run_plain(gpu, files, dxts, passes=3, warmup=1)
        plain = texels / (ms_plain * 1e-3) / 1e6
        result["per_image"] = {"value": round(plain, 1), "unit": "Mtexel/s", "ms_per_pass": round(ms_plain, 3),
                               "path": "LoadCompressedDXT per image, two kernels each (rANS, then wavelet + "
                                       "index scan + DXT1 assembly)"}
        result["speedup_vs_per_image"] = round(value / plain, 2)
        log("plain per-image: %.3f ms/pass  %.0f Mtexel/s" % (ms_plain, plain))
        singles = {}
        for name, (seed, w, h) in (("C1_512x512", (0, 512, 512)), ("C2_4096x4096", (1, 4096, 4096))):
            (f,), (d,) = encode.make_textures([(seed, w, h)], encoder="gentenc")
            _, ms_s = run_plain(gpu, [f], [d], passes=5, warmup=2)
            singles[name] = {"ms": round(ms_s, 4), "Mtexel_s": round(w * h / (ms_s * 1e-3) / 1e6, 1)}
            if name.startswith("C2"):
                c2 = (f, d)
        result["single_image"] = singles
        log("single images: %s" % singles)

    if not args.no_ans_api:
        result["ans_api"] = run_ans_api(files)

    if not args.no_served:
        result["served"] = run_served(gpu, [200.0, 1000.0, 5000.0], args.served_seconds, tuned=tuned)

    if not args.no_cpu:
        rate, threads, passes, secs = run_cpu_baseline(files)
        cpu_count, usable = host_cores()
        result["cpu_baseline"] = {
            "value": round(rate, 1), "unit": "Mtexel/s", "cores": threads, "kind": "port",
            "host_cpu_count": cpu_count, "host_cores_usable": usable,
            "host_cpu_affinity": len(os.sched_getaffinity(0)), "host_cgroup_cpu_quota": cgroup_cpu_quota(),
            "sample": "the full C3 corpus (%d textures, %.0f Mtexel) decoded %d times in %.1f s by the C "
                      "restatement on %d threads (every core this process may use), one task per rANS group / "
                      "wavelet tile / block chunk" % (len(files), texels / 1e6, passes, secs, threads),
        }
        log("cpu baseline: %.0f Mtexel/s on %d threads (cpu_count %d)" % (rate, threads, cpu_count))
        if c2 is not None:
            ms_cpu, threads = run_cpu_single(*c2)
            result["cpu_baseline"]["C2_4096x4096"] = {
                "ms": round(ms_cpu, 2), "Mtexel_s": round(4096 * 4096 / (ms_cpu * 1e-3) / 1e6, 1),
                "threads": threads, "gpu_ms": result["single_image"]["C2_4096x4096"]["ms"]}
            log("cpu C2: %.2f ms on %d threads" % (ms_cpu, threads))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
