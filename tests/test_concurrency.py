"""Several host threads driving the drop-in entry points at once, as the reference's
callers do: demo/photos_sf.cpp:747-830 loads pages from a thread pool, and the reference
guards its enqueues and its scratch with mutexes (gpu/gpu.h:118, codec/decoder.cpp:77-82).
Each thread has its own HIP stream and calls LoadCompressedDXT and LoadCompressedDXTs;
every output (poisoned beforehand) must equal the oracle's blocks, with stream-ordered
scratch and with one preallocation shared by all threads."""
import ctypes
import threading

import pytest

from conftest import load_golden

import gst_amd
from gst_amd import encode
from oracle import gst_oracle as go

pytestmark = pytest.mark.gpu

THREADS = 6


@pytest.fixture(scope="module")
def corpus():
    specs = [(1100 + i, w, h) for i, (w, h) in enumerate([(512, 512), (1024, 1024), (512, 256), (768, 512),
                                                          (2048, 1024), (512, 1024), (1024, 512), (512, 512)])]
    files, dxts = encode.make_textures(specs)
    files += [load_golden(n)["gtc"].tobytes() for n in ("ref_c0", "ref_wide", "ref_test1")]
    dxts += [load_golden(n)["dxt"].tobytes() for n in ("ref_c0", "ref_wide", "ref_test1")]
    for f, d in zip(files, dxts):
        rc, expected, _ = go.decode_dxt(f, nthreads=8)
        assert rc == go.OK and expected == d
    return files, dxts


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


@pytest.mark.parametrize("prealloc", [False, True])
def test_threads_on_their_own_streams_decode_bit_exact(device, corpus, prealloc):
    files, dxts = corpus
    hip = _hip()
    # Per thread: two single-image calls and one three-image page, from its own staging.
    jobs = []
    for t in range(THREADS):
        picks = [(t + k) % len(files) for k in range(5)]
        single = [picks[0], picks[1]]
        page = picks[2:5]
        job = {"calls": []}
        for idx in single:
            staging, (hdr,) = gst_amd.PackStaging([files[idx]])
            job["calls"].append(("one", [hdr], device.upload(staging), device.empty(hdr.dxt_bytes), [idx]))
        staging, hdrs = gst_amd.PackStaging([files[i] for i in page])
        job["calls"].append(("page", hdrs, device.upload(staging), device.empty(sum(h.dxt_bytes for h in hdrs)),
                             page))
        jobs.append(job)
    device.sync()
    if prealloc:
        # Every call carves its scratch from the preallocation in call order (three rounds).
        total = 3 * sum(gst_amd.RequiredScratchMem(h) for j in jobs for c in j["calls"] for h in c[1])
        gst_amd.PreallocateDecompressor(total)
    streams = []
    for _ in range(THREADS):
        s = ctypes.c_void_p()
        assert hip.hipStreamCreate(ctypes.byref(s)) == 0
        streams.append(s.value)
    start = threading.Barrier(THREADS)
    errors = []

    def work(t):
        try:
            start.wait()
            for _ in range(3):  # the same calls again: outputs must stay the same
                for kind, hdrs, d_cmp, d_out, _ in jobs[t]["calls"]:
                    if kind == "one":
                        gst_amd.LoadCompressedDXT(hdrs[0], streams[t], d_cmp.data_ptr(), d_out.data_ptr())
                    else:
                        gst_amd.LoadCompressedDXTs(hdrs, streams[t], d_cmp.data_ptr(), d_out.data_ptr())
            assert hip.hipStreamSynchronize(streams[t]) == 0
            jobs[t]["bits"] = gst_amd.ReadErrorBits(streams[t])
        except Exception as e:  # reported below, on the main thread
            errors.append((t, repr(e)))

    try:
        threads = [threading.Thread(target=work, args=(t,)) for t in range(THREADS)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=120)
        assert not any(th.is_alive() for th in threads), "a decoding thread did not finish"
        assert not errors, errors
        for t, job in enumerate(jobs):
            assert job["bits"] == 0, "thread %d raised device error bits" % t
            for kind, hdrs, _, d_out, idx in job["calls"]:
                assert device.download(d_out) == b"".join(dxts[i] for i in idx), (t, kind, idx)
    finally:
        device.sync()
        if prealloc:
            gst_amd.FreeDecompressor()
        for s in streams:
            hip.hipStreamDestroy(s)
