"""The continuous-batching server: scheduling and admission on the CPU (dry run, no GPU
work), and bit-exact results alone vs in the mix on the GPU."""
import time

import numpy as np
import pytest

import gst_amd
from gst_amd import encode
from gst_amd.serve import Server

from conftest import load_golden


def _files():
    return [load_golden(n)["gtc"].tobytes() for n in ("ref_c0", "ref_small", "ref_wide", "ref_test1")]


def _dev_bytes(f):
    h = gst_amd.ParseHeader(f)
    a = lambda v: (v + 255) // 256 * 256
    return a(len(f)) + a(h.dxt_bytes)


# ---- CPU: dry run ---------------------------------------------------------------------

def test_dry_run_completes_every_request_once():
    files = _files()
    with Server(1 << 30, dry_run=True) as s:
        tickets = [s.submit(f, tag=i) for i, f in enumerate(files * 5)]
        s.drain()
        done = s.poll()
        assert sorted(c.ticket for c in done) == sorted(tickets)
        assert sorted(c.tag for c in done) == list(range(len(tickets)))
        st = s.stats()
        assert st["completed"] == len(tickets) and st["failed"] == 0
        assert st["texels"] == sum(gst_amd.ParseHeader(f).width * gst_amd.ParseHeader(f).height
                                   for f in files * 5)


def test_admission_respects_budget_and_waits_for_release():
    files = _files()
    need = max(_dev_bytes(f) for f in files)
    # Room for exactly one request at a time; completions are held until released.
    with Server(need, dry_run=True, auto_release=False) as s:
        tickets = [s.submit(files[0]) for _ in range(3)]
        deadline = time.time() + 5
        done = []
        while len(done) < 1 and time.time() < deadline:
            done += s.poll()
            time.sleep(0.001)
        time.sleep(0.02)
        done += s.poll()
        assert len(done) == 1, "only one request fits the budget until a release"
        st = s.stats()
        assert st["peak_arena_bytes"] <= need and st["admission_waits"] >= 1
        for _ in range(2):
            s.release(done[-1].ticket)
            while True:
                got = s.poll()
                if got:
                    done += got
                    break
                time.sleep(0.001)
        s.release(done[-1].ticket)
        assert sorted(c.ticket for c in done) == sorted(tickets)


def test_requests_that_can_never_fit_are_rejected():
    f = load_golden("ref_wide")["gtc"].tobytes()
    with Server(_dev_bytes(f) - 256, dry_run=True) as s:
        with pytest.raises(gst_amd.GstError) as e:
            s.submit(f)
        assert e.value.code == -4
    with Server(1 << 30, dry_run=True, max_batch_blocks=8192) as s:
        with pytest.raises(gst_amd.GstError):
            s.submit(f)  # 32768 blocks > one batch
    with Server(1 << 30, dry_run=True) as s:
        with pytest.raises(gst_amd.GstError) as e:
            s.submit(f[:100])
        assert e.value.code == -2


def test_batches_respect_image_and_block_limits():
    files = [load_golden("ref_c0")["gtc"].tobytes()] * 40
    with Server(1 << 30, dry_run=True, max_batch_images=4, slots=1) as s:
        for f in files:
            s.submit(f)
        s.drain()
        done = s.poll()
        assert len(done) == 40 and max(c.batch_images for c in done) <= 4
        assert s.stats()["max_batch_images"] <= 4


def test_poisson_load_generator_dry_run():
    files = _files()
    with Server(1 << 30, dry_run=True) as s:
        lat_dev, lat_host, n, wall = s.run_load(files, qps=2000.0, duration_s=0.25, seed=3)
        assert 300 <= n <= 800  # ~500 expected arrivals
        assert np.all(lat_dev >= 0) and wall >= 0.2
    with Server(1 << 30, dry_run=True) as s:
        _, _, n, _ = s.run_load(files, closed_loop=8, duration_s=0.1, seed=3)
        assert n >= 8


# ---- GPU ---------------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("direct_launch", [False, True])
def test_served_results_equal_alone_and_in_the_mix(device, direct_launch):
    """Each request decoded alone (one at a time) and in a concurrent stream of
    arrivals gives the same bit-exact blocks, with graph replays or stream launches."""
    torch = device.torch
    specs = [(501, 512, 512), (502, 1024, 1024), (503, 512, 256), (504, 2048, 2048), (505, 768, 512),
             (506, 512, 1024)]
    files, dxts = encode.make_textures(specs)
    files += [load_golden("ref_test1")["gtc"].tobytes()]
    dxts += [load_golden("ref_test1")["dxt"].tobytes()]

    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def download(c):
        host = np.empty(int(c.output_bytes), np.uint8)
        assert hip.hipMemcpy(host.ctypes.data, c.d_output, int(c.output_bytes), 2) == 0
        return host.tobytes()

    alone = {}
    with Server(1 << 30, auto_release=False, direct_launch=direct_launch) as s:
        for i, f in enumerate(files):
            s.submit(f, tag=i)
            s.drain()
            (c,) = s.poll()
            assert c.status == 0 and c.batch_images == 1
            alone[i] = download(c)
            s.release(c.ticket)
    assert [alone[i] for i in range(len(files))] == dxts

    with Server(1 << 30, auto_release=False, slots=3, direct_launch=direct_launch) as s:
        order = list(range(len(files))) * 4
        np.random.default_rng(0).shuffle(order)
        for i in order:
            s.submit(files[i], tag=i)
        s.drain()
        done = s.poll(4096)
        assert len(done) == len(order)
        assert max(c.batch_images for c in done) > 1, "arrivals were batched"
        for c in done:
            assert c.status == 0
            assert download(c) == alone[c.tag]
            s.release(c.ticket)


@pytest.mark.gpu
def test_plain_mode_matches(device):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    g = load_golden("ref_wide")
    with Server(1 << 28, plain=True, auto_release=False) as s:
        s.submit(g["gtc"].tobytes())
        s.drain()
        (c,) = s.poll()
        host = np.empty(int(c.output_bytes), np.uint8)
        assert hip.hipMemcpy(host.ctypes.data, c.d_output, int(c.output_bytes), 2) == 0
        assert host.tobytes() == g["dxt"].tobytes()
        s.release(c.ticket)
