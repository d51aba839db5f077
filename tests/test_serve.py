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
    need = _dev_bytes(files[0])
    # Room for exactly one request at a time (its input is returned when its decode
    # completes, its output is held until released).
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
        lat_dev, lat_host, res = s.run_load(files, qps=2000.0, duration_s=0.25, seed=3)
        assert 300 <= res["requests"] <= 800  # ~500 expected arrivals
        assert np.all(lat_dev >= 0) and res["wall_s"] >= 0.2
        assert res["failed"] == 0 and res["verified"] == 0
    with Server(1 << 30, dry_run=True) as s:
        _, _, res = s.run_load(files, closed_loop=8, duration_s=0.1, seed=3)
        assert res["requests"] >= 8
    # Several submitting threads: every submission completes once, none is lost.
    with Server(1 << 30, dry_run=True) as s:
        _, _, res = s.run_load(files, closed_loop=64, duration_s=0.2, seed=4, submit_threads=4)
        st = s.stats()
        assert res["requests"] >= 64 and res["failed"] == 0
        assert st["submitted"] == st["completed"] == res["requests"]


def test_failed_launch_fails_its_batch_and_frees_its_memory():
    """A batch whose launch fails part-way completes its requests with the error, frees
    their memory only after the slot's stream has drained, and the server goes on: with
    every second launch failing, the other batches still complete and the budget, which
    holds one request at a time, keeps admitting."""
    f = _files()[0]
    need = _dev_bytes(f)
    with Server(need, dry_run=True, fail_launch_every=2, slots=1) as s:
        tickets = [s.submit(f, tag=i) for i in range(8)]
        s.drain()
        done = s.poll()
        assert sorted(c.ticket for c in done) == sorted(tickets)
        failed = [c for c in done if c.status != 0]
        assert len(failed) == 4 and all(c.status == -5 for c in failed)  # GST_ERR_HIP
        st = s.stats()
        assert st["failed"] == 4 and st["completed"] == 8 and st["batches"] == 8
        assert st["peak_arena_bytes"] <= need


def test_prestaged_requests_take_no_copy_dry_run():
    files = _files()
    with Server(1 << 30, dry_run=True) as s:
        for i, f in enumerate(files):
            s.submit_request(f[:28], tag=i, d_file=0x1000 * (i + 1))  # never dereferenced in a dry run
        s.submit(files[0], tag=99)
        s.drain()
        done = s.poll()
        assert len(done) == len(files) + 1 and all(c.status == 0 for c in done)
        st = s.stats()
        assert st["prestaged"] == len(files) and st["submitted"] == len(files) + 1
    with Server(1 << 30, dry_run=True) as s:
        with pytest.raises(gst_amd.GstError) as e:
            s.submit_request(files[0], d_file=0x1000, d_expected=0x1008)  # misaligned expected output
        assert e.value.code == -1


# ---- GPU ---------------------------------------------------------------------------------

def _hip_download():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]

    def download(c):
        host = np.empty(int(c.output_bytes), np.uint8)
        assert hip.hipMemcpy(host.ctypes.data, c.d_output, int(c.output_bytes), 2) == 0
        return host.tobytes()
    return download


# Every size of the served mix (512^2 ... 4096^2) plus odd shapes.
SERVED_SPECS = [(501, 512, 512), (502, 1024, 1024), (503, 512, 256), (504, 2048, 2048), (505, 768, 512),
                (506, 512, 1024), (507, 4096, 4096)]


@pytest.fixture(scope="module")
def served_corpus():
    files, dxts = encode.make_textures(SERVED_SPECS)
    files += [load_golden("ref_test1")["gtc"].tobytes()]
    dxts += [load_golden("ref_test1")["dxt"].tobytes()]
    return files, dxts


@pytest.mark.gpu
@pytest.mark.parametrize("graph_launch,copy_kernel", [(False, False), (True, False), (False, True)])
def test_served_results_equal_alone_and_in_the_mix(device, served_corpus, graph_launch, copy_kernel):
    """Each request decoded alone (one at a time) and in a concurrent stream of
    arrivals gives the same bit-exact blocks, with graph replays or stream launches,
    inputs copied in by DMA or by the gather kernel.  Outputs are poisoned before every
    decode, and the in-mix run also compares each output with its expected blocks on
    the device."""
    files, dxts = served_corpus
    download = _hip_download()
    alone = {}
    with Server(1 << 30, auto_release=False, graph_launch=graph_launch, poison_outputs=True,
                copy_kernel=copy_kernel) as s:
        for i, f in enumerate(files):
            s.submit(f, tag=i)
            s.drain()
            (c,) = s.poll()
            assert c.status == 0 and c.batch_images == 1
            alone[i] = download(c)
            s.release(c.ticket)
    assert [alone[i] for i in range(len(files))] == dxts

    d_expected = [device.upload(d) for d in dxts]
    with Server(1 << 30, auto_release=False, slots=3, graph_launch=graph_launch, poison_outputs=True,
                copy_kernel=copy_kernel) as s:
        order = list(range(len(files))) * 4
        np.random.default_rng(0).shuffle(order)
        for i in order:
            s.submit_request(files[i], tag=i, d_expected=d_expected[i].data_ptr())
        s.drain()
        done = s.poll(4096)
        assert len(done) == len(order)
        assert max(c.batch_images for c in done) > 1, "arrivals were batched"
        for c in done:
            assert c.status == 0
            assert c.verified == 1, "the device comparison found a difference (tag %d)" % c.tag
            assert download(c) == alone[c.tag]
            s.release(c.ticket)
        st = s.stats()
        assert st["verified"] == len(order) and st["mismatches"] == 0
        assert st["h2d_copies"] <= st["submitted"]


@pytest.mark.gpu
def test_device_comparison_reports_a_wrong_expected_output(device, served_corpus):
    """The comparison is not vacuous: an expected output with one flipped byte is
    reported as differing, the others as equal."""
    files, dxts = served_corpus
    wrong = bytearray(dxts[0])
    wrong[len(wrong) // 2] ^= 0x10
    d_good = device.upload(dxts[0])
    d_wrong = device.upload(bytes(wrong))
    with Server(1 << 30, poison_outputs=True) as s:
        s.submit_request(files[0], tag=0, d_expected=d_good.data_ptr())
        s.submit_request(files[0], tag=1, d_expected=d_wrong.data_ptr())
        s.submit(files[0], tag=2)
        s.drain()
        done = {c.tag: c for c in s.poll()}
    assert done[0].verified == 1 and done[1].verified == 2 and done[2].verified == 0
    assert all(c.status == 0 for c in done.values())


@pytest.mark.gpu
def test_prestaged_inputs_decode_without_copies(device, served_corpus):
    """Files already in HBM are decoded in place: no host-to-device copy, same blocks."""
    files, dxts = served_corpus
    d_files = [device.upload(f) for f in files]
    d_expected = [device.upload(d) for d in dxts]
    download = _hip_download()
    with Server(1 << 30, auto_release=False, poison_outputs=True) as s:
        for i, (f, df) in enumerate(zip(files, d_files)):
            s.submit_request(f[:28], tag=i, d_file=df.data_ptr(), d_expected=d_expected[i].data_ptr())
        s.drain()
        done = s.poll()
        assert len(done) == len(files)
        for c in done:
            assert c.status == 0 and c.verified == 1
            assert download(c) == dxts[c.tag]
            s.release(c.ticket)
        st = s.stats()
        assert st["prestaged"] == len(files) and st["h2d_bytes"] == 0


@pytest.mark.gpu
def test_served_poisson_load_checks_every_completion(device, served_corpus):
    """The load generator under Poisson arrivals and in closed loops, host-submitted
    (inputs copied by DMA or by the gather kernel) and pre-staged: every completion is
    compared with its expected blocks on the device."""
    files, dxts = served_corpus
    d_files = [device.upload(f) for f in files]
    d_expected = [device.upload(d) for d in dxts]
    exp = [t.data_ptr() for t in d_expected]
    with Server(1 << 30, poison_outputs=True, copy_to_host=True) as s:
        _, _, res = s.run_load(files, qps=2000.0, duration_s=0.3, seed=11, pool_expected=exp)
        assert res["requests"] > 100 and res["failed"] == 0
        assert res["verified"] == res["requests"] and res["mismatches"] == 0
    with Server(1 << 30, poison_outputs=True, copy_kernel=True) as s:
        _, _, res = s.run_load(files, closed_loop=64, duration_s=0.3, seed=13, pool_expected=exp)
        assert res["requests"] > 64 and res["failed"] == 0
        assert res["verified"] == res["requests"] and res["mismatches"] == 0
        assert s.stats()["h2d_copies"] > 0
    with Server(1 << 30, poison_outputs=True) as s:
        _, _, res = s.run_load(files, closed_loop=32, duration_s=0.3, seed=12, pool_expected=exp,
                               pool_device=[t.data_ptr() for t in d_files])
        assert res["requests"] > 32 and res["failed"] == 0
        assert res["verified"] == res["requests"] and res["mismatches"] == 0


@pytest.mark.gpu
@pytest.mark.parametrize("copy_kernel", [False, True])
def test_plain_mode_matches(device, served_corpus, copy_kernel):
    files, dxts = served_corpus
    g = load_golden("ref_wide")
    download = _hip_download()
    d_expected = device.upload(dxts[-2])
    with Server(1 << 28, plain=True, auto_release=False, poison_outputs=True, copy_kernel=copy_kernel) as s:
        s.submit(g["gtc"].tobytes(), tag=0)
        s.submit_request(files[-2], tag=1, d_expected=d_expected.data_ptr())  # 4096x4096
        s.drain()
        done = {c.tag: c for c in s.poll()}
        assert done[0].status == 0 and download(done[0]) == g["dxt"].tobytes()
        assert done[1].status == 0 and done[1].verified == 1 and download(done[1]) == dxts[-2]
        for c in done.values():
            s.release(c.ticket)


@pytest.mark.gpu
def test_plain_mode_reports_each_corrupt_request(device):
    """Plain mode gives every request its own device error bits: a corrupt file fails,
    the valid requests before and after it do not."""
    import struct
    gtc = load_golden("ref_c0")["gtc"].tobytes()
    bad = bytearray(gtc)
    struct.pack_into("<H", bad, 28, struct.unpack_from("<H", bad, 28)[0] + 3)  # Y frequencies no longer sum to 2048
    with Server(1 << 28, plain=True) as s:
        s.submit(gtc, tag=0)
        s.submit(bytes(bad), tag=1)
        s.submit(gtc, tag=2)
        s.drain()
        done = {c.tag: c.status for c in s.poll()}
    assert done == {0: 0, 1: -3, 2: 0}
