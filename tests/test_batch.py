"""Batched decode of textures of uneven size: one plan, persistent launches, and the
captured hipGraph replayed for different bound batches."""
import numpy as np
import pytest

from conftest import load_golden

import gst_amd
from gst_amd import encode
from oracle import gst_oracle as go

pytestmark = pytest.mark.gpu

MIXED = [(301, 512, 512), (302, 1024, 1024), (303, 512, 1024), (304, 2048, 2048), (305, 768, 512),
         (306, 512, 256), (307, 1024, 512)]


def _bind(device, batch, files):
    hdrs = [gst_amd.ParseHeader(f) for f in files]
    d_files = [device.upload(f) for f in files]
    d_outs = [device.empty(h.dxt_bytes) for h in hdrs]
    batch.bind(hdrs, [t.data_ptr() for t in d_files], [t.data_ptr() for t in d_outs], device.stream)
    return hdrs, d_files, d_outs


def _capacity(files):
    hdrs = [gst_amd.ParseHeader(f) for f in files]
    return sum(h.num_blocks for h in hdrs), sum(h.palette_bytes for h in hdrs)


def test_uneven_batch_matches_oracle(device):
    files, dxts = encode.make_textures(MIXED)
    files += [load_golden("ref_test1")["gtc"].tobytes()]
    dxts += [load_golden("ref_test1")["dxt"].tobytes()]
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal)
    try:
        hdrs, d_files, d_outs = _bind(device, b, files)
        b.decode(device.stream)
        for f, d, out in zip(files, dxts, d_outs):
            rc, expected, _ = go.decode_dxt(f, nthreads=8)
            assert rc == go.OK and expected == d
            assert device.download(out) == expected
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


@pytest.mark.parametrize("gpw,waves,wwaves,awaves,separate,sep_tables,wide,sep_wavelet", [
    (2, 4, 1, 4, False, False, False, True), (1, 4, 2, 1, False, True, True, True),
    (2, 8, 4, 2, False, False, False, True), (1, 8, 1, 4, True, False, False, True),
    (2, 1, 1, 0, True, False, False, True), (1, 2, 4, 0, True, True, False, True),
    (2, 4, 2, 2, False, False, True, True), (2, 8, 1, 4, False, True, True, True),
    (2, 4, 1, 4, False, False, False, False), (1, 8, 1, 4, False, True, False, False),
    (2, 2, 1, 4, False, False, False, False), (2, 2, 2, 2, False, False, False, True)])
def test_launch_shapes_agree(device, gpw, waves, wwaves, awaves, separate, sep_tables, wide, sep_wavelet):
    files, dxts = encode.make_textures(MIXED[:4])
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal, groups_per_wave=gpw, waves_per_group=waves,
                      grids={"ans_decode": 64, "inv_wavelet": 128, "decode_indices": 16, "assemble": 32},
                      wavelet_waves=wwaves, assembly_waves=awaves, separate_index_scan=separate,
                      separate_table_stage=sep_tables, assembly_wide=wide, separate_wavelet=sep_wavelet)
    try:
        _, _, d_outs = _bind(device, b, files)
        b.decode(device.stream)
        for d, out in zip(dxts, d_outs):
            assert device.download(out) == d
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


def test_graph_replays_match_eager_and_follow_rebinds(device):
    torch = device.torch
    files_a, dxts_a = encode.make_textures(MIXED[:5])
    files_b, dxts_b = encode.make_textures([(401, 1024, 1024), (402, 512, 256), (403, 512, 512)])
    blocks = max(_capacity(files_a)[0], _capacity(files_b)[0])
    pal = max(_capacity(files_a)[1], _capacity(files_b)[1])
    b = gst_amd.Batch(8, blocks, pal)
    try:
        _, keep_a, outs_a = _bind(device, b, files_a)
        b.decode(device.stream)
        eager = [device.download(o) for o in outs_a]
        assert eager == dxts_a
        b.capture_graph(device.stream)
        for rep in range(3):
            for o in outs_a:
                device.poison(o)
            scratch = torch.randn(1 << 20, device=device.dev)  # eager work between replays
            scratch.mul_(2.0)
            b.launch_graph(device.stream)
            device.sync()
            assert [device.download(o) for o in outs_a] == eager, "replay %d" % rep
        # Rebind a different batch to the same plan; the captured graph follows it.
        _, keep_b, outs_b = _bind(device, b, files_b)
        b.launch_graph(device.stream)
        assert [device.download(o) for o in outs_b] == dxts_b
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


@pytest.mark.parametrize("opts,launched", [
    # Default: tables built in the rANS kernel, wavelet and index scan in the assembly.
    ({}, {"ans_decode", "assemble"}),
    ({"separate_wavelet": True}, {"ans_decode", "inv_wavelet", "assemble"}),
    ({"separate_index_scan": True, "separate_table_stage": True}, set(gst_amd.STAGES)),
])
def test_decode_timed_reports_launched_stages_only(device, opts, launched):
    """A stage with no launch of its own reports None, not the empty gap between two
    events; every launched stage reports a positive time."""
    files, dxts = encode.make_textures(MIXED[:3])
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal, **opts)
    try:
        _, _, outs = _bind(device, b, files)
        ms = b.decode_timed(device.stream)
        assert set(ms) == set(gst_amd.STAGES)
        assert {k for k, v in ms.items() if v is not None} == launched
        assert all(ms[k] > 0 for k in launched)
        assert [device.download(o) for o in outs] == dxts
        assert b.error_bits(device.stream) == 0
        # Several decodes back to back: the same stages, mean times, the same blocks.
        for o in outs:
            device.poison(o)
        ms = b.decode_timed(device.stream, reps=5)
        assert {k for k, v in ms.items() if v is not None} == launched
        assert all(ms[k] > 0 for k in launched)
        assert [device.download(o) for o in outs] == dxts
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


def test_bind_rejects_batches_over_capacity(device):
    files, _ = encode.make_textures(MIXED[:2])
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(1, blocks, pal)
    try:
        with pytest.raises(gst_amd.GstError):
            _bind(device, b, files)
    finally:
        b.close()
    b = gst_amd.Batch(4, blocks - 8192, pal)
    try:
        with pytest.raises(gst_amd.GstError) as e:
            _bind(device, b, files)
        assert e.value.code == -4
    finally:
        b.close()


def test_autotune_keeps_results_bit_exact(device):
    files, dxts = encode.make_textures(MIXED)
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal)
    try:
        _, _, outs = _bind(device, b, files)
        tuned = b.autotune(device.stream, reps=3)
        assert tuned["streams_per_wavefront"] in (32, 64)
        assert tuned["wavelet_waves_per_group"] in (1, 2, 4)
        assert tuned["candidates"] >= 7 * 5 + 3 * 5 + 2 * 5
        # A fused index scan, a wavelet inside the assembly and tables built in LDS
        # have no launch of their own.
        folded = {"decode_indices": tuned["assembly_waves_per_group"] > 0 or tuned["wavelet_fused"] == 1,
                  "inv_wavelet": tuned["wavelet_fused"] == 1,
                  "build_table": tuned["ans_builds_tables"] == 1}
        for k, v in tuned["stage_ms"].items():
            assert v is None if folded.get(k) else v > 0, (k, v)
        for o in outs:
            device.poison(o)
        b.decode(device.stream)
        assert [device.download(o) for o in outs] == dxts
        for o in outs:
            device.poison(o)
        b.capture_graph(device.stream)
        b.launch_graph(device.stream)
        assert [device.download(o) for o in outs] == dxts
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


@pytest.mark.parametrize("shapes", [
    # 3 tiles per tile row: work-group pairs of tile positions straddle tile rows.
    [(501, 384, 1024), (502, 640, 1024)],
    # 528 tiles per tile row: a pair's rows touch more than 64 index-stream groups.
    [(503, 67584, 128)],
])
def test_wavelet_inside_assembly_on_odd_and_wide_tilings(device, shapes):
    files, dxts = encode.make_textures(shapes)
    blocks, pal = _capacity(files)
    fused = gst_amd.Batch(len(files), blocks, pal, grids={"assemble": 7})
    separate = gst_amd.Batch(len(files), blocks, pal, separate_wavelet=True)
    try:
        for b in (fused, separate):
            _, _, outs = _bind(device, b, files)
            b.decode(device.stream)
            assert [device.download(o) for o in outs] == dxts
            assert b.error_bits(device.stream) == 0
    finally:
        fused.close()
        separate.close()


@pytest.mark.parametrize("capacity", [255, 256, 300, 5000])
def test_large_plan_capacities_locate_units(device, capacity):
    # Plans of more than 255 images find a unit's image with a 64-way narrowing per
    # round trip instead of one load of all prefix entries; 255 and 256 straddle the
    # switch.  Bound batches smaller than the capacity leave padded prefix entries.
    files, dxts = encode.make_textures(MIXED[:3] + [(601, 512, 256)] * 2)
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(capacity, blocks, pal)
    try:
        _, _, outs = _bind(device, b, files)
        b.decode(device.stream)
        assert [device.download(o) for o in outs] == dxts
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


def test_full_plan_of_many_small_images(device):
    # 300 bound images: every prefix entry is real and the narrowing path runs to the
    # last image.
    specs = [(700 + i, 512, 256) for i in range(300)]
    files, dxts = encode.make_textures(specs)
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal)
    try:
        _, _, outs = _bind(device, b, files)
        b.decode(device.stream)
        got = [device.download(o) for o in outs]
        assert got == dxts
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


@pytest.mark.parametrize("ans_grid,asm_grid", [(9, 13), (3, 1), (1568, 7), (64, 5003), (64, 8)])
def test_odd_persistent_grids(device, ans_grid, asm_grid):
    """Grids that are not multiples of the 8 XCDs, below 8, exactly 8, and far above the
    unit count: the fused kernel's per-XCD unit ranges must still visit every unit exactly
    once, pass after pass."""
    files, dxts = encode.make_textures(MIXED[:5])
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal, grids={"ans_decode": ans_grid, "assemble": asm_grid})
    try:
        _, _, d_outs = _bind(device, b, files)
        b.decode(device.stream)
        for d, out in zip(dxts, d_outs):
            assert device.download(out) == d
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


def test_two_plans_in_flight(device):
    """Two plans bound to the same inputs, each with its own outputs and stream, used
    alternately so that their decodes overlap on the device."""
    torch = device.torch
    files, dxts = encode.make_textures(MIXED[:6])
    blocks, pal = _capacity(files)
    hdrs = [gst_amd.ParseHeader(f) for f in files]
    d_files = [device.upload(f) for f in files]
    streams = [torch.cuda.Stream(device=device.dev), torch.cuda.Stream(device=device.dev)]
    plans, outs = [], []
    try:
        for st in streams:
            p = gst_amd.Batch(len(files), blocks, pal)
            o = [device.empty(h.dxt_bytes) for h in hdrs]
            p.bind(hdrs, [t.data_ptr() for t in d_files], [t.data_ptr() for t in o], st.cuda_stream)
            plans.append(p)
            outs.append(o)
        torch.cuda.synchronize()
        for k in range(8):
            plans[k % 2].decode(streams[k % 2].cuda_stream)
        torch.cuda.synchronize()
        for o in outs:
            for d, out in zip(dxts, o):
                assert device.download(out) == d
        for p, st in zip(plans, streams):
            assert p.error_bits(st.cuda_stream) == 0
    finally:
        for p in plans:
            p.close()


def test_full_size_batch_of_reference_encoded_textures(device):
    """The benchmark's workload at its full size: 128 seeded textures of 512x512 ..
    2048x2048 (207.9 Mtexel), each file byte for byte the reference encoder's, decoded as
    one batch with the default shapes (short stream remainders paired) and as a captured
    graph: every block equals the encoder's."""
    import os
    sizes = [(512, 512), (1024, 1024), (512, 1024), (2048, 2048)]
    picks = np.random.default_rng(42).integers(0, len(sizes), 128)
    specs = [(1000 + i, sizes[p][0], sizes[p][1]) for i, p in enumerate(picks)]
    files, dxts = encode.make_textures(specs, nthreads=min(16, os.cpu_count() or 1), encoder="gentenc")
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal)
    try:
        hdrs, keep, outs = _bind(device, b, files)
        assert sum(h.width * h.height for h in hdrs) == 207880192
        b.decode(device.stream)
        assert [device.download(o) for o in outs] == dxts
        b.capture_graph(device.stream)
        for o in outs:
            device.poison(o)
        b.launch_graph(device.stream)
        assert [device.download(o) for o in outs] == dxts
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()


def _garbled(gtc, stream, seed):
    """ref_c0 with 200 bytes of noise in the renormalisation words of one stream (the
    group offsets stay valid): a different bitstream that still decodes everywhere."""
    h = gst_amd.ParseHeader(gtc)
    starts = {"y": 28 + 2048 + 4 * (2 * h.num_blocks // 8192),
              "chroma": 28 + 2048 + h.y_cmp_sz + 4 * (4 * h.num_blocks // 8192),
              "palette": 28 + 2048 + h.y_cmp_sz + h.chroma_cmp_sz + 4 * (h.palette_bytes // 8192),
              "indices": 28 + 2048 + h.y_cmp_sz + h.chroma_cmp_sz + h.palette_sz + 4 * (h.num_blocks // 8192)}
    start = starts[stream]
    b = bytearray(gtc)
    b[start:start + 200] = np.random.default_rng(seed).integers(0, 256, 200, dtype=np.uint8).tobytes()
    return bytes(b)


@pytest.mark.parametrize("form", ["auto", "paired", "plain"])
def test_ans_forms_match_oracle(device, form):
    """Every rANS work-group form of the 4-wave x 2-group shape on an uneven batch, with
    reference files and garbled streams (noise in the words: each form must decode the
    bits exactly as the oracle does, including palette indices that leave the palette)."""
    files, dxts = encode.make_textures(MIXED)
    files += [load_golden("ref_test1")["gtc"].tobytes()]
    dxts += [load_golden("ref_test1")["dxt"].tobytes()]
    c0 = load_golden("ref_c0")["gtc"].tobytes()
    bad_palette = False
    # Seeds whose noise keeps every group's words inside the group (the oracle decodes
    # them without a stream error).
    for stream, seed in (("y", 10), ("chroma", 11), ("palette", 13)):
        g = _garbled(c0, stream, seed)
        rc, expected, _ = go.decode_dxt(g, nthreads=8)
        assert rc >= 0
        bad_palette |= rc > 0
        files.append(g)
        dxts.append(expected)
    blocks, pal = _capacity(files)
    b = gst_amd.Batch(len(files), blocks, pal, ans_form=form)
    try:
        _, _, d_outs = _bind(device, b, files)  # outputs poisoned
        b.decode(device.stream)
        for i, (d, out) in enumerate(zip(dxts, d_outs)):
            assert device.download(out) == d, "image %d" % i
        assert b.error_bits(device.stream) == (0x4 if bad_palette else 0)
    finally:
        b.close()


@pytest.mark.parametrize("form", ["auto", "paired"])
def test_ans_forms_report_bad_frequencies_alike(device, form):
    """A frequency table that does not sum to 2048 raises the frequency bit in every form,
    and every form decodes it with the same identity table (same output bytes)."""
    c0 = load_golden("ref_c0")["gtc"].tobytes()
    b = bytearray(c0)
    b[28:30] = (int.from_bytes(b[28:30], "little") + 3).to_bytes(2, "little")
    files = [bytes(b), encode.make_textures([(310, 512, 512)])[0][0]]
    blocks, pal = _capacity(files)
    outs = {}
    for f in ("plain", form):
        batch = gst_amd.Batch(len(files), blocks, pal, ans_form=f)
        try:
            _, _, d_outs = _bind(device, batch, files)
            batch.decode(device.stream)
            outs[f] = [device.download(o) for o in d_outs]
            assert batch.error_bits(device.stream) & 0x2
        finally:
            batch.close()
    assert outs[form] == outs["plain"]


def test_auto_form_switches_under_one_captured_graph(device):
    """GST_ANS_FORM_AUTO decodes a small batch in work-groups of 8 groups of one stream and
    pairs short remainders once those work-groups no longer fit the device in one pass;
    the rANS launch reads its form from the plan, so one captured graph replays both,
    bit-exact."""
    torch = device.torch
    (big,), (big_dxt,) = encode.make_textures([(320, 4096, 4096)])
    (small,), (small_dxt,) = encode.make_textures([(321, 512, 512)])
    h = gst_amd.ParseHeader(big)
    per_big = sum((g + 7) // 8 for g in (2 * h.num_blocks // 8192, 4 * h.num_blocks // 8192,
                                         h.palette_bytes // 8192, h.num_blocks // 8192))
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    n_big = 6 * cus // per_big + 1  # past one pass of 6 work-groups per CU
    assert n_big * per_big > 6 * cus
    b = gst_amd.Batch(n_big + 1, (n_big + 1) * h.num_blocks, (n_big + 1) * h.palette_bytes)
    try:
        _, keep_s, outs_s = _bind(device, b, [small])
        b.capture_graph(device.stream)
        b.launch_graph(device.stream)
        assert device.download(outs_s[0]) == small_dxt
        _, keep_b, outs_b = _bind(device, b, [big] * n_big + [small])
        for rep in range(2):
            for o in outs_b:
                device.poison(o)
            b.launch_graph(device.stream)
            device.sync()
            got = [device.download(o) for o in outs_b]
            assert got[:-1] == [big_dxt] * n_big and got[-1] == small_dxt, "replay %d" % rep
        _, keep_s2, outs_s2 = _bind(device, b, [small])
        b.launch_graph(device.stream)
        assert device.download(outs_s2[0]) == small_dxt
        assert b.error_bits(device.stream) == 0
    finally:
        b.close()
