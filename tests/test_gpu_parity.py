"""Parity of the gfx950 decoder (through the C ABI) with the oracle and the reference's
golden vectors.  Bit-exact everywhere: all stages are integer work."""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN_TEXTURES, load_golden

import gst_amd
from gst_amd import encode
from oracle import gst_oracle as go

pytestmark = pytest.mark.gpu


def packed_tables(gtc):
    """The decoder's 4-byte slot entries from the oracle's table (f | s-cum << 12 | x << 24)."""
    out = []
    for k in range(4):
        freqs = np.frombuffer(gtc[28 + 512 * k:28 + 512 * (k + 1)], "<u2")
        f, c, s = go.build_table(freqs)
        slot = np.arange(2048, dtype=np.uint32)
        out.append(f.astype(np.uint32) | ((slot - c.astype(np.uint32)) << 12) | (s.astype(np.uint32) << 24))
    return np.concatenate(out)


@pytest.mark.parametrize("name", GOLDEN_TEXTURES)
def test_decompress_dxt_reference_files(device, name):
    g = load_golden(name)
    assert gst_amd.DecompressDXT(g["gtc"].tobytes()) == g["dxt"].tobytes()


@pytest.mark.parametrize("name", ["ref_test1", "ref_c0"])
def test_each_stage_matches_oracle_and_reference(device, name):
    g = load_golden(name)
    gtc = g["gtc"].tobytes()
    hdr = gst_amd.ParseHeader(gtc)
    nv, P = hdr.num_blocks, hdr.palette_bytes
    s = device.stream
    d_file = device.upload(gtc)
    base = d_file.data_ptr()
    d_tables = device.empty(4 * 2048 * 4 + 16)
    d_err = device.upload(b"\0\0\0\0")
    d_sym = device.empty(7 * nv + P)
    d_gs = device.empty(4 * (nv // 8192))
    d_idx = device.empty(4 * nv)
    d_planes = device.empty(6 * nv)
    d_out = device.empty(8 * nv)

    gst_amd.StageBuildTable(base + 28, d_tables.data_ptr(), d_err.data_ptr(), s)
    tab = np.frombuffer(device.download(d_tables), "<u4")
    np.testing.assert_array_equal(tab[:8192], packed_tables(gtc))
    assert list(tab[8192:8196]) == [1, 1, 1, 1]

    gst_amd.StageAnsDecode(hdr, base + 28 + 2048, d_tables.data_ptr(), d_sym.data_ptr(),
                           d_gs.data_ptr(), d_err.data_ptr(), s)
    sym = np.frombuffer(device.download(d_sym), np.uint8)
    np.testing.assert_array_equal(sym, g["symbols"])
    idx_syms = g["symbols"][6 * nv + P:].astype(np.int64) - 128
    gs = np.frombuffer(device.download(d_gs), "<i4")
    np.testing.assert_array_equal(gs, idx_syms.reshape(-1, 8192).sum(1))

    gst_amd.StageDecodeIndices(hdr, d_sym.data_ptr(), d_gs.data_ptr(), d_idx.data_ptr(), s)
    idx = np.frombuffer(device.download(d_idx), "<i4")
    np.testing.assert_array_equal(idx, go.decode_indices(g["symbols"][6 * nv + P:]))

    gst_amd.StageInvWavelet(hdr, d_sym.data_ptr(), d_planes.data_ptr(), s)
    planes = np.frombuffer(device.download(d_planes), np.int8)
    np.testing.assert_array_equal(planes, g["planes"])

    gst_amd.StageAssembleDXT(hdr, d_sym.data_ptr(), d_planes.data_ptr(), d_idx.data_ptr(),
                             d_out.data_ptr(), d_err.data_ptr(), s)
    assert device.download(d_out) == g["dxt"].tobytes()
    assert device.download(d_err) == b"\0\0\0\0"


def test_load_compressed_dxt_from_staging_buffer(device):
    g = load_golden("ref_test1")
    staging, (hdr,) = gst_amd.PackStaging([g["gtc"].tobytes()])
    d_cmp = device.upload(staging)
    d_out = device.empty(hdr.dxt_bytes)
    ev = gst_amd.LoadCompressedDXT(hdr, device.stream, d_cmp.data_ptr(), d_out.data_ptr(), want_event=True)
    assert ev
    assert device.download(d_out) == g["dxt"].tobytes()
    assert gst_amd.ReadErrorBits(device.stream) == 0


def test_load_compressed_dxts_uneven_sizes(device):
    names = ["ref_wide", "ref_small", "ref_c0", "ref_test1"]
    files = [load_golden(n)["gtc"].tobytes() for n in names]
    staging, hdrs = gst_amd.PackStaging(files)
    d_cmp = device.upload(staging)
    total = sum(h.dxt_bytes for h in hdrs)
    d_out = device.empty(total)
    gst_amd.LoadCompressedDXTs(hdrs, device.stream, d_cmp.data_ptr(), d_out.data_ptr())
    out = device.download(d_out)
    assert out == b"".join(load_golden(n)["dxt"].tobytes() for n in names)


def test_load_compressed_dxts_same_size_pages(device):
    """The reference's batched caller pattern: pages of equal-size images
    (demo/photos_sf.cpp:688-830)."""
    specs = [(100 + i, 512, 512) for i in range(16)]
    files, dxts = encode.make_textures(specs)
    staging, hdrs = gst_amd.PackStaging(files)
    d_cmp = device.upload(staging)
    d_out = device.empty(sum(h.dxt_bytes for h in hdrs))
    gst_amd.LoadCompressedDXTs(hdrs, device.stream, d_cmp.data_ptr(), d_out.data_ptr())
    assert device.download(d_out) == b"".join(dxts)


@pytest.mark.parametrize("name", ["ref_c0", "ref_small", "mixed"])
def test_load_rgb_matches_oracle(device, name):
    """LoadRGB against the oracle, and against the reference's own CPU decompression
    (ref_rgb.npz) for ref_small and for a file with many blocks in three-colour order."""
    import hashlib
    rgbs = load_golden("ref_rgb")
    gtc = (rgbs["mixed_gtc"] if name == "mixed" else load_golden(name)["gtc"]).tobytes()
    staging, (hdr,) = gst_amd.PackStaging([gtc])
    d_cmp = device.upload(staging)
    d_out = device.empty(hdr.width * hdr.height * 3)
    gst_amd.LoadRGB(hdr, device.stream, d_cmp.data_ptr(), d_out.data_ptr())
    rc, expected = go.decode_rgb(gtc)
    assert rc == go.OK
    got = device.download(d_out)
    assert got == expected
    if name == "ref_small":
        assert got == rgbs["small_rgb"].tobytes()
    if name == "mixed":
        assert hashlib.sha256(got).digest() == rgbs["mixed_rgb_sha256"].tobytes()


def test_load_rgbs_uneven_batch_matches_oracle(device):
    """LoadRGBs (codec/decoder.cpp:552-556) over images of uneven size in one staging
    buffer: image i's 3*w*h RGB8 bytes follow image i-1's, each equal to the oracle and,
    for ref_small, to the reference's own CPU decompression."""
    rgbs = load_golden("ref_rgb")
    files = [load_golden("ref_wide")["gtc"].tobytes(), load_golden("ref_small")["gtc"].tobytes(),
             rgbs["mixed_gtc"].tobytes(), load_golden("ref_c0")["gtc"].tobytes()]
    staging, hdrs = gst_amd.PackStaging(files)
    assert len({(h.width, h.height) for h in hdrs}) > 1
    d_cmp = device.upload(staging)
    sizes = [3 * h.width * h.height for h in hdrs]
    d_out = device.empty(sum(sizes) + 4096)  # a poisoned tail: nothing may be written past the images
    gst_amd.LoadRGBs(hdrs, device.stream, d_cmp.data_ptr(), d_out.data_ptr())
    got = device.download(d_out)
    off = 0
    for f, n in zip(files, sizes):
        rc, expected = go.decode_rgb(f, nthreads=8)
        assert rc == go.OK
        assert got[off:off + n] == expected
        off += n
    assert got[off:] == bytes([0xA5]) * 4096
    small = sizes[0]
    assert got[small:small + sizes[1]] == rgbs["small_rgb"].tobytes()
    assert gst_amd.ReadErrorBits(device.stream) == 0


@pytest.mark.parametrize("name", ["ref_c0", "ref_small"])
def test_stage_assemble_rgb_matches_oracle(device, name):
    """The per-stage RGB8 assembly (gst_stage_assemble_rgb, codec/assemble.cl:83-129)
    fed by the per-stage rANS, index and wavelet launches, into a poisoned output."""
    gtc = load_golden(name)["gtc"].tobytes()
    hdr = gst_amd.ParseHeader(gtc)
    nv, P = hdr.num_blocks, hdr.palette_bytes
    s = device.stream
    d_file = device.upload(gtc)
    base = d_file.data_ptr()
    d_tables = device.empty(4 * 2048 * 4 + 16)
    d_err = device.upload(b"\0\0\0\0")
    d_sym = device.empty(7 * nv + P)
    d_gs = device.empty(4 * (nv // 8192))
    d_idx = device.empty(4 * nv)
    d_planes = device.empty(6 * nv)
    d_out = device.empty(3 * hdr.width * hdr.height)
    gst_amd.StageBuildTable(base + 28, d_tables.data_ptr(), d_err.data_ptr(), s)
    gst_amd.StageAnsDecode(hdr, base + 28 + 2048, d_tables.data_ptr(), d_sym.data_ptr(), d_gs.data_ptr(),
                           d_err.data_ptr(), s)
    gst_amd.StageDecodeIndices(hdr, d_sym.data_ptr(), d_gs.data_ptr(), d_idx.data_ptr(), s)
    gst_amd.StageInvWavelet(hdr, d_sym.data_ptr(), d_planes.data_ptr(), s)
    gst_amd.StageAssembleRGB(hdr, d_sym.data_ptr(), d_planes.data_ptr(), d_idx.data_ptr(), d_out.data_ptr(),
                             d_err.data_ptr(), s)
    rc, expected = go.decode_rgb(gtc)
    assert rc == go.OK
    assert device.download(d_out) == expected
    if name == "ref_small":
        assert device.download(d_out) == load_golden("ref_rgb")["small_rgb"].tobytes()
    assert device.download(d_err) == b"\0\0\0\0"


@pytest.mark.parametrize("w,h,seed", [(512, 256, 21), (768, 512, 22), (1024, 1024, 23), (2048, 1024, 24),
                                      (2048, 2048, 25), (4096, 4096, 1)])
def test_encoder_textures_match_oracle(device, w, h, seed):
    gtc, dxt, _ = encode.make_texture(seed, w, h)
    rc, expected, _ = go.decode_dxt(gtc, nthreads=8)
    assert rc == go.OK and expected == dxt
    staging, (hdr,) = gst_amd.PackStaging([gtc])
    d_cmp = device.upload(staging)
    d_out = device.empty(hdr.dxt_bytes)
    gst_amd.LoadCompressedDXT(hdr, device.stream, d_cmp.data_ptr(), d_out.data_ptr())
    assert device.download(d_out) == expected
    assert gst_amd.ReadErrorBits(device.stream) == 0


def test_preallocated_scratch_is_carved_in_call_order(device):
    names = ["ref_c0", "ref_small", "ref_wide"]
    files = [load_golden(n)["gtc"].tobytes() for n in names]
    hdrs = [gst_amd.ParseHeader(f) for f in files]
    gst_amd.PreallocateDecompressor(sum(gst_amd.RequiredScratchMem(h) for h in hdrs))
    try:
        outs = []
        for f, h in zip(files, hdrs):
            staging, _ = gst_amd.PackStaging([f])
            d_cmp = device.upload(staging)
            d_out = device.empty(h.dxt_bytes)
            gst_amd.LoadCompressedDXT(h, device.stream, d_cmp.data_ptr(), d_out.data_ptr())
            outs.append((d_cmp, d_out))
        for (d_cmp, d_out), n in zip(outs, names):
            assert device.download(d_out) == load_golden(n)["dxt"].tobytes()
    finally:
        gst_amd.FreeDecompressor()


def test_init_events_order_the_decode(device):
    torch = device.torch
    g = load_golden("ref_c0")
    staging, (hdr,) = gst_amd.PackStaging([g["gtc"].tobytes()])
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        d_cmp = device.upload(staging)
        ready = torch.cuda.Event()
        ready.record(side)
    d_out = device.empty(hdr.dxt_bytes)
    gst_amd.LoadCompressedDXT(hdr, device.stream, d_cmp.data_ptr(), d_out.data_ptr(),
                              init=[ready.cuda_event])
    assert device.download(d_out) == g["dxt"].tobytes()


# ---- corrupt input: errors are reported, reads stay inside the buffers --------------------

def _corrupt(gtc, fn):
    b = bytearray(gtc)
    fn(b)
    return bytes(b)


def test_bad_frequency_table_is_reported(device):
    gtc = load_golden("ref_c0")["gtc"].tobytes()
    bad = _corrupt(gtc, lambda b: struct.pack_into("<H", b, 28, struct.unpack_from("<H", b, 28)[0] + 3))
    with pytest.raises(gst_amd.GstError) as e:
        gst_amd.DecompressDXT(bad)
    assert e.value.code == -3


def test_bad_group_offsets_are_reported(device):
    gtc = load_golden("ref_c0")["gtc"].tobytes()
    y0 = 28 + 2048
    # End offset of Y group 0 past the stream; and one that truncates group 1.
    for value in (0x7FFFFFF0, struct.unpack_from("<I", gtc, y0)[0] + 64):
        bad = _corrupt(gtc, lambda b: struct.pack_into("<I", b, y0, value))
        with pytest.raises(gst_amd.GstError) as e:
            gst_amd.DecompressDXT(bad)
        assert e.value.code == -3


@pytest.mark.parametrize("end", [128, 512])
def test_group_pointer_running_below_the_stream_start_is_reported(device, end):
    """Y group 0 cut to `end` bytes (its words really span 2328): its lanes renormalise
    past the start of the stream, so the word pointer runs below 0.  The oracle reports
    the file as corrupt; so must the device, although the pointer wraps round."""
    gtc = load_golden("ref_c0")["gtc"].tobytes()
    y0 = 28 + 2048
    assert struct.unpack_from("<I", gtc, y0)[0] == 2328
    bad = _corrupt(gtc, lambda b: struct.pack_into("<I", b, y0, end))
    rc, _, _ = go.decode_dxt(bad, nthreads=4)
    assert rc == go.ERR_STREAM
    with pytest.raises(gst_amd.GstError) as e:
        gst_amd.DecompressDXT(bad)
    assert e.value.code == -3


@pytest.mark.parametrize("stream", ["palette", "indices"])
def test_garbled_renormalisation_words_stay_in_bounds(device, stream):
    """Noise in a stream's words decodes to other symbols; palette indices that leave
    the palette raise the device error bit instead of reading past it."""
    gtc = load_golden("ref_c0")["gtc"].tobytes()
    h = gst_amd.ParseHeader(gtc)
    if stream == "palette":
        start = 28 + 2048 + h.y_cmp_sz + h.chroma_cmp_sz + 4 * (h.palette_bytes // 8192)
    else:
        start = 28 + 2048 + h.y_cmp_sz + h.chroma_cmp_sz + h.palette_sz + 4 * (h.num_blocks // 8192)
    rng = np.random.default_rng(5)
    noise = rng.integers(0, 256, 200, dtype=np.uint8).tobytes()
    bad = _corrupt(gtc, lambda b: b.__setitem__(slice(start, start + 200), noise))
    try:
        out = gst_amd.DecompressDXT(bad)
        assert len(out) == h.dxt_bytes
    except gst_amd.GstError as e:
        assert e.code == -3


@pytest.mark.parametrize("name", ["ref_test1", "ref_wide"])
def test_cpp_codec_test_against_reference_blocks(device, name, tmp_path):
    """tests/cpp/codec_test.cpp: codec/test/codec_test.cpp:36-48 through the C++ API."""
    import subprocess
    from conftest import ROOT
    g = load_golden(name)
    (tmp_path / "f.gtc").write_bytes(g["gtc"].tobytes())
    (tmp_path / "f.dxt").write_bytes(g["dxt"].tobytes())
    exe = os.path.join(ROOT, "tests", "cpp", "codec_test")
    proc = subprocess.run([exe, str(tmp_path / "f.gtc"), str(tmp_path / "f.dxt")], capture_output=True,
                          text=True, timeout=60)
    assert proc.returncode == 0, proc.stdout + proc.stderr
