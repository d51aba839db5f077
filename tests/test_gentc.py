"""The reference encoder restated (gste_compress_dxt, include/gst_encode.h): the .gtc
file and the DXT1 blocks must be gentenc's, byte for byte.

CPU: the committed fixtures written by the compiled reference (its test image
codec/test/test1.png, three synthetic textures, and one texture encoded from given
blocks); where the reference build oracle/_ref/ref_driver is present, more inputs
against it live (other sizes and noise levels, white noise, a flat image, random
blocks, wavelet coefficients that wrap, and an input the reference refuses).
GPU: files written by the restated encoder decode on the device to its blocks."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

from gst_amd import encode
from oracle import gst_oracle as go

REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def fixture_input(g):
    """The encoder input of a fixture: its pixels, size and starting blocks."""
    if "rgb" in g:
        return g["rgb"], 512, 512, None
    meta = json.loads(g["meta"].tobytes())
    w, h = meta["width"], meta["height"]
    rgb = encode.generate_rgb(meta["seed"], w, h, meta["noise_milli"])
    blocks = None
    if meta.get("blocks") == "gste_compress_dxt1":
        blocks = np.empty(w * h // 2, np.uint8)
        encode.lib().gste_compress_dxt1(rgb.ctypes.data, w, h, blocks.ctypes.data)
        blocks = blocks.tobytes()
    return rgb, w, h, blocks


@pytest.mark.parametrize("name", ["ref_test1", "ref_c0", "ref_small", "ref_wide", "ref_given"])
def test_reference_fixtures_byte_identical(name):
    g = load_golden(name)
    rgb, w, h, blocks = fixture_input(g)
    gtc, dxt = encode.compress_dxt(rgb, w, h, dxt=blocks)
    assert dxt == g["dxt"].tobytes()
    assert gtc == g["gtc"].tobytes()


def test_thread_pool_matches_single_calls_and_fixtures():
    """gste_make_textures_gentenc (the benchmark's corpus path): largest first over a
    thread pool, every file as gste_compress_dxt writes it, so the fixtures' textures come
    out as the reference's files."""
    specs = [(3, 512, 256), (0, 512, 512), (7, 1024, 512), (41, 256, 512)]
    files, dxts = encode.make_textures(specs, nthreads=3, encoder="gentenc")
    for name, i in (("ref_small", 0), ("ref_c0", 1), ("ref_wide", 2)):
        g = load_golden(name)
        assert files[i] == g["gtc"].tobytes() and dxts[i] == g["dxt"].tobytes(), name
    gtc, dxt = encode.compress_dxt(encode.generate_rgb(41, 256, 512))
    assert files[3] == gtc and dxts[3] == dxt
    with pytest.raises(ValueError):
        encode.make_textures(specs[:1], encoder="other")


def test_oracle_decodes_to_the_encoder_blocks():
    for seed, w, h in ((41, 512, 256), (42, 256, 1024)):
        gtc, dxt = encode.compress_dxt(encode.generate_rgb(seed, w, h, 500))
        rc, blocks, _ = go.decode_dxt(gtc, nthreads=4)
        assert rc == go.OK and blocks == dxt


def checkerboard(w, h, word=0):
    """Blocks whose two endpoints alternate between Cg = +63 and Cg = -63 in a
    checkerboard, every index word `word` (0: all pixels take the first endpoint;
    0x55555555: the second, equal to it), with pixels of exactly that colour: no block
    is reused, and the Cg plane's finest diagonal detail reaches 252, outside a byte."""
    bw, bh = w // 4, h // 4
    up = (63 << 5)                      # r 0, g 63, b 0 -> Cg +63
    down = (31 << 11) | 31              # r 31, g 0, b 31 -> Cg -63
    yy, xx = np.mgrid[0:bh, 0:bw]
    ep = np.where((xx + yy) % 2 == 0, up, down).astype(np.uint16)
    blocks = np.zeros((bh, bw, 4), np.uint16)
    blocks[..., 0] = ep
    blocks[..., 1] = ep                 # equal endpoints: 3-colour order, index 1 = ep too
    blocks[..., 2] = word & 0xFFFF
    blocks[..., 3] = word >> 16
    rgb565 = ep.repeat(4, 0).repeat(4, 1)
    r = ((rgb565 >> 11) & 31).astype(np.uint16)
    g = ((rgb565 >> 5) & 63).astype(np.uint16)
    b = (rgb565 & 31).astype(np.uint16)
    rgb = np.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], -1).astype(np.uint8)
    return rgb, blocks.view(np.uint8).tobytes()


def test_wavelet_overflow_wraps_as_in_the_reference():
    """The coefficients wrap into the file; the blocks are kept, so the file decodes
    to other endpoints than the blocks reported (the reference's behaviour)."""
    rgb, blocks = checkerboard(512, 256, 0x55555555)
    gtc, dxt = encode.compress_dxt(rgb, dxt=blocks)
    assert dxt == blocks
    rc, decoded, _ = go.decode_dxt(gtc, nthreads=4)
    assert rc == go.OK and decoded != dxt
    # Index words and palette are untouched by the wrap: only endpoints differ.
    a = np.frombuffer(decoded, "<u4").reshape(-1, 2)
    b = np.frombuffer(dxt, "<u4").reshape(-1, 2)
    assert np.array_equal(a[:, 1], b[:, 1]) and not np.array_equal(a[:, 0], b[:, 0])


def test_stream_of_zero_bytes_is_uncodable():
    rgb, blocks = checkerboard(512, 256, 0)  # palette of zero words only
    with pytest.raises(encode.EncodeError) as e:
        encode.compress_dxt(rgb, dxt=blocks)
    assert e.value.code == encode.GSTE_ERR_UNCODABLE


def test_argument_and_capacity_errors():
    lib = encode.lib()
    need = ctypes.c_size_t(123)
    rgb = np.zeros(128 * 128 * 3, np.uint8)
    # 128 x 128 has 1024 blocks: the index stream would not fill one 8192-symbol group.
    assert lib.gste_compress_dxt(rgb.ctypes.data, 128, 128, None, None, 0, ctypes.byref(need), None) == \
        encode.GSTE_ERR_ARGUMENT
    assert need.value == 0
    assert lib.gste_compress_dxt(None, 512, 256, None, None, 0, ctypes.byref(need), None) == encode.GSTE_ERR_ARGUMENT
    rgb = encode.generate_rgb(5, 512, 256)
    assert lib.gste_compress_dxt(rgb.ctypes.data, 500, 256, None, None, 0, ctypes.byref(need), None) == \
        encode.GSTE_ERR_ARGUMENT
    small = np.empty(64, np.uint8)
    assert lib.gste_compress_dxt(rgb.ctypes.data, 512, 256, None, small.ctypes.data, 64, ctypes.byref(need),
                                 None) == encode.GSTE_ERR_CAPACITY
    gtc, _ = encode.compress_dxt(rgb)
    assert need.value == len(gtc)
    with pytest.raises(ValueError):
        encode.compress_dxt(rgb, dxt=b"\0" * 10)


def test_cpp_encoder_api(tmp_path):
    """tests/cpp/encoder_test.cpp: GenTC::CompressDXT through include/gst/encoder.hpp writes
    the reference's file and blocks for the seed-0 texture; its second pass, from those
    blocks, writes a file that decodes to the blocks it reports (and, where the reference
    build is present, the file the reference writes from the same blocks)."""
    exe = os.path.join(ROOT, "tests", "cpp", "encoder_test")
    outs = [tmp_path / n for n in ("a.gtc", "a.dxt", "b.gtc", "b.dxt")]
    proc = subprocess.run([exe, *map(str, outs)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    g = load_golden("ref_c0")
    assert outs[0].read_bytes() == g["gtc"].tobytes()
    assert outs[1].read_bytes() == g["dxt"].tobytes()
    rc, blocks, _ = go.decode_dxt(outs[2].read_bytes(), nthreads=4)
    assert rc == go.OK and blocks == outs[3].read_bytes()
    if os.path.exists(REF_DRIVER):
        rgb = encode.generate_rgb(0, 512, 512, 300)
        want = _reference(tmp_path, rgb, 512, 512, outs[1].read_bytes())
        assert want == (outs[2].read_bytes(), outs[3].read_bytes())


# ---- against the reference build, where it is present --------------------------

def _reference(tmp_path, rgb, w, h, blocks=None):
    src = tmp_path / "in.rgb"
    src.write_bytes(np.ascontiguousarray(rgb).tobytes())
    gtc, dxt = tmp_path / "out.gtc", tmp_path / "out.dxt"
    if blocks is None:
        args = ["encode_rgb", src, w, h, gtc, dxt]
    else:
        (tmp_path / "in.dxt").write_bytes(blocks)
        args = ["encode_rgb_dxt", src, w, h, tmp_path / "in.dxt", gtc, dxt]
    p = subprocess.run([REF_DRIVER, *map(str, args)], capture_output=True, timeout=120)
    if p.returncode != 0:
        return None
    return gtc.read_bytes(), dxt.read_bytes()


def _live_cases():
    rng = np.random.default_rng(5)
    cases = [("seed%d_%dx%d_noise%d" % (s, w, h, n), (s, w, h, n), None)
             for s, w, h, n in ((11, 512, 256, 0), (12, 256, 512, 100), (13, 512, 512, 600),
                                (14, 1024, 128, 1000), (15, 128, 1024, 50))]
    cases.append(("random_blocks", (21, 512, 256, 300), rng.integers(0, 256, 512 * 128, dtype=np.uint8).tobytes()))
    cases.append(("white_noise", rng.integers(0, 256, (512, 256, 3), dtype=np.uint8), None))
    cases.append(("flat", np.full((256, 512, 3), 77, np.uint8), None))
    rgb, blocks = checkerboard(512, 256, 0x55555555)
    cases.append(("wavelet_wrap", rgb, blocks))
    return cases


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="reference build oracle/_ref/ref_driver not present")
@pytest.mark.parametrize("name,source,blocks", _live_cases(), ids=[c[0] for c in _live_cases()])
def test_against_reference_build(tmp_path, name, source, blocks):
    if isinstance(source, tuple):
        seed, w, h, noise = source
        rgb = encode.generate_rgb(seed, w, h, noise)
    else:
        rgb = source
        h, w = rgb.shape[:2]
    want = _reference(tmp_path, rgb, w, h, blocks)
    assert want is not None, "the reference encoder failed"
    got = encode.compress_dxt(rgb, w, h, dxt=blocks)
    assert got[1] == want[1]
    assert got[0] == want[0]


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="reference build oracle/_ref/ref_driver not present")
def test_reference_build_also_refuses_a_zero_stream(tmp_path):
    rgb, blocks = checkerboard(512, 256, 0)
    assert _reference(tmp_path, rgb, 512, 256, blocks) is None


# ---- GPU ------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("seed,w,h", [(0, 512, 512), (31, 512, 256), (32, 1024, 1024), (33, 2048, 2048)])
def test_encoder_files_decode_on_gpu(device, seed, w, h):
    """Files the restated encoder writes on this host decode on the device to its
    blocks; the 512x512 seed-0 file is the committed fixture written by the reference."""
    import gst_amd
    gtc, dxt = encode.compress_dxt(encode.generate_rgb(seed, w, h, 300))
    if (seed, w, h) == (0, 512, 512):
        assert gtc == load_golden("ref_c0")["gtc"].tobytes()
    staging, (hdr,) = gst_amd.PackStaging([gtc])
    d_cmp = device.upload(staging)
    d_out = device.empty(hdr.dxt_bytes)
    gst_amd.LoadCompressedDXT(hdr, device.stream, d_cmp.data_ptr(), d_out.data_ptr())
    assert device.download(d_out) == dxt
    assert gst_amd.ReadErrorBits(device.stream) == 0
