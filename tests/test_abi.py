"""The C-ABI libraries load and export what include/*.h declares; host-side logic
(header rules, staging layout, scratch sizing) without touching a GPU."""
import ctypes
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden

import gst_amd
from gst_amd import encode
from oracle import gst_oracle as go


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gst|gste)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("gst_decoder.h", "libgst_amd.so"), ("gst_serve.h", "libgst_amd.so"),
                                        ("gst_ans.h", "libgst_amd.so"), ("gst_encode.h", "libgst_encode.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = declared_functions(header)
    assert len(names) >= 5
    handle = ctypes.CDLL(os.path.join(ROOT, "gst_amd", lib))
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, missing


def test_header_parse_matches_oracle_rules():
    g = load_golden("ref_c0")["gtc"].tobytes()
    h = gst_amd.ParseHeader(g)
    assert (h.width, h.height) == (512, 512)
    assert h.stream_bytes + 28 + 2048 == len(g)
    for field, value in ((0, 500), (1, 0), (2, 100), (3, 6)):
        bad = bytearray(g)
        struct.pack_into("<I", bad, 4 * field, value)
        with pytest.raises(gst_amd.GstError) as e:
            gst_amd.ParseHeader(bytes(bad))
        assert e.value.code == -2
        assert go.parse_header(bytes(bad))[0] == go.ERR_HEADER


def test_required_scratch_mem_is_reference_formula():
    h = gst_amd.ParseHeader(load_golden("ref_c0")["gtc"].tobytes())
    assert gst_amd.RequiredScratchMem(h) == 4 * 2048 * 6 + 17 * 512 * 512 // 16 + h.palette_bytes


def test_staging_layout_matches_reference():
    """codec/decoder.cpp:109-209 and demo/photos_sf.cpp:749-795: offsets block padded
    to 512 B, 2048 B of frequencies per image, then the streams back to back."""
    files = [load_golden(n)["gtc"].tobytes() for n in ("ref_c0", "ref_small", "ref_wide")]
    buf, hdrs = gst_amd.PackStaging(files)
    total, soff = gst_amd.StagingSize(hdrs)
    assert len(buf) == total and soff == 512 + 3 * 2048
    off = np.frombuffer(buf[:96], "<u4")
    out_off, in_off = off[:12], off[12:24]
    o = i = 0
    pos = soff
    for k, (f, h) in enumerate(zip(files, hdrs)):
        nv = h.width * h.height // 16
        assert list(out_off[4 * k:4 * k + 4]) == [o, o + 2 * nv, o + 6 * nv, o + 6 * nv + h.palette_bytes]
        o += 7 * nv + h.palette_bytes
        assert list(in_off[4 * k:4 * k + 4]) == [i, i + h.y_cmp_sz, i + h.y_cmp_sz + h.chroma_cmp_sz,
                                                  i + h.y_cmp_sz + h.chroma_cmp_sz + h.palette_sz]
        i += h.stream_bytes
        assert buf[512 + 2048 * k:512 + 2048 * (k + 1)] == f[28:28 + 2048]
        assert buf[pos:pos + h.stream_bytes] == f[28 + 2048:]
        pos += h.stream_bytes
    assert not any(buf[96:512])


def test_single_image_staging_is_file_at_offset_512():
    f = load_golden("ref_c0")["gtc"].tobytes()
    buf, _ = gst_amd.PackStaging([f])
    assert buf[512:] == f[28:]


# ---- the build's own encoder (input tooling) ------------------------------------------

@pytest.mark.parametrize("w,h,seed", [(512, 256, 1), (512, 512, 0), (1024, 512, 2), (768, 512, 3),
                                      (512, 1024, 4), (1024, 1024, 5)])
def test_encoder_files_decode_to_encoder_blocks(w, h, seed):
    gtc, dxt, st = encode.make_texture(seed, w, h)
    assert st.wavelet_overflows == 0
    rc, out, _ = go.decode_dxt(gtc, nthreads=4)
    assert rc == go.OK
    assert out == dxt


def test_encoder_is_deterministic():
    a = encode.make_texture(11, 512, 512)[0]
    b = encode.make_texture(11, 512, 512)[0]
    assert a == b
    c = encode.make_textures([(11, 512, 512), (12, 512, 256)], nthreads=2)[0]
    assert c[0] == a


def test_generator_matches_committed_fixture_pixels():
    """ref_c0's RGB is the generator's seed-0 512x512 texture; the reference encoder's
    blocks must be a DXT1 encoding of it within stb_dxt's error (loose bound)."""
    rgb = encode.generate_rgb(0, 512, 512).astype(np.int32)
    dxt = load_golden("ref_c0")["dxt"]
    rc, dec_rgb = go.decode_rgb(load_golden("ref_c0")["gtc"].tobytes())
    assert rc == go.OK
    img = np.frombuffer(dec_rgb, np.uint8).reshape(512, 512, 3).astype(np.int32)
    assert len(dxt) == 512 * 512 // 2
    assert np.abs(img - rgb).mean() < 12


def test_rans_units_cover_every_group_once():
    """tests/cpp/ans_units_test.cpp (host code): the rANS work units of an image, with and
    without short stream remainders paired in one work-group, cover each group of each
    stream exactly once for every group-count combination up to 20 per stream."""
    exe = os.path.join(ROOT, "tests", "cpp", "ans_units_test")
    proc = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "combinations checked" in proc.stdout
