"""gst_load_compressed_dxt_noalloc: the single-image decode with the library's static
device scratch instead of any allocation, against gst_load_compressed_dxt."""
import numpy as np
import pytest

from conftest import POISON, load_golden

import gst_amd
from gst_amd import encode

pytestmark = pytest.mark.gpu


def _decode_both(device, gtc):
    """Decodes `gtc` with gst_load_compressed_dxt and with the no-allocation entry point;
    returns (reference blocks, blocks, the poisoned tail after them, error word bytes,
    the reference path's error bits)."""
    staging, (hdr,) = gst_amd.PackStaging([gtc])
    d_cmp = device.upload(staging)
    d_ref = device.empty(hdr.dxt_bytes)
    gst_amd.LoadCompressedDXT(hdr, device.stream, d_cmp.data_ptr(), d_ref.data_ptr())
    ref_bits = gst_amd.ReadErrorBits(device.stream)
    d_out = device.empty(hdr.dxt_bytes + 4096)  # a poisoned tail: nothing may be written past the image
    d_err = device.upload(b"\0\0\0\0")
    gst_amd.LoadCompressedDXTNoAlloc(hdr, device.stream, d_cmp.data_ptr(), d_out.data_ptr(), d_err.data_ptr())
    out = device.download(d_out)
    return device.download(d_ref), out[:hdr.dxt_bytes], out[hdr.dxt_bytes:], device.download(d_err), ref_bits


def test_noalloc_equals_load_compressed_dxt_on_two_calls(device):
    """Two calls in a row, the second with a smaller input: the static scratch keeps the
    first call's larger contents, which the second must not depend on."""
    (big,), (big_dxt,) = encode.make_textures([(911, 4096, 4096)])
    small = load_golden("ref_c0")
    for gtc, expected in ((big, big_dxt), (small["gtc"].tobytes(), small["dxt"].tobytes())):
        ref, got, tail, err, ref_bits = _decode_both(device, gtc)
        assert ref == expected and ref_bits == 0
        assert got == ref
        assert tail == bytes([POISON]) * 4096
        assert err == b"\0\0\0\0"


def test_noalloc_reports_device_errors_and_rejects_oversized_images(device):
    import struct
    gtc = load_golden("ref_c0")["gtc"].tobytes()
    bad = bytearray(gtc)
    struct.pack_into("<H", bad, 28, struct.unpack_from("<H", bad, 28)[0] + 3)  # frequencies no longer sum to 2048
    _, _, _, err, ref_bits = _decode_both(device, bytes(bad))
    assert np.frombuffer(err, "<u4")[0] == ref_bits and ref_bits & 0x2  # GST_DEVERR_FREQUENCIES
    hdr = gst_amd.ParseHeader(gtc)
    big = gst_amd.Header(8192, 8192, hdr.palette_bytes, 4 * 2048 * 132, 8 * 2048 * 132, hdr.palette_sz,
                         hdr.indices_sz * 32 * 8)
    d = device.empty(64)
    with pytest.raises(gst_amd.GstError) as e:
        gst_amd.LoadCompressedDXTNoAlloc(big, device.stream, d.data_ptr(), d.data_ptr(), d.data_ptr())
    assert e.value.code == -8  # GST_ERR_TOO_LARGE, refused before any launch
