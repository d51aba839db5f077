"""gst_upload_gather: pinned host ranges copied into device ranges by the gather kernel
(the server's ingest copy), checked byte for byte, bytes outside the ranges untouched."""
import numpy as np
import pytest

from conftest import POISON

import gst_amd

pytestmark = pytest.mark.gpu


def _ranges(rng, sizes, span):
    """Non-overlapping, 16-byte aligned (src, dst, size) triples inside `span` bytes."""
    out, src, dst = [], 0, 0
    for n in sizes:
        out.append((src, dst, n))
        src += (n + 15) // 16 * 16 + 16 * int(rng.integers(0, 4))
        dst += (n + 15) // 16 * 16 + 16 * int(rng.integers(0, 4))
    assert src <= span and dst <= span
    return out


def test_gather_copies_every_range_and_nothing_else(device):
    torch = device.torch
    rng = np.random.default_rng(7)
    span = 8 << 20
    host = torch.from_numpy(rng.integers(0, 256, span, dtype=np.uint8)).pin_memory()
    dev = device.empty(span)
    # Whole words, tails of 1-15 bytes, an empty range, one range past 1 MiB, and enough
    # ranges (about 160) that the copy takes three launches of 64.
    sizes = [1, 15, 16, 17, 0, 4095, 4096, (1 << 20) + 5, 65536 + 3]
    sizes += [int(x) for x in rng.integers(1, 20000, 150)]
    ranges = _ranges(rng, sizes, span)
    hb, db = host.data_ptr(), dev.data_ptr()
    gst_amd.UploadGather([hb + s for s, _, _ in ranges], [db + d for _, d, _ in ranges],
                         [n for _, _, n in ranges], device.stream)
    device.sync()
    h = host.numpy()
    got = dev.cpu().numpy()
    covered = np.zeros(span, bool)
    for s, d, n in ranges:
        assert np.array_equal(got[d:d + n], h[s:s + n]), (s, d, n)
        covered[d:d + n] = True
    assert np.all(got[~covered] == POISON)


def test_gather_rejects_what_the_device_cannot_read(device):
    torch = device.torch
    host = torch.zeros(4096, dtype=torch.uint8).pin_memory()
    dev = device.empty(4096)
    with pytest.raises(gst_amd.GstError):  # misaligned source
        gst_amd.UploadGather([host.data_ptr() + 8], [dev.data_ptr()], [64], device.stream)
    with pytest.raises(gst_amd.GstError):  # misaligned destination
        gst_amd.UploadGather([host.data_ptr()], [dev.data_ptr() + 4], [64], device.stream)
    big = torch.zeros(1 << 20, dtype=torch.uint8).pin_memory()
    room = device.empty(2 << 20)
    with pytest.raises(gst_amd.GstError):  # runs past the end of its pinned allocation
        gst_amd.UploadGather([big.data_ptr()], [room.data_ptr()], [(1 << 20) + 64], device.stream)
    pageable = np.zeros(4096 + 16, np.uint8)
    addr = (pageable.ctypes.data + 15) // 16 * 16
    with pytest.raises(gst_amd.GstError):  # not pinned: no device address
        gst_amd.UploadGather([addr], [dev.data_ptr()], [64], device.stream)
    gst_amd.UploadGather([], [], [], device.stream)  # nothing to do
    device.sync()
    assert np.all(dev.cpu().numpy() == POISON)
