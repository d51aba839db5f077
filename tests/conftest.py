"""Shared fixtures.  GPU tests carry @pytest.mark.gpu and run on an MI355X only."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_TEXTURES = ("ref_test1", "ref_c0", "ref_small", "ref_wide")
# Byte every GPU test output starts as (0xA5 blocks are valid DXT1, so a path that
# writes nothing yields plausible bytes that only a comparison catches).
POISON = 0xA5


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libgst_amd.so")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def load_hist():
    with open(os.path.join(GOLDEN, "ref_hist.json")) as f:
        return json.load(f)


class Device:
    """torch-backed device memory and the current stream, for C-ABI calls."""

    def __init__(self, torch):
        self.torch = torch
        self.dev = torch.device("cuda:0")
        # A created stream (graph capture is not possible on the null stream).
        self.side = torch.cuda.Stream(device=self.dev)
        torch.cuda.set_stream(self.side)

    @property
    def stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def empty(self, nbytes):
        """Device bytes filled with POISON: a launch that writes nothing, or writes
        short, cannot pass on what an earlier test left in a recycled block."""
        t = self.torch.empty(max(int(nbytes), 1), dtype=self.torch.uint8, device=self.dev)
        return self.poison(t)

    def poison(self, t):
        t.fill_(POISON)
        self.torch.cuda.current_stream().synchronize()
        return t

    def upload(self, data):
        host = self.torch.frombuffer(bytearray(bytes(data)), dtype=self.torch.uint8)
        return host.to(self.dev)

    def download(self, t, nbytes=None):
        self.torch.cuda.synchronize()
        arr = t.cpu().numpy()
        return arr[: nbytes if nbytes is not None else len(arr)].tobytes()

    def sync(self):
        self.torch.cuda.synchronize()


@pytest.fixture(scope="session")
def device():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gst_amd
    assert gst_amd.InitializeDecoder(0), "device 0 is not a usable gfx950"
    return Device(torch)
