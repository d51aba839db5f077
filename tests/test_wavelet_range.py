"""Range of the inverse 5/3 wavelet: the gfx950 kernel computes in packed 16-bit
halves, which equals the reference's int32 arithmetic while every value stays below
2^15.  The CPU tests bound the growth on the oracle with sign patterns chosen to
maximise single outputs; the GPU test runs those worst cases, constant tiles and
full-range random bytes through the wavelet stage and compares with the oracle."""
import numpy as np
import pytest

import gst_amd
from oracle import gst_oracle as go

LEVELS = (2, 4, 8, 16, 32)
# One level (row pass then column pass) adds at most 674 to the magnitude bound:
# a 1-D pass gives at most Lo + 1.5 Hi + 0.5 for low/high-band bounds Lo, Hi.
LEVEL_GROWTH = 674


def _impulse_responses():
    """resp[i, o]: output o for a unit impulse at input i (scaled; the map is linear
    up to truncation)."""
    amp = 1 << 12
    resp = np.zeros((1024, 1024))
    for i in range(1024):
        t = np.zeros(1024, np.int32)
        t[i] = amp
        for n in LEVELS:
            t = go.inv_wavelet_level(t, n)
        resp[i] = t / amp
    return resp


def worst_case_tiles(count=8):
    """Input tiles (int8 values) whose signs follow the impulse responses of the
    outputs with the largest l1 norm, both polarities."""
    resp = _impulse_responses()
    l1 = np.abs(resp).sum(0)
    tiles = []
    for o in np.argsort(-l1)[:count]:
        for sgn in (1, -1):
            tiles.append(np.where(sgn * resp[:, o] > 0, 127, -128).astype(np.int32))
    return tiles, l1.max()


def test_impulse_response_norm_fits_the_bound():
    _, l1_max = worst_case_tiles(1)
    assert l1_max * 128 < 128 + 5 * LEVEL_GROWTH


def test_adversarial_tiles_stay_within_the_level_bound():
    tiles, _ = worst_case_tiles()
    for x in tiles + [np.full(1024, -128, np.int32), np.full(1024, 127, np.int32)]:
        t = x.copy()
        for k, n in enumerate(LEVELS, start=1):
            t = go.inv_wavelet_level(t, n)
            assert np.abs(t).max() <= 128 + k * LEVEL_GROWTH
        # Sums formed inside a level stay below 2 * bound + 2 < 2^15.
        assert 2 * np.abs(t).max() + 2 < 1 << 15


@pytest.mark.gpu
def test_wavelet_stage_on_extreme_tiles(device):
    # 512x256: nv = 8192 blocks, 8 tiles per plane, 48 tiles in the 6 planes.
    hdr = gst_amd.Header(512, 256, 8192, 4 * 4 + 4 * 128, 8 * 4 + 8 * 128, 4 + 128, 4 + 128)
    nv, bx = hdr.num_blocks, 512 // 4
    worst, _ = worst_case_tiles()
    rng = np.random.default_rng(11)
    tiles = [(w + 128).astype(np.uint8) for w in worst]
    tiles += [np.zeros(1024, np.uint8), np.full(1024, 255, np.uint8)]
    tiles += [np.where(rng.random(1024) < 0.5, 0, 255).astype(np.uint8) for _ in range(6)]
    while len(tiles) < 48:
        tiles.append(rng.integers(0, 256, 1024, dtype=np.uint8))
    sym = np.zeros(7 * nv + hdr.palette_bytes, np.uint8)
    sym[:6 * nv] = np.concatenate(tiles)

    per_plane, tiles_x = nv // 1024, bx // 32
    expect = np.zeros((6, nv // bx, bx), np.int8)
    for i, t in enumerate(tiles):
        plane, k = divmod(i, per_plane)
        ty, tx = divmod(k, tiles_x)
        expect[plane, 32 * ty:32 * ty + 32, 32 * tx:32 * tx + 32] = go.inv_wavelet_tile(t).reshape(32, 32)

    d_sym = device.upload(sym.tobytes())
    d_planes = device.empty(6 * nv)
    gst_amd.StageInvWavelet(hdr, d_sym.data_ptr(), d_planes.data_ptr(), device.stream)
    got = np.frombuffer(device.download(d_planes), np.int8).reshape(expect.shape)
    np.testing.assert_array_equal(got, expect)
