"""The stand-alone interleaved rANS decoder (include/gst_ans.h, the counterpart of the
reference's ans::ocl::OpenCLDecoder, ans/ans_ocl.h:26-72).

CPU: the frequency normalisation against the reference's GenerateHistogram outputs
(tests/golden/ref_hist.json, written by the compiled reference), and the test-side rANS
encoder below against the reference's own encoded groups (ref_ans_groups.npz, from
ans::EncodeInterleaved).  GPU: the five tests of ans/ans_ocl_test.cpp restated, the
reference's encoded groups decoded through the API, and corrupt input."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden, load_hist

import gst_amd

M = 2048
L = 1 << 15


def rans_encode(norm, streams):
    """rANS encoder of ans/encode.cpp:49-69 (b = 2^16, k = 2^4, M = 2048), the streams
    interleaved symbol by symbol in stream order (ans/ans_ocl_test.cpp:237-248).
    `norm`: normalised frequencies; `streams`: lists of symbols.  Returns the 16-bit
    words (little-endian bytes, emission order) and the final states."""
    cum = np.concatenate([[0], np.cumsum(norm)[:-1]]).astype(np.int64)
    states = [L] * len(streams)
    words = []
    for i in range(len(streams[0])):
        for j, syms in enumerate(streams):
            s = int(syms[i])
            f = int(norm[s])
            x = states[j]
            while x >= (f << 20):
                words.append(x & 0xFFFF)
                x >>= 16
            states[j] = (x // f) * M + int(cum[s]) + x % f
    return np.array(words, dtype="<u2").tobytes(), states


def draw_symbols(counts, n, rng):
    """Symbols with probability proportional to `counts` (ans/ans_ocl_test.cpp:43-62)."""
    p = np.asarray(counts, np.float64)
    return rng.choice(len(counts), size=n, p=p / p.sum()).astype(np.uint8)


def closed_form_table(norm):
    sym, freq, cum = [], [], []
    s = 0
    for i, f in enumerate(norm):
        sym += [i] * f
        freq += [f] * f
        cum += [s] * f
        s += f
    assert s == M
    return np.array(sym, np.uint8), np.array(freq, np.uint16), np.array(cum, np.uint16)


def golden_hist_cases():
    h = load_hist()
    return [v for k, v in h.items() if k != "random"] + h["random"]


# ---- CPU -------------------------------------------------------------------------

def test_normalization_matches_reference_histograms():
    cases = golden_hist_cases()
    assert len(cases) >= 50
    assert any(max(c["counts"]) >= 1 << 21 for c in cases)  # the wrapping count * M product
    for c in cases:
        assert gst_amd.NormalizeFrequencies(c["counts"], c["M"]) == c["normalized"], c["counts"][:8]


def test_normalization_rejects_bad_input():
    with pytest.raises(gst_amd.GstError):
        gst_amd.NormalizeFrequencies([0, 0, 0])
    with pytest.raises(gst_amd.GstError):
        gst_amd.NormalizeFrequencies([1, 2, 3], 0)


def test_test_encoder_reproduces_reference_groups():
    """The encoder above writes the reference's bytes for its own groups (32 streams):
    words in emission order, then the 32 final states."""
    g = load_golden("ref_ans_groups")
    seeds = sorted(int(k.split("_")[1]) for k in g if k.startswith("group_"))
    assert seeds
    for seed in seeds:
        norm = g["freq_%d" % seed].astype(np.int64)
        assert norm.sum() == M
        syms = g["symbols_%d" % seed].reshape(32, 256)
        words, states = rans_encode(norm, list(syms))
        expect = np.frombuffer(words, np.uint8).tobytes() + np.array(states, "<u4").tobytes()
        assert expect == g["group_%d" % seed].tobytes(), seed


# ---- GPU -------------------------------------------------------------------------

@pytest.mark.gpu
def test_initialization(device):
    """ans/ans_ocl_test.cpp:64-105."""
    h = load_hist()["ocl_init"]
    dec = gst_amd.AnsDecoder(h["counts"], 1)
    for got, want in zip(dec.table(), closed_form_table(h["normalized"])):
        np.testing.assert_array_equal(got, want)
    dec.close()


@pytest.mark.gpu
def test_table_rebuilding(device):
    """ans/ans_ocl_test.cpp:107-155."""
    h = load_hist()
    dec = gst_amd.AnsDecoder([3, 2, 1, 4, 3, 406], 1)
    dec.rebuild_table(h["ocl_rebuild"]["counts"])
    for got, want in zip(dec.table(), closed_form_table(h["ocl_rebuild"]["normalized"])):
        np.testing.assert_array_equal(got, want)
    dec.close()


@pytest.mark.gpu
def test_decode_single_stream(device):
    """ans/ans_ocl_test.cpp:157-218, decoded twice by two decoders."""
    h = load_hist()["ocl_single"]
    rng = np.random.default_rng(0)
    syms = draw_symbols(h["counts"], 256, rng)
    words, states = rans_encode(np.array(h["normalized"]), [syms])
    for _ in range(2):
        dec = gst_amd.AnsDecoder(h["counts"], 1)
        out = dec.decode(states, words)
        assert out[0] == syms.tobytes()
        dec.close()


@pytest.mark.gpu
def test_decode_interleaved_streams(device):
    """ans/ans_ocl_test.cpp:220-275: 24 streams interleaved in one buffer."""
    h = load_hist()["ocl_interleaved"]
    rng = np.random.default_rng(1)
    streams = [draw_symbols(h["counts"], 256, rng) for _ in range(24)]
    words, states = rans_encode(np.array(h["normalized"]), streams)
    dec = gst_amd.AnsDecoder(h["counts"], 24)
    out = dec.decode(states, words)
    assert len(out) == 24
    for got, want in zip(out, streams):
        assert got == want.tobytes()
    dec.close()


@pytest.mark.gpu
@pytest.mark.parametrize("num_interleaved,n_data", [(1, 2), (32, 5), (7, 3)])
def test_decode_multiple_interleaved_streams(device, num_interleaved, n_data):
    """ans/ans_ocl_test.cpp:277-343 (1 x 2), and wider: independent buffers of
    num_interleaved streams each."""
    h = load_hist()["ocl_multiple"]
    rng = np.random.default_rng(2 + num_interleaved)
    norm = np.array(h["normalized"])
    all_streams, blobs, all_states = [], [], []
    for _ in range(n_data):
        streams = [draw_symbols(h["counts"], 256, rng) for _ in range(num_interleaved)]
        words, states = rans_encode(norm, streams)
        all_streams += streams
        blobs.append(words)
        all_states += states
    dec = gst_amd.AnsDecoder(h["counts"], num_interleaved)
    out = dec.decode(all_states, blobs)
    assert len(out) == len(all_streams)
    for got, want in zip(out, all_streams):
        assert got == want.tobytes()
    dec.close()


@pytest.mark.gpu
def test_reference_encoded_groups_decode(device):
    """The reference's own 32-stream groups (ans::EncodeInterleaved) through the API."""
    g = load_golden("ref_ans_groups")
    for seed in sorted(int(k.split("_")[1]) for k in g if k.startswith("group_")):
        norm = g["freq_%d" % seed].astype(np.uint32)
        raw = g["group_%d" % seed].tobytes()
        words, states = raw[:-128], list(np.frombuffer(raw[-128:], "<u4"))
        dec = gst_amd.AnsDecoder(list(norm), 32)  # already normalised: kept as is
        sym, freq, cum = dec.table()
        np.testing.assert_array_equal(freq, closed_form_table(list(norm))[1])
        out = b"".join(dec.decode(states, [words]))
        assert out == g["symbols_%d" % seed].tobytes(), seed
        dec.close()


@pytest.mark.gpu
def test_corrupt_stream_is_reported(device):
    h = load_hist()["ocl_interleaved"]
    rng = np.random.default_rng(9)
    streams = [draw_symbols(h["counts"], 256, rng) for _ in range(8)]
    words, states = rans_encode(np.array(h["normalized"]), streams)
    dec = gst_amd.AnsDecoder(h["counts"], 8)
    bad_states = list(states)
    bad_states[3] ^= 0x5A5A
    with pytest.raises(gst_amd.GstError):
        dec.decode(bad_states, words)
    with pytest.raises(gst_amd.GstError):  # too short: words missing
        dec.decode(states, words[:2])
    # the decoder still works afterwards
    assert [o for o in dec.decode(states, words)] == [s.tobytes() for s in streams]
    dec.close()


@pytest.mark.gpu
def test_bad_arguments(device):
    with pytest.raises(gst_amd.GstError):
        gst_amd.AnsDecoder([1, 2], 33)
    with pytest.raises(gst_amd.GstError):
        gst_amd.AnsDecoder([0, 0], 1)
    dec = gst_amd.AnsDecoder([1, 2], 4)
    with pytest.raises(gst_amd.GstError):  # 5 streams in one buffer > num_interleaved
        dec.decode([L] * 5, b"")
    with pytest.raises(gst_amd.GstError):  # 6 streams over 4 buffers
        dec.decode([L] * 6, [b"", b"", b"", b""])
    dec.close()


@pytest.mark.gpu
def test_cpp_ans_ocl_test(device):
    """tests/cpp/ans_ocl_test.cpp: ans/ans_ocl_test.cpp's five tests through the C++ API."""
    exe = os.path.join(ROOT, "tests", "cpp", "ans_ocl_test")
    proc = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert proc.stdout.count(" ok") == 6, proc.stdout
