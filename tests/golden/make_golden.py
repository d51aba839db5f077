#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from the reference's own code.

Runs only where the reference sources are present (/root/reference): it builds
oracle/_ref/ref_driver (oracle/Makefile target `ref`, the reference's encoder and
CPU stage functions compiled from their sources plus a small driver) and records
its outputs as data fixtures:

  ref_test1.npz     the reference's own test image codec/test/test1.png encoded by
                    the reference encoder (DXTImage + CompressDXT); `dxt` is
                    PhysicalBlocks(), the value codec/test/codec_test.cpp:36-48
                    requires of the decoder; `symbols` / `planes` are the reference
                    CPU rANS decoder's and CPU inverse wavelet's outputs.
  ref_c0.npz        the 512x512 seed-0 synthetic texture (noise 0.3) encoded the same way.
  ref_small.npz     a 512x256 texture (smallest legal size, one index group).
  ref_wide.npz      a 1024x512 texture (three index-scan stages).
                    (These two keep only `gtc` and `dxt`.)
  ref_given.npz     a 512x256 texture encoded from given blocks, DXTImage(w, h, rgb, dxt)
                    with this repository's fast DXT1 compressor's blocks.
  (ref_test1.npz also keeps `rgb`, the pixels stbi_load gives the encoder.)
  ref_ans_groups.npz  single-group rANS known answers from ans::EncodeInterleaved.
  ref_hist.json     ans::GenerateHistogram outputs used by the table known answers
                    of ans/ans_ocl_test.cpp:64-155.
  ref_rgb.npz       RGB8 known answers of the LoadRGB path (see make_rgb).

The synthetic RGB comes from this repository's generator (libgst_encode.so), so
the same pixels can be regenerated anywhere; the encoded bytes are the reference's.
"""
import ctypes
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def run(*args):
    subprocess.run([DRIVER, *[str(a) for a in args]], check=True)


def read(path):
    with open(path, "rb") as f:
        return f.read()


def encode_case(name, rgb_path=None, png=None, w=None, h=None, meta=None, tmp=None,
                stages=True, dxt_in=None):
    gtc, dxt = os.path.join(tmp, name + ".gtc"), os.path.join(tmp, name + ".dxt")
    sym, planes = os.path.join(tmp, name + ".sym"), os.path.join(tmp, name + ".planes")
    extra = {}
    if png:
        run("encode_png", png, gtc, dxt)
        rgb = os.path.join(tmp, name + ".rgb")
        run("png_rgb", png, rgb)
        extra["rgb"] = np.frombuffer(read(rgb), np.uint8)  # the encoder's input pixels
    elif dxt_in:
        run("encode_rgb_dxt", rgb_path, w, h, dxt_in, gtc, dxt)
    else:
        run("encode_rgb", rgb_path, w, h, gtc, dxt)
    arrays = dict(gtc=np.frombuffer(read(gtc), np.uint8), dxt=np.frombuffer(read(dxt), np.uint8),
                  meta=np.frombuffer(json.dumps(meta or {}).encode(), np.uint8), **extra)
    if stages:
        run("stages", gtc, sym, planes)
        arrays["symbols"] = np.frombuffer(read(sym), np.uint8)
        arrays["planes"] = np.frombuffer(read(planes), np.int8)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)


def synthetic_rgb(seed, w, h, noise_milli, path):
    lib = ctypes.CDLL(os.path.join(ROOT, "gst_amd", "libgst_encode.so"))
    buf = ctypes.create_string_buffer(w * h * 3)
    lib.gste_generate_rgb(ctypes.c_uint64(seed), w, h, noise_milli, buf)
    with open(path, "wb") as f:
        f.write(buf.raw)
    return buf


def fast_dxt1(rgb, w, h, path):
    """Blocks of this repository's fast DXT1 compressor (gste_compress_dxt1)."""
    lib = ctypes.CDLL(os.path.join(ROOT, "gst_amd", "libgst_encode.so"))
    buf = ctypes.create_string_buffer(w * h // 2)
    lib.gste_compress_dxt1(rgb, w, h, buf)
    with open(path, "wb") as f:
        f.write(buf.raw)


def main():
    if not os.path.isdir(REF):
        sys.exit("reference sources not present; golden vectors are committed data")
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        encode_case("ref_test1", png=os.path.join(REF, "codec", "test", "test1.png"),
                    meta={"source": "codec/test/test1.png"}, tmp=tmp)
        for name, seed, w, h in (("ref_c0", 0, 512, 512), ("ref_small", 3, 512, 256),
                                 ("ref_wide", 7, 1024, 512)):
            rgb = os.path.join(tmp, name + ".rgb")
            synthetic_rgb(seed, w, h, 300, rgb)
            encode_case(name, rgb_path=rgb, w=w, h=h, tmp=tmp, stages=(name == "ref_c0"),
                        meta={"generator": "gste_generate_rgb", "seed": seed, "width": w,
                              "height": h, "noise_milli": 300})
        # Given blocks: DXTImage(w, h, rgb, dxt) with the fast compressor's blocks.
        rgb = os.path.join(tmp, "given.rgb")
        pixels = synthetic_rgb(21, 512, 256, 300, rgb)
        blocks = os.path.join(tmp, "given.dxt_in")
        fast_dxt1(pixels, 512, 256, blocks)
        encode_case("ref_given", rgb_path=rgb, w=512, h=256, tmp=tmp, stages=False, dxt_in=blocks,
                    meta={"generator": "gste_generate_rgb", "seed": 21, "width": 512, "height": 256,
                          "noise_milli": 300, "blocks": "gste_compress_dxt1"})

        # Single-group rANS known answers.  The first count vector is the one of
        # ans/ans_ocl_test.cpp DecodeSingleStream; the others cover a 256-symbol
        # alphabet and a one-symbol-dominant alphabet.
        cases = [
            (1, [12, 14, 17, 1, 1, 2, 372]),
            (2, [1] * 256),
            (3, [5000] + [1] * 20),
            (4, list(range(1, 200))),
        ]
        groups = {}
        for seed, counts in cases:
            path = os.path.join(tmp, "g%d.bin" % seed)
            run("ans_group", seed, path, *counts)
            raw = np.frombuffer(read(path), np.uint8)
            freqs = raw[:512].view("<u2")
            n = int(raw[512:516].view("<u4")[0])
            groups["freq_%d" % seed] = freqs.copy()
            groups["group_%d" % seed] = raw[516:516 + n].copy()
            groups["symbols_%d" % seed] = raw[516 + n:516 + n + 8192].copy()
        np.savez_compressed(os.path.join(HERE, "ref_ans_groups.npz"), **groups)

    make_hist()


def ref_hist(m, counts):
    out = subprocess.run([DRIVER, "hist", str(m), *map(str, counts)], check=True,
                         capture_output=True, text=True).stdout.split()
    return [int(v) for v in out]


def make_hist():
    """ref_hist.json: GenerateHistogram outputs.  The ANS decoder tests' count vectors
    (ans/ans_ocl_test.cpp:65,108,158,221,278), the histogram KAT, and seeded random
    vectors: small and skewed counts, up to 256 symbols, and counts of 2^21 and more,
    whose count * M wraps in the reference's 32-bit product."""
    hist = {}
    for key, counts in (("ocl_init", [3, 2, 1, 4, 3]),
                        ("ocl_rebuild", [80, 300, 2, 14, 1, 1, 1, 20]),
                        ("ocl_single", [12, 14, 17, 1, 1, 2, 372]),
                        ("ocl_interleaved", [32, 186, 54, 8, 94, 35, 13, 21, 456, 789, 33, 215, 6, 54, 987,
                                             54, 65, 13, 2, 1]),
                        ("ocl_multiple", [65, 4, 6, 132, 135, 64, 879, 87, 456, 13, 2, 12, 33, 16, 546, 987,
                                          98, 74, 65, 43, 21, 32, 1]),
                        ("hist_kat", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10])):
        m = 256 if key == "hist_kat" else 2048
        hist[key] = {"counts": counts, "M": m, "normalized": ref_hist(m, counts)}
    rng = np.random.default_rng(2024)
    cases = []
    for i in range(48):
        n = int(rng.integers(2, 257))
        kind = i % 4
        if kind == 0:
            counts = rng.integers(0, 20, n)
        elif kind == 1:
            counts = rng.geometric(0.05, n) - 1
        elif kind == 2:
            counts = rng.integers(0, 1 << 24, n)
        else:
            counts = np.where(rng.random(n) < 0.3, rng.integers(1, 4, n), rng.integers(0, 5000, n))
        counts = [int(c) for c in counts]
        if not any(counts):
            counts[0] = 1
        m = 2048 if i % 6 else 4096
        cases.append({"counts": counts, "M": m, "normalized": ref_hist(m, counts)})
    hist["random"] = cases
    with open(os.path.join(HERE, "ref_hist.json"), "w") as f:
        json.dump(hist, f, separators=(",", ":"))
        f.write("\n")


def four_colour_blocks(dxt):
    """Blocks that the reference's CPU decompression (DXTImage::DecompressedImage, which
    follows the DXT1 endpoint-order rule) turns into what assemble_rgb
    (codec/assemble.cl:83-130, always four colours) computes for `dxt`: where ep1 < ep2
    the endpoints swap and every index flips its low bit (then ep1 > ep2 and each pixel
    keeps its colour: 0<->1, 2<->3); where ep1 == ep2 index 3 becomes 0 (both are ep1
    among four colours, but black in the three-colour rule)."""
    blk = np.frombuffer(bytes(dxt), "<u2").reshape(-1, 4).copy()
    e0, e1 = blk[:, 0].copy(), blk[:, 1].copy()
    word = blk[:, 2].astype(np.uint32) | (blk[:, 3].astype(np.uint32) << 16)
    lt, eq = e0 < e1, e0 == e1
    blk[lt, 0], blk[lt, 1] = e1[lt], e0[lt]
    word[lt] ^= np.uint32(0x55555555)
    weq = word[eq]
    for i in range(16):
        weq[((weq >> np.uint32(2 * i)) & 3) == 3] &= ~np.uint32(3 << (2 * i))
    word[eq] = weq
    blk[:, 2] = word & 0xFFFF
    blk[:, 3] = word >> 16
    return blk.tobytes()


def reference_rgb(dxt, w, h, tmp):
    src, out = os.path.join(tmp, "rgb_in.dxt"), os.path.join(tmp, "rgb_out.rgb")
    with open(src, "wb") as f:
        f.write(four_colour_blocks(dxt))
    run("decompress_rgb", src, w, h, out)
    return np.frombuffer(read(out), np.uint8)


def make_rgb():
    """ref_rgb.npz: RGB8 known answers for the LoadRGB path, computed by the reference's
    own CPU decompression (Decode565 + LerpChannels) through four_colour_blocks:
      small_rgb          for ref_small.npz's file;
      mixed_gtc / _dxt   a 512x256 file the reference encoder wrote from random starting
                         blocks (seed 77 of numpy's default_rng) over the seed-22 texture,
                         so that many blocks keep ep1 < ep2 or ep1 == ep2;
      mixed_rgb_sha256   SHA-256 of its expected RGB8 (the pixels are noise and would not
                         compress)."""
    import hashlib
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        small = np.load(os.path.join(HERE, "ref_small.npz"))
        arrays = {"small_rgb": reference_rgb(small["dxt"].tobytes(), 512, 256, tmp)}
        rgb = os.path.join(tmp, "mixed.rgb")
        synthetic_rgb(22, 512, 256, 300, rgb)
        blocks = os.path.join(tmp, "mixed.dxt_in")
        with open(blocks, "wb") as f:
            f.write(np.random.default_rng(77).integers(0, 256, 512 * 256 // 2, dtype=np.uint8).tobytes())
        gtc, dxt = os.path.join(tmp, "mixed.gtc"), os.path.join(tmp, "mixed.dxt")
        run("encode_rgb_dxt", rgb, 512, 256, blocks, gtc, dxt)
        arrays["mixed_gtc"] = np.frombuffer(read(gtc), np.uint8)
        arrays["mixed_dxt"] = np.frombuffer(read(dxt), np.uint8)
        expect = reference_rgb(read(dxt), 512, 256, tmp)
        arrays["mixed_rgb_sha256"] = np.frombuffer(hashlib.sha256(expect.tobytes()).digest(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "ref_rgb.npz"), **arrays)


if __name__ == "__main__":
    if sys.argv[1:] == ["hist"]:
        make_hist()
    elif sys.argv[1:] == ["rgb"]:
        make_rgb()
    else:
        main()
        make_rgb()
