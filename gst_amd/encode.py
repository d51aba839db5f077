"""The GST encoder and seeded synthetic textures (libgst_encode.so, include/gst_encode.h).

`compress_dxt` is the reference encoder gentenc (GenTC::CompressDXT) byte for byte.
The rest is input tooling for tests and benchmarks: a fast encoder that makes .gtc
files in the reference's on-disk layout from deterministic textures, on any host.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgst_encode.so")
DEFAULT_NOISE_MILLI = 300   # noise amplitude 0.3 of the 0..255 range
DEFAULT_REUSE = 2           # palette-reuse MSE threshold (tracks gentenc's stream sizes)

_lib = None


class EncodeStats(ctypes.Structure):
    _fields_ = [("palette_entries", ctypes.c_uint32), ("reused_blocks", ctypes.c_uint32),
                ("wavelet_overflows", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgst_encode.so is not built; run gst_amd/build.py")
        h = ctypes.CDLL(LIB_PATH)
        h.gste_max_encoded_size.restype = ctypes.c_size_t
        h.gste_max_encoded_size.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        h.gste_generate_rgb.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_void_p]
        h.gste_compress_dxt1.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        h.gste_encode.restype = ctypes.c_size_t
        h.gste_encode.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        h.gste_make_texture.restype = ctypes.c_size_t
        h.gste_make_texture.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                        ctypes.c_void_p]
        h.gste_compress_dxt.restype = ctypes.c_int
        h.gste_compress_dxt.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                        ctypes.c_void_p]
        h.gste_make_textures_gentenc.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_int]
        h.gste_make_textures.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        _lib = h
    return _lib


GSTE_OK, GSTE_ERR_ARGUMENT, GSTE_ERR_CAPACITY, GSTE_ERR_RANGE, GSTE_ERR_UNCODABLE = 0, -1, -2, -3, -4


class EncodeError(ValueError):
    """gste_compress_dxt failed; `code` is the GSTE_* status."""

    def __init__(self, code, what):
        super().__init__("%s (status %d)" % (what, code))
        self.code = code


def compress_dxt(rgb, width=None, height=None, dxt=None):
    """GenTC::CompressDXT (codec/encoder.cpp:176-187): the .gtc file gentenc writes for
    an RGB8 image, and the DXT1 blocks it decodes to.  `rgb`: (height, width, 3) uint8
    array, or flat bytes with `width` and `height`.  `dxt`: optional starting blocks
    (width*height/2 bytes, DXTImage(w, h, rgb, dxt)); None compresses them with stb_dxt
    as gentenc does.  Returns (gtc bytes, dxt bytes)."""
    a = np.ascontiguousarray(np.frombuffer(rgb, np.uint8) if isinstance(rgb, (bytes, bytearray)) else rgb,
                             dtype=np.uint8)
    if a.ndim == 3:
        height, width = a.shape[0], a.shape[1]
        if a.shape[2] != 3:
            raise ValueError("rgb must have 3 channels")
    if width is None or height is None or a.size != width * height * 3:
        raise ValueError("rgb size does not match width x height x 3")
    blocks = None
    if dxt is not None:
        blocks = np.ascontiguousarray(np.frombuffer(bytes(dxt), np.uint8))
        if blocks.size != width * height // 2:
            raise ValueError("dxt must hold width*height/2 bytes")
    need = ctypes.c_size_t()
    cap = lib().gste_max_encoded_size(width, height)
    out = np.empty(cap, np.uint8)
    blk = np.empty(width * height // 2, np.uint8)
    rc = lib().gste_compress_dxt(a.ctypes.data, width, height, None if blocks is None else blocks.ctypes.data,
                                 out.ctypes.data, cap, ctypes.byref(need), blk.ctypes.data)
    if rc == GSTE_ERR_CAPACITY:  # not expected: the bound covers every legal input
        out = np.empty(need.value, np.uint8)
        rc = lib().gste_compress_dxt(a.ctypes.data, width, height,
                                     None if blocks is None else blocks.ctypes.data,
                                     out.ctypes.data, need.value, ctypes.byref(need), blk.ctypes.data)
    if rc != GSTE_OK:
        raise EncodeError(rc, {GSTE_ERR_ARGUMENT: "size gentenc cannot code",
                               GSTE_ERR_RANGE: "palette-index step outside a byte",
                               GSTE_ERR_UNCODABLE: "a stream of zero bytes only"}.get(rc, "failed"))
    return out[:need.value].tobytes(), blk.tobytes()


def generate_rgb(seed, width, height, noise_milli=DEFAULT_NOISE_MILLI):
    buf = np.empty(width * height * 3, np.uint8)
    lib().gste_generate_rgb(seed, width, height, noise_milli, buf.ctypes.data)
    return buf.reshape(height, width, 3)


def make_texture(seed, width, height, noise_milli=DEFAULT_NOISE_MILLI, reuse=DEFAULT_REUSE):
    """(gtc bytes, expected DXT1 bytes, EncodeStats)."""
    cap = lib().gste_max_encoded_size(width, height)
    out = np.empty(cap, np.uint8)
    dxt = np.empty(width * height // 2, np.uint8)
    st = EncodeStats()
    n = lib().gste_make_texture(seed, width, height, noise_milli, reuse, out.ctypes.data, cap,
                                dxt.ctypes.data, ctypes.byref(st))
    if n == 0:
        raise ValueError("cannot encode %dx%d" % (width, height))
    return out[:n].tobytes(), dxt.tobytes(), st


def make_textures(specs, noise_milli=DEFAULT_NOISE_MILLI, reuse=DEFAULT_REUSE, nthreads=None,
                  want_dxt=True, encoder="fast"):
    """specs: list of (seed, width, height).  Returns (files, dxts) lists of bytes.
    encoder "gentenc": every file byte for byte the reference encoder's (stb_dxt blocks,
    gentenc's palette reuse; `reuse` is ignored); "fast": the corpus encoder."""
    if encoder not in ("fast", "gentenc"):
        raise ValueError("encoder must be 'fast' or 'gentenc'")
    n = len(specs)
    seeds = np.array([s[0] for s in specs], np.uint64)
    ws = np.array([s[1] for s in specs], np.uint32)
    hs = np.array([s[2] for s in specs], np.uint32)
    caps = np.array([lib().gste_max_encoded_size(int(w), int(h)) for w, h in zip(ws, hs)], np.uint64)
    out_off = np.concatenate([[0], np.cumsum(caps)[:-1]]).astype(np.uint64)
    dx = (ws.astype(np.uint64) * hs.astype(np.uint64)) // 2
    dxt_off = np.concatenate([[0], np.cumsum(dx)[:-1]]).astype(np.uint64)
    out = np.empty(int(caps.sum()), np.uint8)
    dxt = np.empty(int(dx.sum()) if want_dxt else 1, np.uint8)
    sizes = np.zeros(n, np.uint64)
    threads = nthreads or min(16, os.cpu_count() or 1)
    if encoder == "gentenc":
        lib().gste_make_textures_gentenc(n, seeds.ctypes.data, ws.ctypes.data, hs.ctypes.data, noise_milli,
                                         out.ctypes.data, out_off.ctypes.data,
                                         dxt.ctypes.data if want_dxt else None, dxt_off.ctypes.data,
                                         sizes.ctypes.data, threads)
    else:
        lib().gste_make_textures(n, seeds.ctypes.data, ws.ctypes.data, hs.ctypes.data, noise_milli, reuse,
                                 out.ctypes.data, out_off.ctypes.data, dxt.ctypes.data if want_dxt else None,
                                 dxt_off.ctypes.data, sizes.ctypes.data, threads)
    if not all(sizes):
        raise ValueError("encoding failed for some textures")
    files = [out[int(o):int(o) + int(s)].tobytes() for o, s in zip(out_off, sizes)]
    dxts = [dxt[int(o):int(o) + int(d)].tobytes() for o, d in zip(dxt_off, dx)] if want_dxt else None
    return files, dxts
