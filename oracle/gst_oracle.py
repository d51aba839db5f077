"""ctypes view of liboracle_gst.so, the CPU restatement of the GST decode path.

TEST INFRASTRUCTURE ONLY: the checker for the HIP decoder.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gst.so")

OK, ERR_ARG, ERR_HEADER, ERR_STREAM, ERR_FREQ = 0, -1, -2, -3, -4


class Header(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "width", "height", "palette_bytes", "y_cmp_sz", "chroma_cmp_sz", "palette_sz", "indices_sz")]


class TableEntry(ctypes.Structure):
    _fields_ = [("freq", ctypes.c_uint16), ("cum_freq", ctypes.c_uint16), ("symbol", ctypes.c_uint8)]


class Stages(ctypes.Structure):
    _fields_ = [("symbols", ctypes.c_void_p), ("planes", ctypes.c_void_p), ("indices", ctypes.c_void_p)]


_lib = None
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liboracle_gst.so is not built (make -C oracle)")
        h = ctypes.CDLL(LIB_PATH)
        h.gsto_parse_header.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Header)]
        h.gsto_build_table.argtypes = [_P, _P]
        h.gsto_ans_decode_group.argtypes = [_P, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32,
                                            ctypes.c_uint32, _P, _P, _P]
        h.gsto_decode_indices.argtypes = [_P, ctypes.c_uint32, _P]
        h.gsto_decode_indices_staged.argtypes = [_P, ctypes.c_uint32, _P, ctypes.c_int, ctypes.c_int32]
        h.gsto_inv_wavelet_tile.argtypes = [_P, _P]
        h.gsto_inv_wavelet_level.argtypes = [_P, ctypes.c_int]
        h.gsto_ycocg_to_rgb565.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        h.gsto_ycocg_to_rgb565.restype = ctypes.c_uint16
        h.gsto_assemble_dxt.argtypes = [_P, _P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P]
        h.gsto_assemble_dxt.restype = ctypes.c_uint32
        h.gsto_assemble_rgb.argtypes = [_P, _P, _P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P]
        h.gsto_assemble_rgb.restype = ctypes.c_uint32
        h.gsto_decode_dxt.argtypes = [ctypes.c_char_p, ctypes.c_size_t, _P, ctypes.POINTER(Stages), ctypes.c_int]
        h.gsto_decode_rgb.argtypes = [ctypes.c_char_p, ctypes.c_size_t, _P, ctypes.c_int]
        _lib = h
    return _lib


def parse_header(data):
    hdr = Header()
    rc = lib().gsto_parse_header(bytes(data), len(data), ctypes.byref(hdr))
    return rc, hdr


def build_table(freqs):
    """freqs: 256 uint16 -> (freq, cum, symbol) arrays of 2048."""
    f = np.ascontiguousarray(freqs, np.uint16)
    t = (TableEntry * 2048)()
    lib().gsto_build_table(f.ctypes.data, ctypes.addressof(t))
    freq = np.array([e.freq for e in t], np.uint16)
    cum = np.array([e.cum_freq for e in t], np.uint16)
    sym = np.array([e.symbol for e in t], np.uint8)
    return freq, cum, sym


def table_struct(freqs):
    f = np.ascontiguousarray(freqs, np.uint16)
    t = (TableEntry * 2048)()
    lib().gsto_build_table(f.ctypes.data, ctypes.addressof(t))
    return t


def ans_decode_group(freqs, stream, num_groups, group):
    """-> (rc, symbols[8192], final_states[32], final_word)."""
    t = table_struct(freqs)
    out = np.zeros(8192, np.uint8)
    states = np.zeros(32, np.uint32)
    word = ctypes.c_uint32()
    rc = lib().gsto_ans_decode_group(ctypes.addressof(t), bytes(stream), len(stream), num_groups, group,
                                     out.ctypes.data, states.ctypes.data, ctypes.byref(word))
    return rc, out, states, word.value


def decode_indices(diffs):
    d = np.ascontiguousarray(diffs, np.uint8)
    out = np.zeros(len(d), np.int32)
    lib().gsto_decode_indices(d.ctypes.data, len(d), out.ctypes.data)
    return out


def decode_indices_staged(diffs, race_policy=0, garbage=0x5A5A5A5A):
    d = np.ascontiguousarray(diffs, np.uint8)
    out = np.zeros(len(d), np.int32)
    lib().gsto_decode_indices_staged(d.ctypes.data, len(d), out.ctypes.data, race_policy, garbage)
    return out


def inv_wavelet_tile(tile_bytes):
    t = np.ascontiguousarray(tile_bytes, np.uint8)
    out = np.zeros(1024, np.int8)
    lib().gsto_inv_wavelet_tile(t.ctypes.data, out.ctypes.data)
    return out


def inv_wavelet_level(tile32, length):
    t = np.ascontiguousarray(tile32, np.int32).copy()
    lib().gsto_inv_wavelet_level(t.ctypes.data, length)
    return t


def ycocg_to_rgb565(y, co, cg):
    return lib().gsto_ycocg_to_rgb565(int(y), int(co), int(cg))


def decode_dxt(data, nthreads=1, stages=False):
    """-> (rc, dxt bytes[, dict of stage arrays])."""
    rc, hdr = parse_header(data)
    if rc != OK:
        return rc, None, None
    nv = hdr.width * hdr.height // 16
    out = np.zeros(nv * 8, np.uint8)
    st = None
    arrays = None
    if stages:
        arrays = {"symbols": np.zeros(7 * nv + hdr.palette_bytes, np.uint8),
                  "planes": np.zeros(6 * nv, np.int8), "indices": np.zeros(nv, np.int32)}
        st = Stages(arrays["symbols"].ctypes.data, arrays["planes"].ctypes.data,
                    arrays["indices"].ctypes.data)
    rc = lib().gsto_decode_dxt(bytes(data), len(data), out.ctypes.data,
                               ctypes.byref(st) if st is not None else None, nthreads)
    return rc, out.tobytes(), arrays


def decode_rgb(data, nthreads=1):
    rc, hdr = parse_header(data)
    if rc != OK:
        return rc, None
    out = np.zeros(hdr.width * hdr.height * 3, np.uint8)
    rc = lib().gsto_decode_rgb(bytes(data), len(data), out.ctypes.data, nthreads)
    return rc, out.tobytes()
