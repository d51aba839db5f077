"""gst_amd: MI355X (gfx950) decoder for GST supercompressed textures.

Python view of libgst_amd.so's C ABI (include/gst_decoder.h), named after the
reference decoder's entry points (GammaUNC/GST codec/decoder.h:15-37):

    InitializeDecoder      codec/decoder.h:15
    DecompressDXT          codec/decoder.h:16   host bytes -> host DXT1 bytes
    LoadCompressedDXT[s]   codec/decoder.h:19,23 device staging buffer -> device output
    LoadRGB[s]             codec/decoder.h:27,31
    RequiredScratchMem     codec/decoder.h:35
    PreallocateDecompressor / FreeDecompressor   codec/decoder.h:36-37

Device buffers and streams are passed as integer addresses / handles (for example
`tensor.data_ptr()` and `torch.cuda.current_stream().cuda_stream`).  Every call
raises GstError on a non-zero status.  The library must be built
(gst_amd/build.py); importing without it raises, there is no fallback path.
"""
import ctypes
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgst_amd.so")

GST_OK = 0
STATUS_NAMES = {
    -1: "GST_ERR_INVALID_ARGUMENT",
    -2: "GST_ERR_BAD_HEADER",
    -3: "GST_ERR_CORRUPT_STREAM",
    -4: "GST_ERR_OUT_OF_MEMORY",
    -5: "GST_ERR_HIP",
    -6: "GST_ERR_NOT_INITIALIZED",
    -7: "GST_ERR_BUSY",
    -8: "GST_ERR_TOO_LARGE",
}
DEVERR_STREAM_OFFSETS = 0x1
DEVERR_FREQUENCIES = 0x2
DEVERR_PALETTE_INDEX = 0x4
STAGES = ("build_table", "ans_decode", "decode_indices", "inv_wavelet", "assemble")


class GstError(RuntimeError):
    def __init__(self, code, where):
        super().__init__("%s failed: %s (%d)" % (where, STATUS_NAMES.get(code, "unknown"), code))
        self.code = code


class Header(ctypes.Structure):
    """GenTCHeader (codec/codec_base.h:9-20)."""
    _fields_ = [(name, ctypes.c_uint32) for name in (
        "width", "height", "palette_bytes", "y_cmp_sz", "chroma_cmp_sz", "palette_sz", "indices_sz")]

    @property
    def num_blocks(self):
        return self.width * self.height // 16

    @property
    def stream_bytes(self):
        return self.y_cmp_sz + self.chroma_cmp_sz + self.palette_sz + self.indices_sz

    @property
    def dxt_bytes(self):
        return self.width * self.height // 2

    def __repr__(self):
        return "Header(%s)" % ", ".join("%s=%d" % (n, getattr(self, n)) for n, _ in self._fields_)


# rANS work-group forms of the 4-wave x 2-group shape (GST_ANS_FORM_*).
ANS_FORMS = {0: "auto", 1: "paired", 2: "plain"}
ANS_FORM_CODES = {v: k for k, v in ANS_FORMS.items()}


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("capacity_images", ctypes.c_uint32),
        ("capacity_blocks", ctypes.c_uint64),
        ("capacity_palette", ctypes.c_uint64),
        ("ans_groups_per_wave", ctypes.c_uint32),
        ("ans_waves_per_group", ctypes.c_uint32),
        ("grid_ans", ctypes.c_uint32),
        ("grid_wavelet", ctypes.c_uint32),
        ("grid_indices", ctypes.c_uint32),
        ("grid_assemble", ctypes.c_uint32),
        ("rgb_output", ctypes.c_uint32),
        ("wavelet_waves_per_group", ctypes.c_uint32),
        ("assembly_waves_per_group", ctypes.c_uint32),
        ("separate_index_scan", ctypes.c_uint32),
        ("separate_table_stage", ctypes.c_uint32),
        ("assembly_wide", ctypes.c_uint32),
        ("separate_wavelet", ctypes.c_uint32),
        ("ans_form", ctypes.c_uint32),
    ]


def _stage_times(ms):
    """Stage name -> ms, None for a stage with no launch of its own (NaN from the ABI)."""
    return {k: (None if math.isnan(v) else v) for k, v in zip(STAGES, list(ms))}


class TuneResult(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint32) for name in (
        "ans_groups_per_wave", "ans_waves_per_group", "wavelet_waves_per_group", "assembly_waves_per_group",
        "ans_builds_tables", "assembly_wide", "grid_ans", "grid_indices", "grid_wavelet", "grid_assemble",
        "candidates")] + [("stage_ms", ctypes.c_float * 5), ("total_ms", ctypes.c_float),
                          ("wavelet_fused", ctypes.c_uint32), ("ans_form", ctypes.c_uint32)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_[:11]}
        d["stage_ms"] = _stage_times(self.stage_ms)
        d["total_ms"] = self.total_ms
        d["wavelet_fused"] = self.wavelet_fused
        d["ans_form"] = ANS_FORMS.get(self.ans_form) if self.ans_form else None
        d["streams_per_wavefront"] = 32 * self.ans_groups_per_wave
        return d


_lib = None
_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_SZ = ctypes.c_size_t
_HP = ctypes.POINTER(Header)


def _declare(lib):
    sig = {
        "gst_initialize_decoder": ([ctypes.c_int], ctypes.c_int),
        "gst_parse_header": ([ctypes.c_char_p, _SZ, _HP], ctypes.c_int),
        "gst_decompress_dxt": ([ctypes.c_char_p, _SZ, _P, _SZ], ctypes.c_int),
        "gst_load_compressed_dxt": ([_HP, _P, _P, _P, _U32, _P, _P], ctypes.c_int),
        "gst_load_compressed_dxts": ([_HP, _U32, _P, _P, _P, _U32, _P, _P], ctypes.c_int),
        "gst_load_compressed_dxt_noalloc": ([_HP, _P, _P, _P, _P], ctypes.c_int),
        "gst_read_error_bits_async": ([_P, _P], ctypes.c_int),
        "gst_load_rgb": ([_HP, _P, _P, _P, _U32, _P, _P], ctypes.c_int),
        "gst_load_rgbs": ([_HP, _U32, _P, _P, _P, _U32, _P, _P], ctypes.c_int),
        "gst_required_scratch_mem": ([_HP], _SZ),
        "gst_preallocate_decompressor": ([_SZ], ctypes.c_int),
        "gst_free_decompressor": ([], ctypes.c_int),
        "gst_read_error_bits": ([_P, ctypes.POINTER(_U32)], ctypes.c_int),
        "gst_staging_size": ([_HP, _U32, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], ctypes.c_int),
        "gst_pack_staging": ([ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_SZ), _U32, _P, _SZ],
                             ctypes.c_int),
        "gst_upload_gather": ([ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_uint64), _U32, _P],
                              ctypes.c_int),
        "gst_stage_build_table": ([_P, _P, _P, _P], ctypes.c_int),
        "gst_stage_ans_decode": ([_HP, _P, _P, _P, _P, _P, _P], ctypes.c_int),
        "gst_stage_decode_indices": ([_HP, _P, _P, _P, _P], ctypes.c_int),
        "gst_stage_inv_wavelet": ([_HP, _P, _P, _P], ctypes.c_int),
        "gst_stage_assemble_dxt": ([_HP, _P, _P, _P, _P, _P, _P], ctypes.c_int),
        "gst_stage_assemble_rgb": ([_HP, _P, _P, _P, _P, _P, _P], ctypes.c_int),
        "gst_batch_create": ([ctypes.POINTER(BatchConfig), ctypes.POINTER(_P)], ctypes.c_int),
        "gst_batch_bind": ([_P, _HP, _U32, ctypes.POINTER(_P), ctypes.POINTER(_P), _P], ctypes.c_int),
        "gst_batch_decode": ([_P, _P], ctypes.c_int),
        "gst_batch_capture_graph": ([_P, _P], ctypes.c_int),
        "gst_batch_launch_graph": ([_P, _P], ctypes.c_int),
        "gst_batch_decode_timed": ([_P, _P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "gst_batch_decode_timed_reps": ([_P, _P, _U32, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "gst_batch_error_bits": ([_P, _P, ctypes.POINTER(_U32)], ctypes.c_int),
        "gst_batch_autotune": ([_P, _P, _U32, ctypes.POINTER(TuneResult)], ctypes.c_int),
        "gst_batch_device_bytes": ([_P], _SZ),
        "gst_batch_destroy": ([_P], ctypes.c_int),
        "gst_ans_normalize_frequencies": ([ctypes.POINTER(_U32), _U32, ctypes.c_int32, ctypes.POINTER(_U32)],
                                          ctypes.c_int),
        "gst_ans_decoder_create": ([ctypes.POINTER(_U32), _U32, _U32, ctypes.POINTER(_P)], ctypes.c_int),
        "gst_ans_rebuild_table": ([_P, ctypes.POINTER(_U32), _U32], ctypes.c_int),
        "gst_ans_get_table": ([_P, _P, _P, _P], ctypes.c_int),
        "gst_ans_decode": ([_P, ctypes.POINTER(_U32), _U32, ctypes.POINTER(_P), ctypes.POINTER(_SZ), _U32, _P],
                           ctypes.c_int),
        "gst_ans_decoder_destroy": ([_P], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def _torch_runtime_first():
    """PyTorch-ROCm ships its own HIP and HSA runtimes.  When torch is installed it
    loads and initialises them before libgst_amd.so is loaded, so that the library binds
    to the runtime already in the process; loaded the other way round, the system
    runtime holds the device and torch reports no GPU."""
    try:
        import torch
        torch.cuda.is_available()
    except Exception:  # torch absent or without GPU support: nothing to order
        pass


def lib():
    """The loaded libgst_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgst_amd.so is not built; run gst_amd/build.py (no fallback exists)")
        _torch_runtime_first()
        handle = ctypes.CDLL(LIB_PATH)
        _declare(handle)
        _lib = handle
    return _lib


def _check(rc, where):
    if rc != GST_OK:
        raise GstError(rc, where)


def _events(init):
    init = list(init or ())
    if not init:
        return 0, None
    arr = (_P * len(init))(*init)
    return len(init), ctypes.cast(arr, _P)


# ---- codec/decoder.h mirror -------------------------------------------------------

def InitializeDecoder(device=0):
    """True when the device is a gfx950 that meets every kernel's requirements."""
    return lib().gst_initialize_decoder(device) == GST_OK


def ParseHeader(data):
    """GenTCHeader::LoadFrom plus the decoder's size rules (raises GstError)."""
    hdr = Header()
    _check(lib().gst_parse_header(bytes(data), len(data), ctypes.byref(hdr)), "ParseHeader")
    return hdr


def DecompressDXT(data):
    """Host .gtc bytes -> host DXT1 blocks (bytes)."""
    hdr = ParseHeader(data)
    out = ctypes.create_string_buffer(hdr.dxt_bytes)
    _check(lib().gst_decompress_dxt(bytes(data), len(data), out, len(out)), "DecompressDXT")
    return out.raw


def _load(fn, name, hdrs, stream, d_cmp, d_out, init, want_event):
    hdrs = list(hdrs) if isinstance(hdrs, (list, tuple)) else [hdrs]
    arr = (Header * len(hdrs))(*hdrs)
    n_init, init_ptr = _events(init)
    ev = _P()
    if name.endswith("s"):
        rc = fn(arr, len(hdrs), stream, d_cmp, d_out, n_init, init_ptr, ctypes.byref(ev) if want_event else None)
    else:
        rc = fn(arr, stream, d_cmp, d_out, n_init, init_ptr, ctypes.byref(ev) if want_event else None)
    _check(rc, name)
    return ev.value if want_event else None


def LoadCompressedDXT(hdr, stream, d_cmp, d_out, init=(), want_event=False):
    return _load(lib().gst_load_compressed_dxt, "LoadCompressedDXT", hdr, stream, d_cmp, d_out, init, want_event)


def LoadCompressedDXTs(hdrs, stream, d_cmp, d_out, init=(), want_event=False):
    return _load(lib().gst_load_compressed_dxts, "LoadCompressedDXTs", hdrs, stream, d_cmp, d_out, init,
                 want_event)


def LoadCompressedDXTNoAlloc(hdr, stream, d_cmp, d_out, d_err_bits):
    """gst_load_compressed_dxt_noalloc: one image, no allocation (static scratch);
    error bits are OR-ed into the device word at d_err_bits."""
    _check(lib().gst_load_compressed_dxt_noalloc(ctypes.byref(hdr), stream, d_cmp, d_out, d_err_bits),
           "LoadCompressedDXTNoAlloc")


def LoadRGB(hdr, stream, d_cmp, d_out, init=(), want_event=False):
    return _load(lib().gst_load_rgb, "LoadRGB", hdr, stream, d_cmp, d_out, init, want_event)


def LoadRGBs(hdrs, stream, d_cmp, d_out, init=(), want_event=False):
    return _load(lib().gst_load_rgbs, "LoadRGBs", hdrs, stream, d_cmp, d_out, init, want_event)


def RequiredScratchMem(hdr):
    return lib().gst_required_scratch_mem(ctypes.byref(hdr))


def PreallocateDecompressor(nbytes):
    _check(lib().gst_preallocate_decompressor(nbytes), "PreallocateDecompressor")


def FreeDecompressor():
    _check(lib().gst_free_decompressor(), "FreeDecompressor")


def ReadErrorBits(stream):
    bits = _U32()
    _check(lib().gst_read_error_bits(stream, ctypes.byref(bits)), "ReadErrorBits")
    return bits.value


def StagingSize(hdrs):
    hdrs = list(hdrs)
    arr = (Header * len(hdrs))(*hdrs)
    total, off = _SZ(), _SZ()
    _check(lib().gst_staging_size(arr, len(hdrs), ctypes.byref(total), ctypes.byref(off)), "StagingSize")
    return total.value, off.value


def UploadGather(host_ptrs, dev_ptrs, sizes, stream):
    """gst_upload_gather: pinned host ranges -> device ranges in one gather kernel per 64."""
    n = len(sizes)
    src = (_P * max(n, 1))(*host_ptrs)
    dst = (_P * max(n, 1))(*dev_ptrs)
    nb = (ctypes.c_uint64 * max(n, 1))(*[int(b) for b in sizes])
    _check(lib().gst_upload_gather(src, dst, nb, n, stream), "gst_upload_gather")


def PackStaging(files):
    """Host staging buffer (bytes) for LoadCompressedDXT[s] from .gtc files."""
    files = [bytes(f) for f in files]
    hdrs = [ParseHeader(f) for f in files]
    total, _ = StagingSize(hdrs)
    out = ctypes.create_string_buffer(total)
    ptrs = (ctypes.c_char_p * len(files))(*files)
    lens = (_SZ * len(files))(*[len(f) for f in files])
    _check(lib().gst_pack_staging(ptrs, lens, len(files), out, total), "PackStaging")
    return out.raw, hdrs


# ---- one launch per stage ------------------------------------------------------------

def StageBuildTable(d_freqs, d_tables, d_err, stream):
    _check(lib().gst_stage_build_table(d_freqs, d_tables, d_err, stream), "StageBuildTable")


def StageAnsDecode(hdr, d_streams, d_tables, d_symbols, d_group_sums, d_err, stream):
    _check(lib().gst_stage_ans_decode(ctypes.byref(hdr), d_streams, d_tables, d_symbols, d_group_sums,
                                      d_err, stream), "StageAnsDecode")


def StageDecodeIndices(hdr, d_symbols, d_group_sums, d_indices, stream):
    _check(lib().gst_stage_decode_indices(ctypes.byref(hdr), d_symbols, d_group_sums, d_indices, stream),
           "StageDecodeIndices")


def StageInvWavelet(hdr, d_symbols, d_planes, stream):
    _check(lib().gst_stage_inv_wavelet(ctypes.byref(hdr), d_symbols, d_planes, stream), "StageInvWavelet")


def StageAssembleDXT(hdr, d_symbols, d_planes, d_indices, d_out, d_err, stream):
    _check(lib().gst_stage_assemble_dxt(ctypes.byref(hdr), d_symbols, d_planes, d_indices, d_out, d_err,
                                        stream), "StageAssembleDXT")


def StageAssembleRGB(hdr, d_symbols, d_planes, d_indices, d_out, d_err, stream):
    _check(lib().gst_stage_assemble_rgb(ctypes.byref(hdr), d_symbols, d_planes, d_indices, d_out, d_err,
                                        stream), "StageAssembleRGB")


# ---- batched decode of uneven sizes ------------------------------------------------

class Batch:
    """A device-resident plan for up to `capacity_images` textures of any legal size."""

    def __init__(self, capacity_images, capacity_blocks, capacity_palette=0, groups_per_wave=0,
                 waves_per_group=0, grids=None, rgb=False, wavelet_waves=0, assembly_waves=0,
                 separate_index_scan=False, separate_table_stage=False, assembly_wide=False,
                 separate_wavelet=False, ans_form="auto"):
        grids = grids or {}
        cfg = BatchConfig(capacity_images, capacity_blocks, capacity_palette, groups_per_wave,
                          waves_per_group, grids.get("ans_decode", 0), grids.get("inv_wavelet", 0),
                          grids.get("decode_indices", 0), grids.get("assemble", 0), 1 if rgb else 0,
                          wavelet_waves, assembly_waves, 1 if separate_index_scan else 0,
                          1 if separate_table_stage else 0, 1 if assembly_wide else 0,
                          1 if separate_wavelet else 0, ANS_FORM_CODES[ans_form])
        self._h = _P()
        _check(lib().gst_batch_create(ctypes.byref(cfg), ctypes.byref(self._h)), "gst_batch_create")

    def bind(self, hdrs, d_files, d_outs, stream):
        hdrs = list(hdrs)
        n = len(hdrs)
        arr = (Header * n)(*hdrs)
        files = (_P * n)(*d_files)
        outs = (_P * n)(*d_outs)
        _check(lib().gst_batch_bind(self._h, arr, n, files, outs, stream), "gst_batch_bind")

    def decode(self, stream):
        _check(lib().gst_batch_decode(self._h, stream), "gst_batch_decode")

    def capture_graph(self, stream):
        _check(lib().gst_batch_capture_graph(self._h, stream), "gst_batch_capture_graph")

    def launch_graph(self, stream):
        _check(lib().gst_batch_launch_graph(self._h, stream), "gst_batch_launch_graph")

    def decode_timed(self, stream, reps=1):
        """Per-stage ms of one decode, or the mean over `reps` decodes run back to back."""
        ms = (ctypes.c_float * 5)()
        if reps == 1:
            _check(lib().gst_batch_decode_timed(self._h, stream, ms), "gst_batch_decode_timed")
        else:
            _check(lib().gst_batch_decode_timed_reps(self._h, stream, reps, ms), "gst_batch_decode_timed_reps")
        return _stage_times(ms)

    def error_bits(self, stream):
        bits = _U32()
        _check(lib().gst_batch_error_bits(self._h, stream, ctypes.byref(bits)), "gst_batch_error_bits")
        return bits.value

    def autotune(self, stream, reps=5):
        """Times every launch shape on the bound batch and keeps the fastest per stage."""
        res = TuneResult()
        _check(lib().gst_batch_autotune(self._h, stream, reps, ctypes.byref(res)), "gst_batch_autotune")
        return res.as_dict()

    @property
    def device_bytes(self):
        return lib().gst_batch_device_bytes(self._h)

    def close(self):
        if self._h:
            lib().gst_batch_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- stand-alone interleaved rANS (ans/ans_ocl.h:26-72) -----------------------------

ANS_TABLE_SIZE = 2048
ANS_SYMBOLS_PER_STREAM = 256


def _u32_array(values):
    values = [int(v) for v in values]
    return (_U32 * max(1, len(values)))(*values), len(values)


def NormalizeFrequencies(counts, M=ANS_TABLE_SIZE):
    """ans::GenerateHistogram(counts, M) (ans/histogram.cpp:40-121)."""
    arr, n = _u32_array(counts)
    out = (_U32 * max(1, n))()
    _check(lib().gst_ans_normalize_frequencies(arr, n, int(M), out), "gst_ans_normalize_frequencies")
    return [int(v) for v in out[:n]]


class AnsDecoder:
    """ans::ocl::OpenCLDecoder (ans/ans_ocl.h:26-72): one device table, blocking decodes."""

    def __init__(self, F, num_interleaved):
        arr, n = _u32_array(F)
        self._h = _P()
        _check(lib().gst_ans_decoder_create(arr, n, int(num_interleaved), ctypes.byref(self._h)),
               "gst_ans_decoder_create")

    def rebuild_table(self, F):
        arr, n = _u32_array(F)
        _check(lib().gst_ans_rebuild_table(self._h, arr, n), "gst_ans_rebuild_table")

    def table(self):
        """(symbols u8[2048], frequencies u16[2048], cumulative frequencies u16[2048])."""
        import numpy as np
        sym = np.zeros(ANS_TABLE_SIZE, np.uint8)
        freq = np.zeros(ANS_TABLE_SIZE, np.uint16)
        cum = np.zeros(ANS_TABLE_SIZE, np.uint16)
        _check(lib().gst_ans_get_table(self._h, sym.ctypes.data, freq.ctypes.data, cum.ctypes.data),
               "gst_ans_get_table")
        return sym, freq, cum

    def decode(self, states, data):
        """`data`: one bytes object (all states interleaved in it) or a list of them (each
        holding len(states) / len(data) streams).  Returns one bytes object of 256
        symbols per state."""
        blobs = [bytes(data)] if isinstance(data, (bytes, bytearray)) else [bytes(d) for d in data]
        st, ns = _u32_array(states)
        bufs = [ctypes.create_string_buffer(b, max(1, len(b))) for b in blobs]
        ptrs = (_P * len(bufs))(*[ctypes.cast(b, _P) for b in bufs])
        lens = (_SZ * len(bufs))(*[len(b) for b in blobs])
        out = ctypes.create_string_buffer(ns * ANS_SYMBOLS_PER_STREAM)
        _check(lib().gst_ans_decode(self._h, st, ns, ptrs, lens, len(bufs), out), "gst_ans_decode")
        raw = out.raw
        return [raw[i * ANS_SYMBOLS_PER_STREAM:(i + 1) * ANS_SYMBOLS_PER_STREAM] for i in range(ns)]

    def close(self):
        if self._h:
            lib().gst_ans_decoder_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
