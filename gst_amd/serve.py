"""Python view of the continuous-batching decode server (include/gst_serve.h)."""
import ctypes

import numpy as np

from . import GstError, _check, lib as _decoder_lib


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("host_staging_bytes", ctypes.c_uint64),
        ("slots", ctypes.c_uint32),
        ("max_batch_images", ctypes.c_uint32),
        ("max_batch_blocks", ctypes.c_uint64),
        ("copy_to_host", ctypes.c_uint32),
        ("auto_release", ctypes.c_uint32),
        ("plain", ctypes.c_uint32),
        ("dry_run", ctypes.c_uint32),
        ("ans_groups_per_wave", ctypes.c_uint32),
        ("ans_waves_per_group", ctypes.c_uint32),
        ("graph_launch", ctypes.c_uint32),
        ("poison_outputs", ctypes.c_uint32),
        ("fail_launch_every", ctypes.c_uint32),
        ("idle_spin_us", ctypes.c_uint32),
        ("copy_kernel", ctypes.c_uint32),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("file", ctypes.c_char_p),
        ("d_file", ctypes.c_void_p),
        ("len", ctypes.c_size_t),
        ("tag", ctypes.c_uint64),
        ("d_expected", ctypes.c_void_p),
    ]


class Completion(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_uint64),
        ("ticket", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("batch_images", ctypes.c_uint32),
        ("latency_device_us", ctypes.c_double),
        ("latency_host_us", ctypes.c_double),
        ("queue_us", ctypes.c_double),
        ("d_output", ctypes.c_void_p),
        ("output_bytes", ctypes.c_uint64),
        ("verified", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
        ("batches", ctypes.c_uint64), ("texels", ctypes.c_uint64), ("algorithmic_bytes", ctypes.c_uint64),
        ("device_busy_ms", ctypes.c_double), ("max_batch_images", ctypes.c_double),
        ("mean_batch_images", ctypes.c_double), ("peak_arena_bytes", ctypes.c_uint64),
        ("admission_waits", ctypes.c_uint64), ("verified", ctypes.c_uint64), ("mismatches", ctypes.c_uint64),
        ("h2d_bytes", ctypes.c_uint64), ("h2d_copies", ctypes.c_uint64), ("prestaged", ctypes.c_uint64),
        ("launch_host_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class LoadSpec(ctypes.Structure):
    _fields_ = [
        ("pool", ctypes.POINTER(ctypes.c_char_p)),
        ("pool_lens", ctypes.POINTER(ctypes.c_size_t)),
        ("pool_weights", ctypes.POINTER(ctypes.c_double)),
        ("pool_size", ctypes.c_uint32),
        ("pool_device", ctypes.POINTER(ctypes.c_void_p)),
        ("pool_expected", ctypes.POINTER(ctypes.c_void_p)),
        ("qps", ctypes.c_double),
        ("duration_s", ctypes.c_double),
        ("closed_loop", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("submit_threads", ctypes.c_uint32),
    ]


class LoadResult(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("failed", ctypes.c_uint64), ("verified", ctypes.c_uint64),
                ("mismatches", ctypes.c_uint64), ("wall_s", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_declared = False


def _lib():
    global _declared
    h = _decoder_lib()
    if not _declared:
        P = ctypes.c_void_p
        h.gst_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]
        h.gst_server_submit.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_uint64)]
        h.gst_server_poll.argtypes = [P, ctypes.POINTER(Completion), ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_uint32)]
        h.gst_server_drain.argtypes = [P]
        h.gst_server_release.argtypes = [P, ctypes.c_uint64]
        h.gst_server_stats_get.argtypes = [P, ctypes.POINTER(ServerStats)]
        h.gst_server_destroy.argtypes = [P]
        h.gst_server_submit_request.argtypes = [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
        h.gst_server_run_load.argtypes = [P, ctypes.POINTER(LoadSpec), P, P, ctypes.c_uint64,
                                          ctypes.POINTER(LoadResult)]
        _declared = True
    return h


class Server:
    def __init__(self, hbm_budget_bytes, host_staging_bytes=0, slots=0, max_batch_images=0,
                 max_batch_blocks=0, copy_to_host=False, auto_release=True, plain=False, dry_run=False,
                 groups_per_wave=0, waves_per_group=0, graph_launch=False, poison_outputs=False,
                 fail_launch_every=0, idle_spin_us=0, copy_kernel=False):
        cfg = ServerConfig(hbm_budget_bytes, host_staging_bytes, slots, max_batch_images, max_batch_blocks,
                           int(copy_to_host), int(auto_release), int(plain), int(dry_run), groups_per_wave,
                           waves_per_group, int(graph_launch), int(poison_outputs), fail_launch_every,
                           idle_spin_us, int(copy_kernel))
        self._h = ctypes.c_void_p()
        _check(_lib().gst_server_create(ctypes.byref(cfg), ctypes.byref(self._h)), "gst_server_create")

    def submit(self, data, tag=0):
        t = ctypes.c_uint64()
        data = bytes(data)
        _check(_lib().gst_server_submit(self._h, data, len(data), tag, ctypes.byref(t)), "gst_server_submit")
        return t.value

    def submit_request(self, data, tag=0, d_file=None, d_expected=None):
        """`data`: the whole .gtc on the host, or (with d_file, the file already in device
        memory) at least its header; `d_expected`: the expected output in device memory."""
        data = bytes(data)
        q = Request(data, d_file, len(data) if d_file is None else self._len_of(data), tag, d_expected)
        t = ctypes.c_uint64()
        _check(_lib().gst_server_submit_request(self._h, ctypes.byref(q), ctypes.byref(t)),
               "gst_server_submit_request")
        return t.value

    @staticmethod
    def _len_of(data):
        h = np.frombuffer(data[:28], "<u4")
        return 28 + 2048 + int(h[3]) + int(h[4]) + int(h[5]) + int(h[6])

    def poll(self, max_n=1024):
        buf = (Completion * max_n)()
        n = ctypes.c_uint32()
        _check(_lib().gst_server_poll(self._h, buf, max_n, ctypes.byref(n)), "gst_server_poll")
        return [buf[i] for i in range(n.value)]

    def drain(self):
        _check(_lib().gst_server_drain(self._h), "gst_server_drain")

    def release(self, ticket):
        _check(_lib().gst_server_release(self._h, ticket), "gst_server_release")

    def stats(self):
        st = ServerStats()
        _check(_lib().gst_server_stats_get(self._h, ctypes.byref(st)), "gst_server_stats_get")
        return st.as_dict()

    def run_load(self, pool, weights=None, qps=0.0, duration_s=1.0, closed_loop=0, seed=1, lat_cap=1 << 20,
                 pool_device=None, pool_expected=None, submit_threads=0):
        """Poisson (qps) or closed-loop load over `pool`; returns (device latencies,
        host latencies, LoadResult dict).  `pool_device`: the pool files already in device
        memory (pre-staged); `pool_expected`: their expected outputs in device memory, so
        every completion is compared on the device; `submit_threads`: closed loop only,
        threads that submit in parallel."""
        pool = [bytes(p) for p in pool]
        n = len(pool)
        ptrs = (ctypes.c_char_p * n)(*pool)
        lens = (ctypes.c_size_t * n)(*[len(p) for p in pool])
        w = (ctypes.c_double * n)(*weights) if weights is not None else None
        dev = (ctypes.c_void_p * n)(*pool_device) if pool_device is not None else None
        exp = (ctypes.c_void_p * n)(*pool_expected) if pool_expected is not None else None
        spec = LoadSpec(ptrs, lens, w, n, dev, exp, qps, duration_s, closed_loop, seed, submit_threads)
        lat_dev = np.zeros(lat_cap, np.float64)
        lat_host = np.zeros(lat_cap, np.float64)
        res = LoadResult()
        _check(_lib().gst_server_run_load(self._h, ctypes.byref(spec), lat_dev.ctypes.data, lat_host.ctypes.data,
                                          lat_cap, ctypes.byref(res)), "gst_server_run_load")
        k = min(res.requests, lat_cap)
        return lat_dev[:k], lat_host[:k], res.as_dict()

    def close(self):
        if self._h:
            _lib().gst_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server", "GstError"]
