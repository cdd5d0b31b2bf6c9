"""ctypes binding of libhgp, the MI355X-native hypergraph partitioner (C ABI in include/hgp.h).

The library is native C++/HIP; this module only loads it and moves numpy arrays in and
out of device memory for tests and benchmarks.  There is no Python or CPU fallback: if the
shared library is missing, importing any entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhgp.so")
BIN_DIR = os.path.join(_HERE, "bin")

HGP_OK = 0
HGP_MODE_INCC = 0
HGP_MODE_KWAY = 1
NONE = 0xFFFFFFFF

_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p


HGP_IP_HOST, HGP_IP_DEVICE, HGP_IP_AUTO = 0, 1, 2


class Params(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("max_nodes_per_part", C.c_uint32),
        ("max_inbound_per_part", C.c_uint32),
        ("max_parts", C.c_uint32),
        ("kway", C.c_uint32),
        ("epsilon", C.c_float),
        ("candidates_count", C.c_uint32),
        ("refine_repeats", C.c_uint32),
        ("save_memory_up_to_level", C.c_uint32),
        ("oversized_multiplier", C.c_float),
        ("device_touching_construction", C.c_int32),
        ("initial_partitions_merge", C.c_int32),
        ("verbose", C.c_int32),
        ("initial_partitioner", C.c_int32),
        ("initial_prune", C.c_int32),
        ("lp_rounds", C.c_uint32),
        ("ip_starts", C.c_uint32),
        ("ip_rounds", C.c_uint32),
        ("ip_nodes_per_part", C.c_uint32),
        ("ip_refine_repeats", C.c_uint32),
        ("refine_event_budget", C.c_uint64),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("core_ms", C.c_double),
        ("device_ms", C.c_double),
        ("ip_ms", C.c_double),
        ("levels", C.c_uint32),
        ("num_partitions", C.c_uint32),
        ("total_pins", C.c_uint64),
        ("bytes_algorithmic", C.c_uint64),
        ("level_nodes", C.c_uint32 * 64),
        ("level_pins", C.c_uint64 * 64),
        ("peak_device_bytes", C.c_uint64),
    ]


class Tuning(C.Structure):
    """hgp_tuning: launch parameters chosen by the autotuner (speed only, never results)."""
    _fields_ = [
        ("sort_rounds", C.c_uint32 * 40),
        ("heavy_table_slots", C.c_uint32),
        ("heavy_table_fill16", C.c_uint32),
        ("cascade_block", C.c_uint32),
        ("scan_lookback_tiles", C.c_uint32),
        ("mask_wide4_degree", C.c_uint32),
        ("mask_wide16_degree", C.c_uint32),
        ("side_streams", C.c_uint32),
        ("coop_max_n", C.c_uint32),
        ("coop_tile", C.c_uint32),
        ("tuned", C.c_uint32),
    ]

    def as_dict(self):
        return {"sort_rounds": {i: int(r) for i, r in enumerate(self.sort_rounds) if r},
                "heavy_table_slots": int(self.heavy_table_slots), "heavy_table_fill16": int(self.heavy_table_fill16),
                "cascade_block": int(self.cascade_block), "scan_lookback_tiles": int(self.scan_lookback_tiles),
                "mask_wide4_degree": int(self.mask_wide4_degree), "mask_wide16_degree": int(self.mask_wide16_degree),
                "side_streams": int(self.side_streams), "coop_max_n": int(self.coop_max_n),
                "coop_tile": int(self.coop_tile), "tuned": int(self.tuned)}


# (name, restype, argtypes) of every exported entry point of include/hgp.h
SIGNATURES = [
    ("hgp_default_params", None, [C.POINTER(Params)]),
    ("hgp_ctx_create", C.c_int32, [C.c_int, C.c_uint64, C.POINTER(_vp)]),
    ("hgp_ctx_destroy", None, [_vp]),
    ("hgp_ctx_last_error", C.c_char_p, [_vp]),
    ("hgp_ctx_set_stream", C.c_int32, [_vp, _vp]),
    ("hgp_synchronize", C.c_int32, [_vp]),
    ("hgp_malloc", C.c_int32, [_vp, C.c_uint64, C.POINTER(_vp)]),
    ("hgp_free", C.c_int32, [_vp, _vp]),
    ("hgp_memcpy_htod", C.c_int32, [_vp, _vp, _vp, C.c_uint64]),
    ("hgp_memcpy_dtoh", C.c_int32, [_vp, _vp, _vp, C.c_uint64]),
    ("hgp_memcpy_dtod", C.c_int32, [_vp, _vp, _vp, C.c_uint64]),
    ("hgp_ctx_memory", C.c_int32, [_vp, _u64p, _u64p, _u64p]),
    ("hgp_ctx_trim", C.c_int32, [_vp]),
    ("hgp_job_footprint", C.c_uint64, [_vp, C.POINTER(Params)]),
    ("hgp_hgraph_load", C.c_int32, [C.c_char_p, C.c_int, C.POINTER(_vp)]),
    ("hgp_hgraph_from_csr", C.c_int32, [C.c_uint32, C.c_uint32, _u64p, _u32p, _u32p, C.POINTER(C.c_float), C.c_int,
                                        C.POINTER(_vp)]),
    ("hgp_hgraph_free", None, [_vp]),
    ("hgp_hgraph_sizes", None, [_vp, _u32p, _u32p, _u64p]),
    ("hgp_hgraph_copy_out", C.c_int32, [_vp, _u64p, _u32p, _u32p, C.POINTER(C.c_float)]),
    ("hgp_save_partitioning", C.c_int32, [C.c_char_p, _u32p, C.c_uint32]),
    ("hgp_save_partitioned", C.c_int32, [_vp, _u32p, C.c_char_p]),
    ("hgp_hgraph_metrics", C.c_int32, [_vp, _u32p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]),
    ("hgp_check_validity", C.c_int32, [_vp, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, _i32p]),
    ("hgp_gen_random", C.c_int32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(_vp)]),
    ("hgp_gen_netlist", C.c_int32, [C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32, C.c_double,
                                    C.c_uint64, C.POINTER(_vp)]),
    ("hgp_gen_snn", C.c_int32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64,
                                C.POINTER(_vp)]),
    ("hgp_hgraph_save", C.c_int32, [_vp, C.c_char_p]),
    ("hgp_build_touching", C.c_int32, [_vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _u64p, C.POINTER(_vp),
                                       C.POINTER(_vp), C.POINTER(_vp)]),
    ("hgp_build_neighbors", C.c_int32, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, _u64p, C.POINTER(_vp),
                                        C.POINTER(_vp)]),
    ("hgp_candidates", C.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, _vp, _vp]),
    ("hgp_group_nodes", C.c_int32, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
                                    C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    ("hgp_coarsen_neighbors", C.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _u64p,
                                          C.POINTER(_vp), C.POINTER(_vp)]),
    ("hgp_coarsen_hedges", C.c_int32, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint64, _u64p, C.POINTER(_vp),
                                       C.POINTER(_vp), C.POINTER(_vp)]),
    ("hgp_coarsen_touching", C.c_int32, [_vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _u64p, C.POINTER(_vp),
                                         C.POINTER(_vp), C.POINTER(_vp)]),
    ("hgp_uncoarsen", C.c_int32, [_vp, _vp, _vp, C.c_uint32, _vp]),
    ("hgp_refine", C.c_int32, [_vp, C.POINTER(Params), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32,
                               C.c_uint32, C.c_uint32, C.c_int32, _vp, _vp, _vp]),
    ("hgp_refine_analyze", C.c_int32, [_vp, C.POINTER(Params), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32,
                                       C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, _vp, _vp, _u32p, _i64p, _u32p,
                                       _i64p, _i32p, _i32p, _u32p]),
    ("hgp_greedy_merge_groups", C.c_int32, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _vp, _vp,
                                            _u32p]),
    ("hgp_merge_small_partitions", C.c_int32, [_vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _vp]),
    ("hgp_relabel_partitions", C.c_int32, [_vp, C.c_uint32, _vp, _u32p]),
    ("hgp_inbound_sizes", C.c_int32, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp]),
    ("hgp_connectivity", C.c_int32, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, _i64p]),
    ("hgp_quotient", C.c_int32, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _u32p, _u32p, _u64p,
                                 C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    ("hgp_autotune", C.c_int32, [_vp, C.c_uint32, C.POINTER(Tuning)]),
    ("hgp_tuning_get", C.c_int32, [_vp, C.POINTER(Tuning)]),
    ("hgp_tuning_set", C.c_int32, [_vp, C.POINTER(Tuning)]),
    ("hgp_tuning_defaults", None, [C.POINTER(Tuning)]),
    ("hgp_tuning_set_process_default", C.c_int32, [C.POINTER(Tuning)]),
    ("hgp_tuning_save", C.c_int32, [C.POINTER(Tuning), C.c_char_p]),
    ("hgp_tuning_load", C.c_int32, [C.c_char_p, C.POINTER(Tuning)]),
    ("hgp_connectivity_noalloc", C.c_int32, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, _vp]),
    ("hgp_refine_lp", C.c_int32, [_vp, C.POINTER(Params), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32,
                                  C.c_uint32, _vp, _vp]),
    ("hgp_initial_partition", C.c_int32, [_vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_uint64, _vp, C.c_uint32,
                                          C.c_float, C.c_uint32, C.c_int32, C.POINTER(_vp), C.POINTER(_vp)]),
    ("hgp_ip_input", C.c_int32, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_uint32, C.c_int32, _u32p, _u64p, _u64p,
                                 _u32p, _u64p]),
    ("hgp_partition", C.c_int32, [_vp, C.POINTER(Params), _vp, _u32p, C.POINTER(RunStats)]),
    ("hgp_profile_kernel", C.c_int32, [_vp, C.c_char_p]),
    ("hgp_profile_count", C.c_int32, [_vp, C.c_int32]),
    ("hgp_profile_read", C.c_int32, [_vp, C.POINTER(C.c_double), _u64p, _u64p]),
    ("hgp_server_create", C.c_int32, [C.c_int, C.c_uint64, C.c_uint32, C.POINTER(_vp)]),
    ("hgp_server_submit", C.c_int32, [_vp, _vp, C.POINTER(Params), _u64p]),
    ("hgp_server_wait", C.c_int32, [_vp, C.c_uint64, _u32p, C.POINTER(RunStats), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    ("hgp_server_destroy", None, [_vp]),
    ("hgp_partition_device", C.c_int32, [_vp, C.POINTER(Params), C.c_uint32, C.c_uint32, _vp, _vp, _vp, _vp, _u32p,
                                         C.POINTER(RunStats)]),
]

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libhgp.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"libhgp.so not found at {path}: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        lib = C.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class HgpError(RuntimeError):
    pass


class Hypergraph:
    """Host hypergraph model (hgp_hgraph) owned by the library."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def load(cls, path: str, dedup: bool = True) -> "Hypergraph":
        lib = load_library()
        out = _vp()
        st = lib.hgp_hgraph_load(path.encode(), 1 if dedup else 0, C.byref(out))
        if st != HGP_OK:
            raise HgpError(f"hgp_hgraph_load({path}) failed with status {st}")
        return cls(out)

    @classmethod
    def from_csr(cls, n, offsets, pins, srcs_count=None, weights=None, dedup=True) -> "Hypergraph":
        lib = load_library()
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        pins = np.ascontiguousarray(pins, dtype=np.uint32)
        e = len(offsets) - 1
        srcs_count = np.zeros(e, np.uint32) if srcs_count is None else np.ascontiguousarray(srcs_count, np.uint32)
        weights = np.ones(e, np.float32) if weights is None else np.ascontiguousarray(weights, np.float32)
        out = _vp()
        st = lib.hgp_hgraph_from_csr(n, e, _ptr(offsets, _u64p), _ptr(pins, _u32p), _ptr(srcs_count, _u32p),
                                     _ptr(weights, C.POINTER(C.c_float)), 1 if dedup else 0, C.byref(out))
        if st != HGP_OK:
            raise HgpError(f"hgp_hgraph_from_csr failed with status {st}")
        return cls(out)

    @classmethod
    def generate(cls, kind: str, **kw) -> "Hypergraph":
        lib = load_library()
        out = _vp()
        if kind == "random":
            st = lib.hgp_gen_random(kw["n"], kw["e"], kw.get("max_size", 64), kw.get("seed", 1), C.byref(out))
        elif kind == "netlist":
            st = lib.hgp_gen_netlist(kw["n"], kw["e"], kw.get("alpha", 2.5), kw.get("max_size", 1000),
                                     kw.get("radius", 8), kw.get("frac_local", 0.8), kw.get("seed", 1), C.byref(out))
        elif kind == "snn":
            st = lib.hgp_gen_snn(kw["layers"], kw["per_layer"], kw["grid"], kw["fanout"], kw["sigma"],
                                 kw.get("seed", 1), C.byref(out))
        else:
            raise ValueError(kind)
        if st != HGP_OK:
            raise HgpError(f"generator {kind} failed with status {st}")
        return cls(out)

    def sizes(self):
        lib = load_library()
        n, e, p = C.c_uint32(), C.c_uint32(), C.c_uint64()
        lib.hgp_hgraph_sizes(self.h, C.byref(n), C.byref(e), C.byref(p))
        return n.value, e.value, p.value

    def arrays(self):
        lib = load_library()
        n, e, p = self.sizes()
        off = np.zeros(e + 1, np.uint64)
        pins = np.zeros(p, np.uint32)
        srcc = np.zeros(e, np.uint32)
        w = np.zeros(e, np.float32)
        lib.hgp_hgraph_copy_out(self.h, _ptr(off, _u64p), _ptr(pins, _u32p), _ptr(srcc, _u32p),
                                _ptr(w, C.POINTER(C.c_float)))
        return n, off, pins, srcc, w

    def save(self, path: str):
        st = load_library().hgp_hgraph_save(self.h, path.encode())
        if st != HGP_OK:
            raise HgpError(f"save {path} failed: {st}")

    def metrics(self, part: np.ndarray):
        lib = load_library()
        part = np.ascontiguousarray(part, np.uint32)
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        lib.hgp_hgraph_metrics(self.h, _ptr(part, _u32p), C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def validity(self, part, omega, delta, max_parts) -> int:
        lib = load_library()
        part = np.ascontiguousarray(part, np.uint32)
        v = C.c_int32()
        lib.hgp_check_validity(self.h, _ptr(part, _u32p), omega, delta, max_parts, C.byref(v))
        return v.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.hgp_hgraph_free(self.h)
            self.h = None


class Device:
    """A libhgp context on one GPU plus numpy <-> device helpers."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = load_library()
        ctx = _vp()
        st = self.lib.hgp_ctx_create(device, hbm_budget, C.byref(ctx))
        if st != HGP_OK:
            raise HgpError(f"hgp_ctx_create failed with status {st} (no gfx950 device?)")
        self.ctx = ctx
        self._owned = []

    def check(self, st, what):
        if st != HGP_OK:
            raise HgpError(f"{what} failed with status {st}: {self.lib.hgp_ctx_last_error(self.ctx).decode()}")

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = _vp()
        self.check(self.lib.hgp_malloc(self.ctx, max(a.nbytes, 1), C.byref(p)), "hgp_malloc")
        if a.nbytes:
            self.check(self.lib.hgp_memcpy_htod(self.ctx, p, a.ctypes.data_as(_vp), a.nbytes), "htod")
        return p.value

    def alloc(self, nbytes: int) -> int:
        p = _vp()
        self.check(self.lib.hgp_malloc(self.ctx, max(nbytes, 1), C.byref(p)), "hgp_malloc")
        return p.value

    def download(self, ptr, count: int, dtype) -> np.ndarray:
        out = np.zeros(count, dtype)
        if count:
            self.check(self.lib.hgp_memcpy_dtoh(self.ctx, out.ctypes.data_as(_vp), _vp(ptr), out.nbytes), "dtoh")
        return out

    def free(self, ptr):
        if ptr:
            self.lib.hgp_free(self.ctx, _vp(ptr))

    def close(self):
        if self.ctx:
            self.lib.hgp_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KernelProfile:
    """Times every launch of one kernel (HIP events on the library's stream) from now on;
    with count_bytes the launches also count their algorithmic bytes (extra device work)."""

    def __init__(self, dev: Device, kernel: str, count_bytes: bool = False):
        self.dev = dev
        dev.check(dev.lib.hgp_profile_read(dev.ctx, None, None, None), "hgp_profile_read")
        dev.check(dev.lib.hgp_profile_kernel(dev.ctx, kernel.encode()), "hgp_profile_kernel")
        dev.check(dev.lib.hgp_profile_count(dev.ctx, 1 if count_bytes else 0), "hgp_profile_count")

    def read(self) -> dict:
        ms, n, b = C.c_double(), C.c_uint64(), C.c_uint64()
        self.dev.check(self.dev.lib.hgp_profile_read(self.dev.ctx, C.byref(ms), C.byref(n), C.byref(b)),
                       "hgp_profile_read")
        self.dev.check(self.dev.lib.hgp_profile_kernel(self.dev.ctx, None), "hgp_profile_kernel")
        self.dev.check(self.dev.lib.hgp_profile_count(self.dev.ctx, 0), "hgp_profile_count")
        return {"ms": ms.value, "launches": n.value, "bytes": b.value}


class Server:
    """Continuous-batching partition service (hgp_server_*): submit jobs, wait for results."""

    def __init__(self, device: int = 0, workers: int = 4, hbm_budget: int = 0):
        self.lib = load_library()
        h = _vp()
        st = self.lib.hgp_server_create(device, hbm_budget, workers, C.byref(h))
        if st != HGP_OK:
            raise HgpError(f"hgp_server_create failed with status {st}")
        self.h = h
        self._keep = {}

    def submit(self, hg: Hypergraph, prm: Params) -> int:
        jid = C.c_uint64()
        st = self.lib.hgp_server_submit(self.h, hg.h, C.byref(prm), C.byref(jid))
        if st != HGP_OK:
            raise HgpError(f"hgp_server_submit failed with status {st}")
        self._keep[jid.value] = hg  # the hypergraph must outlive the job
        return jid.value

    def wait(self, job_id: int):
        hg = self._keep[job_id]
        n, _, _ = hg.sizes()
        part = np.zeros(n, np.uint32)
        st = RunStats()
        lat, q = C.c_double(), C.c_double()
        s = self.lib.hgp_server_wait(self.h, job_id, _ptr(part, _u32p), C.byref(st), C.byref(lat), C.byref(q))
        del self._keep[job_id]
        if s != HGP_OK:
            raise HgpError(f"job {job_id} failed with status {s}")
        return part, st, lat.value, q.value

    def close(self):
        if self.h:
            self.lib.hgp_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
