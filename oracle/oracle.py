"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the reference's own host
tools built into oracle/_ref/ (test infrastructure only: tests/, smoke() and bench.py's
cpu_baseline leg use it as the checker; the product never does)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libref_host.so")
REF_GRADER = os.path.join(_HERE, "_ref", "hgrade")
NONE = 0xFFFFFFFF

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)

_lib = None
_ref = None


class OracleError(RuntimeError):
    """A failing oracle call: `status` is the hgp_status code it returned."""

    def __init__(self, status, where):
        super().__init__(f"{where} returned status {status}")
        self.status = status


def P(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB)
        _lib.orc_hg_load.restype = C.c_void_p
        _lib.orc_hg_load.argtypes = [C.c_char_p, C.c_int]
        _lib.orc_hg_sizes.argtypes = [C.c_void_p, u32p, u32p, u64p]
        _lib.orc_hg_copy.argtypes = [C.c_void_p, u64p, u32p, u32p, f32p]
        _lib.orc_hg_free.argtypes = [C.c_void_p]
        _lib.orc_dedup.restype = C.c_uint64
        _lib.orc_fixed_weight.restype = C.c_uint32
        _lib.orc_fixed_weight.argtypes = [C.c_float]
        _lib.orc_noise.restype = C.c_uint32
        _lib.orc_hash_uint32.restype = C.c_uint32
        _lib.orc_connectivity_fixed.restype = C.c_int64
        _lib.orc_connectivity_fixed.argtypes = [C.c_uint32, u64p, u32p, f32p, u32p]
        _lib.orc_initial_partition.argtypes = [C.c_uint32, C.c_uint32, u64p, u32p, f32p, u32p, C.c_uint32, C.c_float,
                                               C.c_uint32, C.c_int, u32p, u32p]
        _lib.orc_partition.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                                       C.c_uint32, C.c_int, C.c_int, C.c_int, u32p, C.c_uint32, C.c_uint32,
                                       C.c_uint32, u64p, u32p, u32p, f32p, u32p, u32p]
        _lib.orc_jacobi_round.argtypes = [C.c_uint32, C.c_uint32, u64p, u32p, u64p, u32p, u32p, f32p, u32p, C.c_uint32,
                                          C.c_uint32, C.c_uint32, u32p, u32p, u32p]
        _lib.orc_lp_refine.argtypes = [C.c_uint32, C.c_uint32, u64p, u32p, u64p, u32p, u32p, f32p, u32p, C.c_uint32,
                                       C.c_uint32, C.c_uint32, u32p, u32p]
        _lib.orc_bfd_start.argtypes = [C.c_uint32, u32p, C.c_uint32, C.c_uint32, C.c_uint64, u32p, u32p]
        _lib.orc_device_initial_partition.argtypes = [C.c_uint32, C.c_uint32, u64p, u32p, f32p, u32p, C.c_uint32,
                                                      C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                                      C.c_uint32, u32p, u32p]
        _lib.orc_metrics.argtypes = [C.c_uint32, C.c_uint32, u64p, u32p, u32p, f32p, u32p, f64p, f64p, f64p]
        _lib.orc_quotient.argtypes = [C.c_uint32, C.c_uint32, u64p, u32p, u32p, f32p, u32p, u32p, u32p, u64p, u64p,
                                      u32p, u32p, f32p]
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_LIB)
        _ref.ref_load.restype = C.c_void_p
        _ref.ref_load.argtypes = [C.c_char_p, C.c_int]
        _ref.ref_sizes.argtypes = [C.c_void_p, u32p, u32p, u64p]
        _ref.ref_copy.argtypes = [C.c_void_p, u64p, u32p, u32p, f32p]
        _ref.ref_free.argtypes = [C.c_void_p]
        _ref.ref_metrics.argtypes = [C.c_void_p, u32p, f64p, f64p, f64p]
        _ref.ref_validity.argtypes = [C.c_void_p, u32p, C.c_uint32, C.c_uint32, C.c_uint32]
        _ref.ref_save.argtypes = [C.c_void_p, u32p, C.c_char_p, C.c_int]
    return _ref


class HG:
    """Plain host CSR (n, offsets, pins, srcs_count, weights)."""

    def __init__(self, n, off, pins, srcc, w):
        self.n = int(n)
        self.off = np.ascontiguousarray(off, np.uint64)
        self.pins = np.ascontiguousarray(pins, np.uint32)
        self.srcc = np.ascontiguousarray(srcc, np.uint32)
        self.w = np.ascontiguousarray(w, np.float32)

    @property
    def e(self):
        return len(self.srcc)


def _load_with(lb, load, sizes, copy, free, path, dedup):
    h = load(path.encode(), 1 if dedup else 0)
    if not h:
        return None
    n, e, p = C.c_uint32(), C.c_uint32(), C.c_uint64()
    sizes(h, C.byref(n), C.byref(e), C.byref(p))
    off = np.zeros(e.value + 1, np.uint64)
    pins = np.zeros(p.value, np.uint32)
    srcc = np.zeros(e.value, np.uint32)
    w = np.zeros(e.value, np.float32)
    copy(h, P(off, u64p), P(pins, u32p), P(srcc, u32p), P(w, f32p))
    free(h)
    return HG(n.value, off, pins, srcc, w)


def load(path, dedup=True):
    L = lib()
    return _load_with(L, L.orc_hg_load, L.orc_hg_sizes, L.orc_hg_copy, L.orc_hg_free, path, dedup)


def ref_load(path, dedup=True):
    R = ref()
    return _load_with(R, R.ref_load, R.ref_sizes, R.ref_copy, R.ref_free, path, dedup)


def ref_metrics(path, part):
    R = ref()
    h = R.ref_load(path.encode(), 0)
    assert h
    part = np.ascontiguousarray(part, np.uint32)
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    rc = R.ref_metrics(h, P(part, u32p), C.byref(a), C.byref(b), C.byref(c))
    R.ref_free(h)
    return (a.value, b.value, c.value) if rc == 0 else None


def ref_validity(path, part, omega, delta, max_parts):
    R = ref()
    h = R.ref_load(path.encode(), 0)
    assert h
    part = np.ascontiguousarray(part, np.uint32)
    v = R.ref_validity(h, P(part, u32p), omega, delta, max_parts)
    R.ref_free(h)
    return v


def ref_save(path_in, part, path_out, which=0):
    """The reference's -s writer (which=0: quotient hypergraph in the format of path_out's
    extension) or -p writer (which=1) on the hypergraph loaded from path_in and deduplicated,
    as the reference's main saves it (sources/main.cu:125,919,933)."""
    R = ref()
    h = R.ref_load(path_in.encode(), 1)
    assert h
    part = np.ascontiguousarray(part, np.uint32)
    rc = R.ref_save(h, P(part, u32p), path_out.encode(), which)
    R.ref_free(h)
    return rc


def dedup(hg: HG) -> HG:
    off = hg.off.copy()
    pins = hg.pins.copy()
    srcc = hg.srcc.copy()
    np_ = lib().orc_dedup(hg.e, P(off, u64p), P(pins, u32p), P(srcc, u32p))
    return HG(hg.n, off, pins[:np_], srcc, hg.w)


def touching(hg: HG):
    p = len(hg.pins)
    toff = np.zeros(hg.n + 1, np.uint64)
    tch = np.zeros(p, np.uint32)
    inb = np.zeros(hg.n, np.uint32)
    lib().orc_build_touching(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(toff, u64p),
                             P(tch, u32p), P(inb, u32p))
    return toff, tch, inb


def neighbors(hg: HG, toff, tch):
    noff = np.zeros(hg.n + 1, np.uint64)
    lib().orc_build_neighbors(hg.n, P(hg.off, u64p), P(hg.pins, u32p), P(toff, u64p), P(tch, u32p), P(noff, u64p), None)
    nbr = np.zeros(int(noff[-1]), np.uint32)
    lib().orc_build_neighbors(hg.n, P(hg.off, u64p), P(hg.pins, u32p), P(toff, u64p), P(tch, u32p), P(noff, u64p),
                              P(nbr, u32p))
    return noff, nbr


def candidates(hg: HG, toff, tch, inb, sizes, cnc, omega, delta):
    pairs = np.zeros(hg.n * cnc, np.uint32)
    scores = np.zeros(hg.n * cnc, np.uint32)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    lib().orc_candidates(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(toff, u64p), P(tch, u32p),
                         P(inb, u32p), P(hg.w, f32p), P(sizes, u32p), cnc, omega, delta, P(pairs, u32p),
                         P(scores, u32p))
    return pairs, scores


def group_nodes(n, inb, pairs, scores, sizes, cnc, omega, delta):
    new_n = C.c_uint32()
    groups = np.zeros(n, np.uint32)
    gs = np.zeros(n, np.uint32)
    ung = np.zeros(n, np.uint32)
    uoff = np.zeros(n + 1, np.uint64)
    raw = np.zeros(n, np.uint32)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    lib().orc_group_nodes(n, P(inb, u32p), P(pairs, u32p), P(scores, u32p), P(sizes, u32p), cnc, omega, delta,
                          C.byref(new_n), P(groups, u32p), P(gs, u32p), P(ung, u32p), P(uoff, u64p), P(raw, u32p))
    k = new_n.value
    return k, groups, gs[:k], ung, uoff[:k + 1], raw


def coarsen_hedges(hg: HG, groups):
    choff = np.zeros(hg.e + 1, np.uint64)
    csrcc = np.zeros(hg.e, np.uint32)
    groups = np.ascontiguousarray(groups, np.uint32)
    lib().orc_coarsen_hedges(hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(groups, u32p),
                             P(choff, u64p), None, P(csrcc, u32p))
    cp = np.zeros(int(choff[-1]), np.uint32)
    lib().orc_coarsen_hedges(hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(groups, u32p),
                             P(choff, u64p), P(cp, u32p), P(csrcc, u32p))
    return choff, cp, csrcc


def coarsen_neighbors(n, noff, nbr, groups, new_n):
    cnoff = np.zeros(new_n + 1, np.uint64)
    groups = np.ascontiguousarray(groups, np.uint32)
    lib().orc_coarsen_neighbors(n, P(noff, u64p), P(nbr, u32p), P(groups, u32p), new_n, P(cnoff, u64p), None)
    cn = np.zeros(int(cnoff[-1]), np.uint32)
    lib().orc_coarsen_neighbors(n, P(noff, u64p), P(nbr, u32p), P(groups, u32p), new_n, P(cnoff, u64p), P(cn, u32p))
    return cnoff, cn


def refine_analyze(mode, omega, delta, repeats, hg: HG, toff, tch, inb, sizes, k, fm_repeat, chainup, part, psizes):
    n = hg.n
    moves = np.zeros(n, np.uint32)
    gains = np.zeros(n, np.int64)
    ranks = np.zeros(n, np.uint32)
    seq = np.zeros(n, np.int64)
    sv = np.zeros(n, np.int32)
    iv = np.zeros(n, np.int32)
    best = C.c_uint32()
    sizes = np.ascontiguousarray(sizes, np.uint32)
    part = np.ascontiguousarray(part, np.uint32)
    psizes = np.ascontiguousarray(psizes, np.uint32)
    lib().orc_refine_analyze(mode, omega, delta, repeats, n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p),
                             P(toff, u64p), P(tch, u32p), P(inb, u32p), P(hg.w, f32p), P(sizes, u32p), k, fm_repeat,
                             1 if chainup else 0, P(part, u32p), P(psizes, u32p), P(moves, u32p), P(gains, i64p),
                             P(ranks, u32p), P(seq, i64p), P(sv, i32p), P(iv, i32p), C.byref(best))
    return dict(moves=moves, gains=gains, ranks=ranks, seq=seq, sv=sv, iv=iv, best=best.value)


def refine(mode, omega, delta, repeats, hg: HG, toff, tch, inb, sizes, k, final_counts, part, psizes):
    part = np.ascontiguousarray(part, np.uint32).copy()
    psizes = np.ascontiguousarray(psizes, np.uint32).copy()
    pinb = np.zeros(k, np.uint32)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    lib().orc_refine(mode, omega, delta, repeats, hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p),
                     P(toff, u64p), P(tch, u32p), P(inb, u32p), P(hg.w, f32p), P(sizes, u32p), k,
                     1 if final_counts else 0, P(part, u32p), P(psizes, u32p), P(pinb, u32p))
    return part, psizes, pinb


def inbound_sizes(hg: HG, part, k):
    out = np.zeros(k, np.uint32)
    part = np.ascontiguousarray(part, np.uint32)
    lib().orc_inbound_sizes(hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(part, u32p), k, P(out, u32p))
    return out


def connectivity_fixed(hg: HG, part):
    part = np.ascontiguousarray(part, np.uint32)
    return lib().orc_connectivity_fixed(hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.w, f32p), P(part, u32p))


def greedy_merge_groups(n, ng, sizes, inb, omega, delta, groups, gsizes):
    groups = np.ascontiguousarray(groups, np.uint32).copy()
    gsizes = np.ascontiguousarray(gsizes, np.uint32).copy()
    gsizes = np.concatenate([gsizes, np.zeros(max(0, ng - len(gsizes)), np.uint32)])
    sizes = np.ascontiguousarray(sizes, np.uint32)
    new = C.c_uint32()
    lib().orc_greedy_merge_groups(n, ng, P(sizes, u32p), P(inb, u32p), omega, delta, P(groups, u32p), P(gsizes, u32p),
                                  C.byref(new))
    return new.value, groups, gsizes[:new.value]


def merge_small_partitions(n, k, psizes, pinb, omega, delta, part):
    part = np.ascontiguousarray(part, np.uint32).copy()
    psizes = np.ascontiguousarray(psizes, np.uint32)
    pinb = np.ascontiguousarray(pinb, np.uint32)
    lib().orc_merge_small_partitions(n, k, P(psizes, u32p), P(pinb, u32p), omega, delta, P(part, u32p))
    return part


def relabel(part):
    part = np.ascontiguousarray(part, np.uint32).copy()
    k = C.c_uint32()
    lib().orc_relabel(len(part), P(part, u32p), C.byref(k))
    return k.value, part


def initial_partition(hg: HG, sizes, k, eps, omega, prune=True):
    part = np.zeros(hg.n, np.uint32)
    ps = np.zeros(k, np.uint32)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    rc = lib().orc_initial_partition(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.w, f32p), P(sizes, u32p),
                                     k, eps, omega, int(prune), P(part, u32p), P(ps, u32p))
    if rc:
        raise OracleError(rc, "orc_initial_partition")
    return part, ps


def partition(hg: HG, mode, omega, delta, max_parts, kway=0, eps=0.0, cnc=16, repeats=16, ipm=False, prune=0,
              ip="auto", lp_rounds=0, ip_options=(0, 0, 0, 0)):
    """Whole pipeline.  ip: "auto" (the product's default: the device multistart partitioner's
    restatement on levels of at least 2^21 pins, the recursive bisection below), "device" or
    "host"; prune: the keep-target
    multiplier of the hedge pruning before a k-way initial partition (0: none); lp_rounds:
    label-propagation rounds before each level's FM rounds (k-way); ip_options: the device
    partitioner's (starts, rounds, nodes per part, refinement rounds), 0 = default."""
    opts = np.ascontiguousarray(ip_options, np.uint32)
    part = np.zeros(hg.n, np.uint32)
    lv = C.c_uint32()
    rc = lib().orc_partition(mode, omega, delta, max_parts, kway, eps, cnc, repeats, 1 if ipm else 0,
                             int(prune), {"host": 0, "device": 1, "auto": 2}[ip], P(opts, u32p), lp_rounds, hg.n, hg.e,
                             P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(hg.w, f32p), P(part, u32p),
                             C.byref(lv))
    if rc:
        raise OracleError(rc, "orc_partition")
    return part, lv.value


def jacobi_round(hg: HG, sizes, k, omega, rnd, part, psizes):
    """One Jacobi round (all pins destinations): (part, psizes, applied) after it."""
    toff, tch, inb = touching(hg)
    part = np.ascontiguousarray(part, np.uint32).copy()
    psizes = np.ascontiguousarray(psizes, np.uint32).copy()
    sizes = np.ascontiguousarray(sizes, np.uint32)
    applied = C.c_uint32()
    lib().orc_jacobi_round(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(toff, u64p), P(tch, u32p), P(inb, u32p),
                           P(hg.w, f32p), P(sizes, u32p), k, omega, rnd, P(part, u32p), P(psizes, u32p),
                           C.byref(applied))
    return part, psizes, applied.value


def lp_refine(hg: HG, sizes, k, omega, rounds, part, psizes):
    toff, tch, inb = touching(hg)
    part = np.ascontiguousarray(part, np.uint32).copy()
    psizes = np.ascontiguousarray(psizes, np.uint32).copy()
    sizes = np.ascontiguousarray(sizes, np.uint32)
    lib().orc_lp_refine(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(toff, u64p), P(tch, u32p), P(inb, u32p),
                        P(hg.w, f32p), P(sizes, u32p), k, omega, rounds, P(part, u32p), P(psizes, u32p))
    return part, psizes


def bfd_start(sizes, k, omega, seed):
    sizes = np.ascontiguousarray(sizes, np.uint32)
    part = np.zeros(len(sizes), np.uint32)
    loads = np.zeros(k, np.uint32)
    lib().orc_bfd_start(len(sizes), P(sizes, u32p), k, omega, seed, P(part, u32p), P(loads, u32p))
    return part, loads


def device_initial_partition(hg: HG, sizes, k, omega, prune=0, starts=2, rounds=8, per_part=16, repeats=4):
    part = np.zeros(hg.n, np.uint32)
    ps = np.zeros(k, np.uint32)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    lib().orc_device_initial_partition(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.w, f32p), P(sizes, u32p),
                                       k, omega, int(prune), starts, rounds, per_part, repeats, P(part, u32p),
                                       P(ps, u32p))
    return part, ps


def quotient(hg: HG, part):
    """The partitioned hypergraph (k, HG of the merged hedges); raises on non-sequential ids."""
    part = np.ascontiguousarray(part, np.uint32)
    k, e, p = C.c_uint32(), C.c_uint32(), C.c_uint64()
    args = (hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(hg.w, f32p), P(part, u32p),
            C.byref(k), C.byref(e), C.byref(p))
    rc = lib().orc_quotient(*args, None, None, None, None)
    if rc:
        raise OracleError(rc, "orc_quotient")
    off = np.zeros(e.value + 1, np.uint64)
    pins = np.zeros(p.value, np.uint32)
    srcc = np.zeros(e.value, np.uint32)
    w = np.zeros(e.value, np.float32)
    lib().orc_quotient(*args, P(off, u64p), P(pins, u32p), P(srcc, u32p), P(w, f32p))
    return k.value, HG(k.value, off, pins, srcc, w)


def validity(hg: HG, part, omega, delta, max_parts):
    part = np.ascontiguousarray(part, np.uint32)
    return lib().orc_check_validity(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(part, u32p),
                                    omega, delta, max_parts)


def metrics(hg: HG, part):
    part = np.ascontiguousarray(part, np.uint32)
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().orc_metrics(hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.srcc, u32p), P(hg.w, f32p), P(part, u32p),
                      C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def kway_omega(n, k, eps):
    """(uint32)ceil((1 + eps) * (float)n / k) in float32, as sources/runconfig.cpp:191."""
    v = (np.float32(1) + np.float32(eps)) * np.float32(n) / np.float32(k)
    return int(np.ceil(np.float32(v)))


def set_ip_threads(threads):
    """1: the initial partitioner stays on one thread; 0: its default (up to 16 threads)."""
    lib().orc_set_ip_threads(int(threads))


def chain_links(part, moves, gains, sizes):
    """Chaining on given moves/gains: (ranks, prev, comp, pos) per node (NONE = no predecessor)."""
    n = len(part)
    part = np.ascontiguousarray(part, np.uint32)
    moves = np.ascontiguousarray(moves, np.uint32)
    gains = np.ascontiguousarray(gains, np.int64)
    sizes = np.ascontiguousarray(sizes, np.uint32)
    ranks, prev, comp, pos = (np.zeros(n, np.uint32) for _ in range(4))
    lib().orc_chain_links(n, P(part, u32p), P(moves, u32p), P(gains, i64p), P(sizes, u32p), P(ranks, u32p),
                          P(prev, u32p), P(comp, u32p), P(pos, u32p))
    return ranks, prev, comp, pos


def ip_input(hg: HG, sizes, prune=True):
    """The initial partitioner's input restated by the oracle from a level (optional hedge
    pruning, merge of identical sets): (offsets, sorted pins, summed 2^18 weights, node weights)."""
    sizes = np.ascontiguousarray(sizes, np.uint32)
    e_out, p_out = C.c_uint32(), C.c_uint64()
    args = (hg.n, hg.e, P(hg.off, u64p), P(hg.pins, u32p), P(hg.w, f32p), P(sizes, u32p), int(prune),
            C.byref(e_out), C.byref(p_out))
    lib().orc_ip_input(*args, None, None, None, None)
    off = np.zeros(e_out.value + 1, np.uint64)
    pins = np.zeros(p_out.value, np.uint32)
    w = np.zeros(e_out.value, np.uint64)
    vw = np.zeros(hg.n, np.uint64)
    lib().orc_ip_input(*args, P(off, u64p), P(pins, u32p), P(w, u64p), P(vw, u64p))
    return off, pins, w, vw
