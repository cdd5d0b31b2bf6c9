"""Thin wrappers that run each libhgp stage through the C ABI on numpy inputs (GPU tests)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import axoncuda_amd as H

V = C.c_void_p


def vp(x):
    return V(x) if x is not None else V()


class Level:
    """A hypergraph level resident in HBM (pins, offsets, source counts, weights)."""

    def __init__(self, dev, hg):
        self.dev = dev
        self.n = hg.n
        self.e = hg.e
        self.p = len(hg.pins)
        self.pins = dev.upload(hg.pins)
        self.off = dev.upload(hg.off)
        self.srcc = dev.upload(hg.srcc)
        self.w = dev.upload(hg.w)
        self.tch = self.toff = self.inb = None

    def build_touching(self):
        d = self.dev
        ts, t, o, i = C.c_uint64(), V(), V(), V()
        d.check(d.lib.hgp_build_touching(d.ctx, V(self.pins), V(self.off), V(self.srcc), self.n, self.e, C.byref(ts),
                                         C.byref(t), C.byref(o), C.byref(i)), "hgp_build_touching")
        self.tch, self.toff, self.inb, self.tsize = t.value, o.value, i.value, ts.value
        return (d.download(self.toff, self.n + 1, np.uint64), d.download(self.tch, self.tsize, np.uint32),
                d.download(self.inb, self.n, np.uint32))


def neighbors(dev, L: Level):
    ns, nb, no = C.c_uint64(), V(), V()
    dev.check(dev.lib.hgp_build_neighbors(dev.ctx, V(L.pins), V(L.off), V(L.tch), V(L.toff), L.n, C.byref(ns),
                                          C.byref(nb), C.byref(no)), "hgp_build_neighbors")
    noff = dev.download(no.value, L.n + 1, np.uint64)
    nbr = dev.download(nb.value, ns.value, np.uint32)
    return noff, nbr, nb.value, no.value


def candidates(dev, L: Level, sizes, cnc, omega, delta):
    d_sizes = dev.upload(np.asarray(sizes, np.uint32))
    d_pairs = dev.alloc(L.n * cnc * 4)
    d_scores = dev.alloc(L.n * cnc * 4)
    dev.check(dev.lib.hgp_candidates(dev.ctx, V(L.pins), V(L.off), V(L.srcc), V(L.tch), V(L.toff), V(L.inb), V(L.w),
                                     V(d_sizes), L.n, L.e, cnc, omega, delta, V(d_pairs), V(d_scores)),
              "hgp_candidates")
    pairs = dev.download(d_pairs, L.n * cnc, np.uint32)
    scores = dev.download(d_scores, L.n * cnc, np.uint32)
    return pairs, scores, d_sizes, d_pairs, d_scores


def group_nodes(dev, n, d_inb, d_pairs, d_scores, d_sizes, cnc, omega, delta):
    nn, g, gs, ug, uo = C.c_uint32(), V(), V(), V(), V()
    dev.check(dev.lib.hgp_group_nodes(dev.ctx, V(d_inb), V(d_pairs), V(d_scores), V(d_sizes), n, cnc, omega, delta,
                                      C.byref(nn), C.byref(g), C.byref(gs), C.byref(ug), C.byref(uo)),
              "hgp_group_nodes")
    k = nn.value
    return (k, dev.download(g.value, n, np.uint32), dev.download(gs.value, k, np.uint32),
            dev.download(ug.value, n, np.uint32), dev.download(uo.value, k + 1, np.uint64), g.value, gs.value,
            ug.value, uo.value)


def coarsen_hedges(dev, L: Level, d_groups):
    cs, ch, co, cc = C.c_uint64(), V(), V(), V()
    dev.check(dev.lib.hgp_coarsen_hedges(dev.ctx, V(L.pins), V(L.off), V(L.srcc), V(d_groups), L.e, L.p, C.byref(cs),
                                         C.byref(ch), C.byref(co), C.byref(cc)), "hgp_coarsen_hedges")
    return (dev.download(co.value, L.e + 1, np.uint64), dev.download(ch.value, cs.value, np.uint32),
            dev.download(cc.value, L.e, np.uint32), ch.value, co.value, cc.value)


def coarsen_touching(dev, d_ch, d_co, d_cc, new_n, e):
    ts, t, o, i = C.c_uint64(), V(), V(), V()
    dev.check(dev.lib.hgp_coarsen_touching(dev.ctx, V(d_ch), V(d_co), V(d_cc), new_n, e, C.byref(ts), C.byref(t),
                                           C.byref(o), C.byref(i)), "hgp_coarsen_touching")
    return (dev.download(o.value, new_n + 1, np.uint64), dev.download(t.value, ts.value, np.uint32),
            dev.download(i.value, new_n, np.uint32))


def coarsen_neighbors(dev, d_nbr, d_noff, d_groups, n, new_n):
    ns, nb, no = C.c_uint64(), V(), V()
    dev.check(dev.lib.hgp_coarsen_neighbors(dev.ctx, V(d_nbr), V(d_noff), V(d_groups), V(), V(), n, new_n,
                                            C.byref(ns), C.byref(nb), C.byref(no)), "hgp_coarsen_neighbors")
    return dev.download(no.value, new_n + 1, np.uint64), dev.download(nb.value, ns.value, np.uint32)


def params(mode, omega, delta, max_parts, repeats=16, cnc=16, kway=0, eps=0.0):
    p = H.Params()
    H.load_library().hgp_default_params(C.byref(p))
    p.mode = mode
    p.max_nodes_per_part = omega
    p.max_inbound_per_part = delta
    p.max_parts = max_parts
    p.refine_repeats = repeats
    p.candidates_count = cnc
    p.kway = kway
    p.epsilon = eps
    return p


def refine_analyze(dev, L: Level, sizes, k, mode, omega, delta, repeats, fm_repeat, chainup, part, psizes):
    n = L.n
    prm = params(mode, omega, delta, k, repeats)
    d_sizes = dev.upload(np.asarray(sizes, np.uint32))
    d_part = dev.upload(np.asarray(part, np.uint32))
    d_ps = dev.upload(np.asarray(psizes, np.uint32))
    out = dict(moves=np.zeros(n, np.uint32), gains=np.zeros(n, np.int64), ranks=np.zeros(n, np.uint32),
               seq=np.zeros(n, np.int64), sv=np.zeros(n, np.int32), iv=np.zeros(n, np.int32))
    best = C.c_uint32()
    u32p, i64p, i32p = C.POINTER(C.c_uint32), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    dev.check(dev.lib.hgp_refine_analyze(
        dev.ctx, C.byref(prm), V(L.pins), V(L.off), V(L.srcc), V(L.tch), V(L.toff), V(L.inb), V(L.w), V(d_sizes), n,
        L.e, k, fm_repeat, 1 if chainup else 0, V(d_part), V(d_ps), out["moves"].ctypes.data_as(u32p),
        out["gains"].ctypes.data_as(i64p), out["ranks"].ctypes.data_as(u32p), out["seq"].ctypes.data_as(i64p),
        out["sv"].ctypes.data_as(i32p), out["iv"].ctypes.data_as(i32p), C.byref(best)), "hgp_refine_analyze")
    for x in (d_sizes, d_part, d_ps):
        dev.free(x)
    out["best"] = best.value
    return out


def refine(dev, L: Level, sizes, k, mode, omega, delta, repeats, final, part, psizes):
    prm = params(mode, omega, delta, k, repeats)
    d_sizes = dev.upload(np.asarray(sizes, np.uint32))
    d_part = dev.upload(np.asarray(part, np.uint32))
    d_ps = dev.upload(np.asarray(psizes, np.uint32))
    d_pin = dev.alloc(k * 4)
    dev.check(dev.lib.hgp_refine(dev.ctx, C.byref(prm), V(L.pins), V(L.off), V(L.srcc), V(L.tch), V(L.toff), V(L.inb),
                                 V(L.w), V(d_sizes), 0, L.n, L.e, k, 1 if final else 0, V(d_part), V(d_ps), V(d_pin)),
              "hgp_refine")
    res = (dev.download(d_part, L.n, np.uint32), dev.download(d_ps, k, np.uint32), dev.download(d_pin, k, np.uint32))
    for x in (d_sizes, d_part, d_ps, d_pin):
        dev.free(x)
    return res


def refine_lp(dev, L: Level, sizes, k, omega, rounds, part, psizes):
    prm = params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k)
    prm.lp_rounds = rounds
    d_sizes = dev.upload(np.asarray(sizes, np.uint32))
    d_part = dev.upload(np.asarray(part, np.uint32))
    d_ps = dev.upload(np.asarray(psizes, np.uint32))
    dev.check(dev.lib.hgp_refine_lp(dev.ctx, C.byref(prm), V(L.pins), V(L.off), V(L.srcc), V(L.tch), V(L.toff),
                                    V(L.inb), V(L.w), V(d_sizes), L.n, L.e, k, V(d_part), V(d_ps)), "hgp_refine_lp")
    res = dev.download(d_part, L.n, np.uint32), dev.download(d_ps, k, np.uint32)
    for x in (d_sizes, d_part, d_ps):
        dev.free(x)
    return res


def partition(dev, hgraph: H.Hypergraph, prm):
    n, _, _ = hgraph.sizes()
    part = np.zeros(n, np.uint32)
    st = H.RunStats()
    dev.check(dev.lib.hgp_partition(dev.ctx, C.byref(prm), hgraph.h, part.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    C.byref(st)), "hgp_partition")
    return part, st
