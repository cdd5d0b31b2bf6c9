"""Seeded numpy generators and file writers for test hypergraphs (independent of libhgp)."""
from __future__ import annotations

import os
import struct

import numpy as np


def random_hypergraph(n, e, rng, max_size=8, src_prob=0.0, max_src=2, dup_prob=0.0, weights="unit", min_size=2):
    """Returns (n, off, pins, srcc, w): sources first in each hedge; optional duplicate pins
    (within a hedge, and sources repeated among destinations) to exercise rule-2 dedup."""
    offs = [0]
    pins = []
    srcc = []
    for _ in range(e):
        sz = int(rng.integers(min_size, max_size + 1))
        sz = min(sz, n)
        nodes = list(rng.choice(n, size=sz, replace=False))
        ns = 0
        if src_prob > 0 and rng.random() < src_prob:
            ns = int(rng.integers(1, max_src + 1))
            ns = min(ns, sz - 1)
        if dup_prob > 0 and rng.random() < dup_prob:
            nodes.append(nodes[int(rng.integers(0, len(nodes)))])
        pins.extend(int(x) for x in nodes)
        offs.append(len(pins))
        srcc.append(ns)
    if weights == "unit":
        w = np.ones(e, np.float32)
    elif weights == "int":
        w = rng.integers(1, 6, size=e).astype(np.float32)
    else:
        w = (rng.random(e) * 4 + 0.25).astype(np.float32)
    return n, np.array(offs, np.uint64), np.array(pins, np.uint32), np.array(srcc, np.uint32), w


def write_hgr(path, n, off, pins, w=None, comments=True, singletons=0):
    """hMETIS writer (all pins as listed, 1-based); optional weights (fmt 1)."""
    e = len(off) - 1
    lines = []
    if comments:
        lines.append("% test hypergraph")
    fmt = " 1" if w is not None else ""
    lines.append(f"{e + singletons} {n}{fmt}")
    for i in range(e):
        ps = " ".join(str(int(p) + 1) for p in pins[off[i]:off[i + 1]])
        if w is not None:
            lines.append(f"{int(w[i])} {ps}")
        else:
            lines.append(ps)
        if comments and i == 0:
            lines.append("   % an indented comment")
    for s in range(singletons):
        lines.append(("1 " if w is not None else "") + str(s + 1))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_snn(path, n, off, pins, w):
    with open(path, "wb") as f:
        e = len(off) - 1
        f.write(struct.pack("<II", n, e))
        for i in range(e):
            seg = pins[off[i]:off[i + 1]]
            f.write(struct.pack("<II", len(seg) - 1, int(seg[0])))
            f.write(np.asarray(seg[1:], np.uint32).tobytes())
            f.write(struct.pack("<f", float(w[i])))


def write_axh(path, n, off, pins, srcc, w):
    with open(path, "wb") as f:
        e = len(off) - 1
        f.write(struct.pack("<II", n, e))
        for i in range(e):
            seg = pins[off[i]:off[i + 1]]
            f.write(struct.pack("<II", len(seg), max(int(srcc[i]), 1)))
            f.write(np.asarray(seg, np.uint32).tobytes())
            f.write(struct.pack("<f", float(w[i])))


def snn_like(layers, per_layer, fanout, rng):
    """Small layered feed-forward net: every neuron of layers 0..L-2 is a hedge with one source."""
    offs = [0]
    pins = []
    srcc = []
    for layer in range(layers - 1):
        for i in range(per_layer):
            src = layer * per_layer + i
            lo = max(0, i - fanout)
            hi = min(per_layer, i + fanout + 1)
            cand = np.arange(lo, hi)
            dst = rng.choice(cand, size=min(fanout, len(cand)), replace=False) + (layer + 1) * per_layer
            pins.append(src)
            pins.extend(int(d) for d in dst)
            offs.append(len(pins))
            srcc.append(1)
    n = layers * per_layer
    e = len(srcc)
    return n, np.array(offs, np.uint64), np.array(pins, np.uint32), np.array(srcc, np.uint32), np.ones(e, np.float32)


def tmp_path_for(tmpdir, name):
    return os.path.join(str(tmpdir), name)
