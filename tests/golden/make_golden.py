"""Writes the golden fixtures under tests/golden/ from the reference's own host code.

The reference ships no tests or known-answer vectors for this path.  Its host-side loaders,
rule-2 dedup, metrics and validity check (includes/hgraph.hpp, sources/constr.cpp,
grading/grade.cpp) do compile with g++ here; oracle/Makefile builds them from the reference
tree into oracle/_ref/ (libref_host.so, hgrade).  This script writes small seeded inputs in
every format the path reads (.hgr formats 0/1/10/11 with comments and singletons, .snn,
.axh with multi-source hedges, duplicate pins everywhere), runs the reference code on them,
and stores inputs plus expected outputs as fixtures:

  tests/golden/<case>.<ext>   the input file
  tests/golden/golden.json    per case: parsed CSR before/after dedup, and for a few seeded
                              partitions the reference's (lambda-1, cut-net, SOED) and its
                              validity verdicts under several constraint sets; plus the
                              "Connectivity:" line of the reference grader binary.

Run from the repository root after `make -C oracle` (needs the reference tree):
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import gen  # noqa: E402
import oracle as O  # noqa: E402


def write_hgr_vw(path, n, off, pins, w, vw, fmt):
    """hMETIS with vertex weights (fmt 10: vertex weights only, fmt 11: both)."""
    e = len(off) - 1
    lines = ["% golden fixture", f"{e} {n} {fmt}"]
    for i in range(e):
        ps = " ".join(str(int(p) + 1) for p in pins[off[i]:off[i + 1]])
        lines.append(f"{int(w[i])} {ps}" if fmt == 11 else ps)
    lines.extend(str(int(x)) for x in vw)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def cases(rng):
    out = []
    # .hgr fmt 0: duplicates inside hedges, comments, trailing singleton hedges
    n, off, pins, srcc, w = gen.random_hypergraph(40, 60, rng, max_size=6, dup_prob=0.3)
    p = os.path.join(HERE, "hgr_fmt0.hgr")
    gen.write_hgr(p, n, off, pins, None, comments=True, singletons=3)
    out.append(("hgr_fmt0", p))
    # .hgr fmt 1: integer hedge weights
    n, off, pins, srcc, w = gen.random_hypergraph(50, 80, rng, max_size=7, dup_prob=0.2, weights="int")
    p = os.path.join(HERE, "hgr_fmt1.hgr")
    gen.write_hgr(p, n, off, pins, w, comments=False)
    out.append(("hgr_fmt1", p))
    # .hgr fmt 10 and 11: vertex weights (ignored by the partitioner, parsed by the loader)
    n, off, pins, srcc, w = gen.random_hypergraph(30, 45, rng, max_size=5, dup_prob=0.2, weights="int")
    vw = rng.integers(1, 5, size=n)
    p = os.path.join(HERE, "hgr_fmt10.hgr")
    write_hgr_vw(p, n, off, pins, w, vw, 10)
    out.append(("hgr_fmt10", p))
    p = os.path.join(HERE, "hgr_fmt11.hgr")
    write_hgr_vw(p, n, off, pins, w, vw, 11)
    out.append(("hgr_fmt11", p))
    # .snn: one source per hedge, duplicate destinations and self-loops
    n, off, pins, srcc, w = gen.snn_like(4, 12, 3, rng)
    pins = pins.copy()
    for i in range(0, len(off) - 1, 5):  # a destination equal to the source, and a repeated one
        s, t = int(off[i]), int(off[i + 1])
        if t - s >= 3:
            pins[s + 1] = pins[s]
            pins[s + 2] = pins[t - 1]
    w = (rng.random(len(off) - 1) * 3 + 0.5).astype(np.float32)
    p = os.path.join(HERE, "net.snn")
    gen.write_snn(p, n, off, pins, w)
    out.append(("snn", p))
    # .axh: multi-source hedges, sources repeated among destinations
    n, off, pins, srcc, w = gen.random_hypergraph(36, 50, rng, max_size=7, src_prob=0.8, max_src=3, dup_prob=0.3,
                                                  weights="float")
    p = os.path.join(HERE, "multi.axh")
    gen.write_axh(p, n, off, pins, srcc, w)
    out.append(("axh", p))
    return out


def csr(hg):
    return {"n": int(hg.n), "off": [int(x) for x in hg.off], "pins": [int(x) for x in hg.pins],
            "srcc": [int(x) for x in hg.srcc], "w": [float(x) for x in hg.w]}


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libref_host.so is missing: run `make -C oracle` with the reference tree present")
    rng = np.random.default_rng(20261202)
    data = {}
    for name, path in cases(rng):
        raw = O.ref_load(path, dedup=False)
        ded = O.ref_load(path, dedup=True)
        entry = {"file": os.path.basename(path), "raw": csr(raw), "dedup": csr(ded), "partitions": []}
        for k in (2, 3, 5):
            part = rng.integers(0, k, size=ded.n).astype(np.uint32)
            met = O.ref_metrics(path, part)
            sizes = np.bincount(part, minlength=k)
            checks = []
            for omega, delta, maxp in ((int(sizes.max()), 2**31 - 1, k), (int(sizes.max()) - 1, 2**31 - 1, k),
                                       (int(sizes.max()), 1, k), (int(sizes.max()), 2**31 - 1, k - 1)):
                checks.append({"omega": omega, "delta": delta, "max_parts": maxp,
                               "valid": int(O.ref_validity(path, part, omega, delta, maxp))})
            entry["partitions"].append({"k": k, "part": [int(x) for x in part],
                                        "metrics": None if met is None else list(met), "validity": checks})
        data[name] = entry
    # the reference grader binary on one case: its printed connectivity
    part_file = os.path.join(HERE, "hgr_fmt1.part")
    part = np.array(data["hgr_fmt1"]["partitions"][1]["part"], np.uint32)
    with open(part_file, "w") as f:
        f.write("\n".join(str(int(x)) for x in part) + "\n")
    res = subprocess.run([O.REF_GRADER, "-r", os.path.join(HERE, "hgr_fmt1.hgr"), "-prt", part_file, "-k-prt", "3",
                          "0.5"], capture_output=True, text=True)
    lines = {ln.split(":")[0].strip(): ln.split(":", 1)[1].strip() for ln in res.stdout.splitlines() if ":" in ln}
    data["grader_hgr_fmt1"] = {"part_file": "hgr_fmt1.part", "args": ["-k-prt", "3", "0.5"],
                               "connectivity": float(lines["Connectivity"]), "cut_net": float(lines["Cut-net"]),
                               "soed": float(lines["SOED"])}
    # the reference's -s writers (quotient hypergraph by getPartitionsHypergraph(part,2,true)) in
    # all three formats, and its -p writer, for one seeded partition of every input format
    saved = {}
    for name in ("hgr_fmt1", "hgr_fmt10", "snn", "axh"):
        entry = data[name]
        # ids made sequential: the reference's quotient builder requires 0..N-1 (includes/hgraph.hpp:292-295)
        part = np.unique(np.array(entry["partitions"][2]["part"], np.uint32), return_inverse=True)[1].astype(np.uint32)
        src = os.path.join(HERE, entry["file"])
        outs = {}
        for ext in (".hgr", ".snn", ".axh"):
            out = os.path.join(HERE, "saved", f"{name}.quotient{ext}")
            os.makedirs(os.path.dirname(out), exist_ok=True)
            if O.ref_save(src, part, out, 0) == 0:
                outs[ext] = os.path.relpath(out, HERE)
            else:  # the reference refuses this output (the CLI exits 1): saveAXH takes no hedge without
                # sources (includes/hgraph.hpp:866-867), and no quotient hedge may lack destinations (:145)
                if os.path.exists(out):
                    os.remove(out)
                outs[ext] = None
        pout = os.path.join(HERE, "saved", f"{name}.part")
        assert O.ref_save(src, part, pout, 1) == 0
        outs["part"] = os.path.relpath(pout, HERE)
        saved[name] = {"part": [int(x) for x in part], "files": outs}
    data["saved"] = saved
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(data, f, indent=0)
    print("wrote", len(data), "fixture entries")


if __name__ == "__main__":
    main()
