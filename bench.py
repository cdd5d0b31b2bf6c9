"""Benchmark of the hypergraph partitioner's hot path on one MI355X.

Workload: netlist-shaped hypergraph, 3.2M nodes and 3.4M nets, net sizes from a discrete
power law (alpha 2.5 on [2, 1000]), 80% of a net's pins within radius 8 of its first pin on a
2D grid and 20% uniform, partitioned k-way with k=64 and epsilon 0.03.  One step is one full
partitioning (coarsening, initial partition, uncoarsening + refinement) of that hypergraph
starting from device-resident CSR arrays; the partition vector lands in host memory.

value = post-dedup pins / second of one step (the reference's "device core" boundary,
sources/main.cu:260-264,891-896), with inputs already in HBM.

The roofline object covers the dominant kernel: its launches are bracketed by HIP events on
the library's stream during the timed steps, and its algorithmic bytes per launch come from
the library's byte model of that kernel.  cpu_baseline is the sequential oracle run on a
bounded sample of the same workload (50K nodes) on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import axoncuda_amd as H  # noqa: E402

N_NODES, N_NETS, ALPHA, MAX_SIZE, RADIUS, FRAC_LOCAL = 3_200_000, 3_400_000, 2.5, 1000, 8, 0.8
K, EPS = 64, 0.03
HBM_PEAK_GBS = 8000.0
# Kernel whose launches the roofline object covers (the largest share of device time on
# this workload in the kernel-trace summary under profiles/).
ROOFLINE_KERNEL = os.environ.get("HGP_BENCH_KERNEL", "k_cand_light")
# Plain copy calibration of FETCH_SIZE / WRITE_SIZE (bytes counted per algorithmic byte),
# measured with rocprofv3 --pmc on a copy kernel of known size; None until measured.
TRAFFIC_PER_LAUNCH = None


def kway_omega(n: int, k: int, eps: float) -> int:
    """(uint32)ceil((1 + eps) * (float)n / k) evaluated in float32 (sources/runconfig.cpp:191)."""
    v = (np.float32(1) + np.float32(eps)) * np.float32(n) / np.float32(k)
    return int(np.ceil(np.float32(v)))


def kway_params(n: int, k: int, eps: float) -> H.Params:
    prm = H.Params()
    H.load_library().hgp_default_params(C.byref(prm))
    prm.mode, prm.kway, prm.epsilon = H.HGP_MODE_KWAY, k, eps
    prm.max_nodes_per_part, prm.max_inbound_per_part, prm.max_parts = kway_omega(n, k, eps), 2**31 - 1, k
    prm.save_memory_up_to_level = 0  # 288 GB of HBM: every level stays resident
    return prm


def cpu_baseline(budget_s: float = 30.0) -> dict:
    """Sequential oracle on a 50K-node sample of the same workload (one host core)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    n = 50_000
    e = int(round(n * N_NETS / N_NODES))
    hg = H.Hypergraph.generate("netlist", n=n, e=e, alpha=ALPHA, max_size=MAX_SIZE, radius=RADIUS,
                               frac_local=FRAC_LOCAL, seed=1)
    g = O.HG(*hg.arrays())
    t0 = time.perf_counter()
    O.partition(g, 1, kway_omega(n, K, EPS), 2**31 - 1, K, kway=K, eps=EPS)
    dt = time.perf_counter() - t0
    return {"value": len(g.pins) / dt, "unit": "pins/s", "cores": 1, "kind": "port",
            "sample": f"netlist {n} nodes / {e} nets / {len(g.pins)} pins, k={K}, eps={EPS}, "
                      f"one run of the oracle pipeline ({dt:.1f} s)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(value is the one GPU's throughput)", file=sys.stderr)
        return 2
    lib = H.load_library()

    t0 = time.perf_counter()
    hg = H.Hypergraph.generate("netlist", n=N_NODES, e=N_NETS, alpha=ALPHA, max_size=MAX_SIZE, radius=RADIUS,
                               frac_local=FRAC_LOCAL, seed=args.seed)
    n, off, pins, srcc, w = hg.arrays()
    hg = H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)
    n, off, pins, srcc, w = hg.arrays()
    e, p = len(srcc), len(pins)
    print(f"# workload generated: n={n} e={e} pins={p} mean|e|={p / e:.3f} ({time.perf_counter() - t0:.1f} s)",
          file=sys.stderr, flush=True)

    dev = H.Device(0)
    d_pins, d_off, d_srcc, d_w = dev.upload(pins), dev.upload(off), dev.upload(srcc), dev.upload(w)
    prm = kway_params(n, K, EPS)
    part = np.zeros(n, np.uint32)
    part_p = part.ctypes.data_as(C.POINTER(C.c_uint32))
    st = H.RunStats()

    def step():
        dev.check(lib.hgp_partition_device(dev.ctx, C.byref(prm), n, e, C.c_void_p(d_pins), C.c_void_p(d_off),
                                           C.c_void_p(d_srcc), C.c_void_p(d_w), part_p, C.byref(st)),
                  "hgp_partition_device")

    for i in range(args.warmup):
        step()
        print(f"# warmup {i} done: {st.core_ms:.1f} ms, {st.levels} levels", file=sys.stderr, flush=True)

    prof = H.KernelProfile(dev, ROOFLINE_KERNEL)
    dev.check(lib.hgp_synchronize(dev.ctx), "sync")
    t0 = time.perf_counter()
    core_ms = []
    for i in range(args.steps):
        step()
        core_ms.append(st.core_ms)
    dev.check(lib.hgp_synchronize(dev.ctx), "sync")
    elapsed = time.perf_counter() - t0
    kp = prof.read()
    print(f"# timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    km1, cut, soed = hg.metrics(part)
    valid = hg.validity(part, prm.max_nodes_per_part, prm.max_inbound_per_part, prm.max_parts)
    ms_per_step = elapsed * 1e3 / args.steps
    value = p * args.steps / elapsed

    roofline = None
    if kp["launches"]:
        avg_s = kp["ms"] / kp["launches"] / 1e3
        bytes_per_launch = kp["bytes"] / kp["launches"]
        achieved = bytes_per_launch / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": ROOFLINE_KERNEL, "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_PER_LAUNCH,
                    "launches_per_step": kp["launches"] / args.steps, "avg_launch_us": avg_s * 1e6,
                    "algorithmic_bytes_per_launch": bytes_per_launch,
                    "kernel_share_of_step": kp["ms"] / (ms_per_step * args.steps)}

    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "pins/s end-to-end at equal λ−1 (vs plain & CPU); p50/p99 job latency batched",
        "value": value, "unit": "pins/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"netlist {n} nodes / {e} nets / {p} pins (power law alpha {ALPHA} on [2,{MAX_SIZE}], "
                               f"radius {RADIUS}, {int(FRAC_LOCAL * 100)}% local), -k {K} {EPS}",
                   "boundary": "device-resident CSR -> host partition vector", "smh": 0,
                   "candidates": prm.candidates_count, "refine_repeats": prm.refine_repeats,
                   "levels": st.levels, "lambda_minus_1": km1, "valid": bool(valid == 1),
                   "core_ms_p50": float(np.median(core_ms)), "ip_ms": st.ip_ms},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
