"""Benchmark of the hypergraph partitioner's hot path on one MI355X.

Workload: netlist-shaped hypergraph, 3.2M nodes and 3.4M nets, net sizes from a discrete
power law (alpha 2.5 on [2, 1000]), 80% of a net's pins within radius 8 of its first pin on a
2D grid and 20% uniform, partitioned k-way with k=64 and epsilon 0.03.  One step is one full
partitioning (coarsening, initial partition, uncoarsening + refinement) of that hypergraph
starting from device-resident CSR arrays; the partition vector lands in host memory.

value = post-dedup pins / second of one step (the reference's "device core" boundary,
sources/main.cu:260-264,891-896), with inputs already in HBM.  Before the warm-up the launch
parameters are autotuned once (hgp_autotune, untimed; results do not depend on them) and
reported under config.tuning.  Beside it, the line carries:

* pipeline: the compulsory HBM traffic of a step (every stage's inputs and outputs once per
  executed level and refinement round) over the step time, against the 8 TB/s peak;
* roofline: the kernel with the largest share of device time (the small-hedge cascade).  Its
  launches in the timed steps are bracketed by HIP events on the library's stream; its
  algorithmic bytes per launch are counted on the device in one extra, untimed step (the
  pipeline is deterministic, so every step launches the same work); traffic comes from
  rocprofv3 FETCH_SIZE / WRITE_SIZE counters scaled by a plain copy of known size;
* end_to_end: the hgraph_gpu command line on the same hypergraph written as .hgr, timed from
  process start to the partition file written (parse, dedup, upload, partition, write);
* cpu_baseline: the sequential oracle on one host thread, on a 100K-node sample of the same
  generator, with its lambda-1 and the GPU's lambda-1 on that same sample;
* serving: the continuous-batching service over jobs of 10^4 to 2 x 10^7 pins.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-e2e] [--no-serve] [--no-tune]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import axoncuda_amd as H  # noqa: E402

N_NODES, N_NETS, ALPHA, MAX_SIZE, RADIUS, FRAC_LOCAL = 3_200_000, 3_400_000, 2.5, 1000, 8, 0.8
K, EPS = 64, 0.03
HBM_PEAK_GBS = 8000.0
# Kernel whose launches the roofline object covers: the largest share of device time on this
# workload in the kernel-trace summary under profiles/.
ROOFLINE_KERNEL = "k_cascade_small"
# HBM bytes per launch of that kernel, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs
# scaled by the same counters on a plain device copy of known size (profiles/traffic_<kernel>.json).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", f"traffic_{ROOFLINE_KERNEL}.json")
COPY_CAL_BYTES = 1 << 30  # the plain copy that calibrates the counters
CPU_SAMPLE_NODES = 100_000


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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


def netlist(n: int, seed: int) -> H.Hypergraph:
    e = int(round(n * N_NETS / N_NODES))
    hg = H.Hypergraph.generate("netlist", n=n, e=e, alpha=ALPHA, max_size=MAX_SIZE, radius=RADIUS,
                               frac_local=FRAC_LOCAL, seed=seed)
    n, off, pins, srcc, w = hg.arrays()
    return H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)


def end_to_end(hg: H.Hypergraph, tmpdir: str) -> dict:
    """The drop-in command line, timed from process start to the partition file written."""
    n, e, p = hg.sizes()
    path = os.path.join(tmpdir, "netlist.hgr")
    pfile = os.path.join(tmpdir, "netlist.part")
    hg.save(path)
    exe = os.path.join(H.BIN_DIR, "hgraph_gpu")
    cmd = [exe, "-r", path, "-k", str(K), str(EPS), "-p", pfile]
    t0 = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    wall = time.perf_counter() - t0
    if r.returncode != 0 or not os.path.exists(pfile):
        raise RuntimeError(f"hgraph_gpu failed ({r.returncode}): {r.stderr[-2000:]}")
    km1 = None
    for line in r.stdout.splitlines():
        if line.strip().startswith("Connectivity:"):
            km1 = float(line.split()[1])
    log(f"# end to end: {wall:.2f} s, lambda-1 {km1}")
    return {"value": p / wall, "unit": "pins/s", "wall_s": wall, "lambda_minus_1": km1,
            "command": "hgraph_gpu -r netlist.hgr -k 64 0.03 -p netlist.part (reference defaults otherwise: -smh 2)",
            "boundary": "process start -> partition file written (.hgr parse, dedup, upload, partition, write)"}


def cpu_baseline(lib, dev) -> dict:
    """Sequential oracle (one host thread) on a 100K-node sample of the same generator, and
    the GPU pipeline on the same sample, both scored by lambda-1."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    hg = netlist(CPU_SAMPLE_NODES, 1)
    g = O.HG(*hg.arrays())
    O.set_ip_threads(1)
    t0 = time.perf_counter()
    part, _ = O.partition(g, 1, kway_omega(g.n, K, EPS), 2**31 - 1, K, kway=K, eps=EPS)
    dt = time.perf_counter() - t0
    O.set_ip_threads(0)
    km1_cpu = O.metrics(g, part)[0]
    prm = kway_params(g.n, K, EPS)
    gpart = np.zeros(g.n, np.uint32)
    st = H.RunStats()
    dev.check(lib.hgp_partition(dev.ctx, C.byref(prm), hg.h, gpart.ctypes.data_as(C.POINTER(C.c_uint32)),
                                C.byref(st)), "hgp_partition")
    km1_gpu = hg.metrics(gpart)[0]
    log(f"# cpu baseline: {dt:.1f} s on one thread, lambda-1 cpu {km1_cpu:.0f} gpu {km1_gpu:.0f}")
    return {"value": len(g.pins) / dt, "unit": "pins/s", "cores": 1, "kind": "port",
            "sample": f"netlist {g.n} nodes / {g.e} nets / {len(g.pins)} pins, k={K}, eps={EPS}; one run of the "
                      f"oracle pipeline on one host thread ({dt:.1f} s)",
            "lambda_minus_1": km1_cpu, "gpu_lambda_minus_1_same_sample": km1_gpu,
            "gpu_core_ms_same_sample": st.core_ms, "same_partition": bool(np.array_equal(part, gpart))}


def serve_jobs(n_jobs: int, max_pins: float, seed: int):
    """The serving stream's jobs: pins log-uniform in [1e4, max_pins]; half random
    hypergraphs of the 20K-node kind (sizes 2 + Geom(1/2), unit weights), half layered SNNs
    of the inbound-constrained kind; modes uniform over k = 2, 8, 32 at eps 0.03 and
    -m 1024 4096 1048576."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n_jobs):
        if i % 32 == 0:
            log(f"# generating job {i}")
        pins = float(np.exp(rng.uniform(np.log(1e4), np.log(max_pins))))
        if i % 2 == 0:
            e = max(64, int(pins / 4))
            hg = H.Hypergraph.generate("random", n=e, e=e, max_size=64, seed=seed * 1000 + i)
        else:
            fan = 64
            per_layer = max(fan + 2, int(pins / (7 * (fan + 1))))
            grid = int(np.ceil(np.sqrt(per_layer)))
            hg = H.Hypergraph.generate("snn", layers=8, per_layer=per_layer, grid=grid, fanout=fan, sigma=4.0,
                                       seed=seed * 1000 + i)
        n, _, p = hg.sizes()
        mode = int(rng.integers(0, 4))
        prm = H.Params()
        H.load_library().hgp_default_params(C.byref(prm))
        prm.save_memory_up_to_level = 0
        if mode < 3:
            k = (2, 8, 32)[mode]
            prm.mode, prm.kway, prm.epsilon = H.HGP_MODE_KWAY, k, EPS
            prm.max_nodes_per_part, prm.max_inbound_per_part, prm.max_parts = kway_omega(n, k, EPS), 2**31 - 1, k
        else:
            prm.mode = H.HGP_MODE_INCC
            prm.max_nodes_per_part, prm.max_inbound_per_part, prm.max_parts = 1024, 4096, 1 << 20
        jobs.append((hg, prm, p))
    return jobs


def serve_bench(jobs_n: int, max_pins: float, workers: int, load: float, seed: int) -> dict:
    """Plain sequential runs of the job stream, then the same jobs through the continuous-
    batching service with Poisson arrivals at `load` times the plain service rate (the
    speedup over plain runs cannot exceed about `load`)."""
    jobs = serve_jobs(jobs_n, max_pins, seed)
    total_pins = sum(p for _, _, p in jobs)
    log(f"# serving stream: {len(jobs)} jobs, {total_pins} pins")
    lib = H.load_library()
    dev = H.Device(0)
    plain_parts, plain_lat = [], []
    t0 = time.perf_counter()
    for i, (hg, prm, _) in enumerate(jobs):
        if i % 32 == 0:
            log(f"# plain: job {i}")
        n, _, _ = hg.sizes()
        part = np.zeros(n, np.uint32)
        st = H.RunStats()
        t1 = time.perf_counter()
        dev.check(lib.hgp_partition(dev.ctx, C.byref(prm), hg.h, part.ctypes.data_as(C.POINTER(C.c_uint32)),
                                    C.byref(st)), "hgp_partition")
        plain_lat.append((time.perf_counter() - t1) * 1e3)
        plain_parts.append(part)
    t_plain = time.perf_counter() - t0
    dev.close()
    log(f"# plain sequential: {t_plain:.2f} s")
    rate = load * len(jobs) / t_plain  # offered load: a multiple of what plain runs sustain
    arrivals = np.cumsum(np.random.default_rng(seed + 7).exponential(1.0 / rate, len(jobs)))
    srv = H.Server(0, workers=workers)
    ids = []
    t0 = time.perf_counter()
    for (hg, prm, _), ta in zip(jobs, arrivals):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        ids.append(srv.submit(hg, prm))
    lat, same = [], 0
    for i, j in enumerate(ids):
        if i % 32 == 0:
            log(f"# served: waiting for job {i}")
        part, st, l, q = srv.wait(j)
        lat.append(l)
        same += int(np.array_equal(part, plain_parts[i]))
    t_serve = time.perf_counter() - t0
    srv.close()
    lat = np.array(lat)
    return {"metric": "aggregate pins/s of the serving stream; p50/p99 job latency", "value": total_pins / t_serve,
            "unit": "pins/s", "plain_value": total_pins / t_plain, "speedup_vs_plain": t_plain / t_serve,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "plain_p50_ms": float(np.percentile(plain_lat, 50)), "plain_p99_ms": float(np.percentile(plain_lat, 99)),
            "jobs": len(jobs), "workers": workers, "total_pins": total_pins,
            "arrival_rate_per_s": rate, "results_equal_to_plain": same, "data": "synthetic",
            "config": {"workload": f"{len(jobs)} jobs, pins log-uniform in [1e4, {max_pins:g}], half random "
                                   "(|e| = 2 + Geom(1/2)), half layered SNN (fan-out 64); modes k=2/8/32 eps 0.03 "
                                   f"or -m 1024 4096 1048576; Poisson arrivals at {load:g}x the plain service rate"}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end command-line leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in launch parameters (no autotuning)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--serve", action="store_true", help="run only the serving-stream benchmark")
    ap.add_argument("--jobs", type=int, default=32, help="serving jobs (the full stream is 256)")
    ap.add_argument("--max-pins", type=float, default=2e7)
    ap.add_argument("--workers", type=int, default=6)
    ap.add_argument("--load", type=float, default=4.0, help="offered load of the serving leg, in multiples of the "
                                                             "plain rate")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain device copy of 1 GiB (for rocprofv3 --pmc counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(value is the one GPU's throughput)", file=sys.stderr)
        return 2
    if args.serve:
        print(json.dumps(serve_bench(args.jobs, args.max_pins, args.workers, args.load, args.seed)))
        return 0
    lib = H.load_library()

    t0 = time.perf_counter()
    hg = netlist(N_NODES, args.seed)
    n, off, pins, srcc, w = hg.arrays()
    e, p = len(srcc), len(pins)
    log(f"# workload generated: n={n} e={e} pins={p} mean|e|={p / e:.3f} ({time.perf_counter() - t0:.1f} s)")

    tmpdir = tempfile.mkdtemp(prefix="hgp_bench_")
    e2e = None if args.no_e2e else end_to_end(hg, tmpdir)  # its own process, before this one uses the GPU

    dev = H.Device(0)
    tuning = None
    if not args.no_tune:  # launch parameters only (untimed; results do not depend on them)
        t = H.Tuning()
        t0 = time.perf_counter()
        dev.check(lib.hgp_autotune(dev.ctx, 26, C.byref(t)), "hgp_autotune")
        dev.check(lib.hgp_tuning_set_process_default(C.byref(t)), "tuning default")  # serving workers too
        tuning = dict(t.as_dict(), seconds=time.perf_counter() - t0)
        log(f"# autotuned in {tuning['seconds']:.1f} s: {t.as_dict()}")
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
        log(f"# warmup {i} done: {st.core_ms:.1f} ms, {st.levels} levels, ip {st.ip_ms:.1f} ms")

    prof = H.KernelProfile(dev, ROOFLINE_KERNEL)
    dev.check(lib.hgp_synchronize(dev.ctx), "sync")
    t0 = time.perf_counter()
    core_ms, ip_ms = [], []
    for i in range(args.steps):
        step()
        core_ms.append(st.core_ms)
        ip_ms.append(st.ip_ms)
    dev.check(lib.hgp_synchronize(dev.ctx), "sync")
    elapsed = time.perf_counter() - t0
    kp = prof.read()
    log(f"# timed {args.steps} steps in {elapsed:.3f} s")
    # one more step, untimed, in which the roofline kernel counts its algorithmic bytes
    cnt = H.KernelProfile(dev, ROOFLINE_KERNEL, count_bytes=True)
    step()
    kc = cnt.read()

    km1, cut, soed = hg.metrics(part)
    valid = hg.validity(part, prm.max_nodes_per_part, prm.max_inbound_per_part, prm.max_parts)
    ms_per_step = elapsed * 1e3 / args.steps
    value = p * args.steps / elapsed

    pipeline = {"bytes_algorithmic_per_step": int(st.bytes_algorithmic),
                "achieved": st.bytes_algorithmic / (ms_per_step / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": st.bytes_algorithmic / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
                "model": "every stage's inputs and outputs once per executed level and refinement round"}
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = None
    if kp["launches"] and kc["launches"]:
        avg_s = kp["ms"] / kp["launches"] / 1e3
        bytes_per_launch = kc["bytes"] / kc["launches"]
        achieved = bytes_per_launch / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": ROOFLINE_KERNEL, "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                    "launches_per_step": kp["launches"] / args.steps, "avg_launch_us": avg_s * 1e6,
                    "algorithmic_bytes_per_launch": bytes_per_launch,
                    "kernel_share_of_step": kp["ms"] / (ms_per_step * args.steps)}

    if args.copy_calibration:
        a, b = dev.alloc(COPY_CAL_BYTES), dev.alloc(COPY_CAL_BYTES)
        dev.check(lib.hgp_memcpy_dtod(dev.ctx, C.c_void_p(b), C.c_void_p(a), COPY_CAL_BYTES), "hgp_memcpy_dtod")
        dev.free(a)
        dev.free(b)
        log(f"# copy calibration: {COPY_CAL_BYTES} bytes copied device to device")

    cpu = None if args.no_cpu else cpu_baseline(lib, dev)
    dev.close()
    serving = None if args.no_serve else serve_bench(args.jobs, args.max_pins, args.workers, args.load, args.seed)
    out = {
        "metric": "pins/s end-to-end at equal λ−1 (vs plain & CPU); p50/p99 job latency batched",
        "value": value, "unit": "pins/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"netlist {n} nodes / {e} nets / {p} pins (power law alpha {ALPHA} on [2,{MAX_SIZE}], "
                               f"radius {RADIUS}, {int(FRAC_LOCAL * 100)}% local), -k {K} {EPS}",
                   "variant": "core: device-resident CSR -> host partition vector (end_to_end holds the command "
                              "line from process start)",
                   "smh": 0, "candidates": prm.candidates_count, "refine_repeats": prm.refine_repeats,
                   "levels": st.levels, "lambda_minus_1": km1, "valid": bool(valid == 1),
                   "core_ms_p50": float(np.median(core_ms)), "ip_ms_p50": float(np.median(ip_ms)),
                   "initial_prune": prm.initial_prune, "initial_partitioner": prm.initial_partitioner,
                   "tuning": tuning},
        "pipeline": pipeline,
        "roofline": roofline,
        "end_to_end": e2e,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
