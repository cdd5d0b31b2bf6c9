"""GPU end-to-end: hgp_partition against the oracle pipeline (bit-exact partitions), plus
validity, determinism and the CLI binary."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import axoncuda_amd as H
import gen
import gpu_stages as G
import oracle as O

pytestmark = pytest.mark.gpu


def hg_pair(n, off, pins, srcc, w):
    """the same CSR as a product model (deduplicated) and as an oracle CSR"""
    prod = H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)
    orc = O.HG(*prod.arrays())
    return prod, orc


@pytest.mark.parametrize("k", [2, 4, 16])
def test_kway_pipeline_matches_oracle(dev, k, rng):
    n, off, pins, srcc, w = gen.random_hypergraph(20000, 20000, rng, max_size=10)
    prod, orc = hg_pair(n, off, pins, srcc, w)
    eps = 0.03
    omega = O.kway_omega(n, k, eps)
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=eps)
    part, st = G.partition(dev, prod, prm)
    exp, levels = O.partition(orc, 1, omega, 2**31 - 1, k, kway=k, eps=eps)
    assert np.array_equal(part, exp)
    assert O.validity(orc, part, omega, 2**31 - 1, k) == 1
    assert st.levels == levels


@pytest.mark.parametrize("ipm", [False, True])
def test_incc_pipeline_matches_oracle(dev, ipm, rng):
    n, off, pins, srcc, w = gen.snn_like(6, 800, 12, rng)
    prod, orc = hg_pair(n, off, pins, srcc, w)
    omega, delta, maxp = 64, 600, 4800
    prm = G.params(H.HGP_MODE_INCC, omega, delta, maxp)
    prm.initial_partitions_merge = 1 if ipm else 0
    part, st = G.partition(dev, prod, prm)
    exp, _ = O.partition(orc, 0, omega, delta, maxp, ipm=ipm)
    assert np.array_equal(part, exp)
    assert O.validity(orc, part, omega, delta, maxp) == 1


@pytest.mark.parametrize("budget", [1, 40000])
def test_incc_inbound_events_in_partition_ranges(dev, budget, rng):
    """An event budget far below what a round's inbound-set events need splits them into
    ranges of partitions (one partition per range at a budget of one byte) processed one
    after the other: the partition is still the oracle's."""
    n, off, pins, srcc, w = gen.snn_like(6, 800, 12, rng)
    prod, orc = hg_pair(n, off, pins, srcc, w)
    omega, delta, maxp = 64, 600, 4800
    prm = G.params(H.HGP_MODE_INCC, omega, delta, maxp)
    prm.refine_event_budget = budget
    print(f"case: INCC pipeline event budget {budget}", flush=True)
    part, _ = G.partition(dev, prod, prm)
    exp, _ = O.partition(orc, 0, omega, delta, maxp)
    assert np.array_equal(part, exp)


def test_pipeline_deterministic(dev, rng):
    n, off, pins, srcc, w = gen.random_hypergraph(30000, 32000, rng, max_size=12, weights="int")
    prod, orc = hg_pair(n, off, pins, srcc, w)
    omega = O.kway_omega(n, 8, 0.03)
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, 8, kway=8, eps=0.03)
    a, _ = G.partition(dev, prod, prm)
    b, _ = G.partition(dev, prod, prm)
    assert np.array_equal(a, b)
    assert O.validity(orc, a, omega, 2**31 - 1, 8) == 1


@pytest.mark.parametrize("k", [4, 32])
def test_device_initial_partitioner(dev, k, rng):
    """The device initial partitioner (best-fit-decreasing starts, label propagation, a
    further coarsened hierarchy), forced on a level where the automatic choice takes the host
    bisection: valid partitions, the same on a second run, and a lambda-1 in the range of the
    bisection's on the same input (measured: +2.3% at k=4, +16% at k=32; on the 3.2M-node
    netlist at k=64, where the automatic choice takes the device partitioner, +1.0%)."""
    hg = H.Hypergraph.generate("netlist", n=120_000, e=128_000, seed=7)
    n, e, p = hg.sizes()
    omega = int(np.ceil(np.float32(1.03) * np.float32(n) / np.float32(k)))
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=0.03)
    prm.initial_partitioner = H.HGP_IP_HOST
    host, _ = G.partition(dev, hg, prm)
    prm.initial_partitioner = H.HGP_IP_DEVICE
    a, _ = G.partition(dev, hg, prm)
    b, _ = G.partition(dev, hg, prm)
    assert np.array_equal(a, b), "two runs differ"
    assert hg.validity(a, omega, 2**31 - 1, k) == 1 and int(a.max()) < k
    ka, kh = hg.metrics(a)[0], hg.metrics(host)[0]
    assert ka <= 1.25 * kh, (ka, kh)


def test_cli_roundtrip(dev, tmp_path, rng):
    n, off, pins, srcc, w = gen.random_hypergraph(5000, 5000, rng, max_size=8)
    f = str(tmp_path / "c0.hgr")
    gen.write_hgr(f, n, off, pins)
    pfile = str(tmp_path / "c0.part")
    sfile = str(tmp_path / "c0.quotient.snn")
    exe = os.path.join(H.BIN_DIR, "hgraph_gpu")
    r = subprocess.run([exe, "-r", f, "-k", "2", "0.03", "-p", pfile, "-s", sfile], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Connectivity:" in r.stdout and "Total pins:" in r.stdout
    part = np.loadtxt(pfile, dtype=np.uint32)
    # -s wrote the quotient hypergraph of that partition (the writer is pinned by the golden tests)
    qfile = str(tmp_path / "again.snn")
    loaded = H.Hypergraph.load(f, dedup=True)
    assert H.load_library().hgp_save_partitioned(loaded.h, part.ctypes.data_as(H._u32p), qfile.encode()) == 0
    assert open(sfile, "rb").read() == open(qfile, "rb").read()
    hg = O.load(f, True)
    omega = O.kway_omega(n, 2, 0.03)
    assert O.validity(hg, part, omega, 2**31 - 1, 2) == 1
    grader = os.path.join(H.BIN_DIR, "hgrade")
    g = subprocess.run([grader, "-r", f, "-prt", pfile, "-k-prt", "2", "0.03"], capture_output=True, text=True, timeout=120)
    assert g.returncode == 0 and "Connectivity:" in g.stdout


def upload_csr(dev, hg):
    n, off, pins, srcc, w = hg.arrays()
    return n, len(srcc), [dev.upload(x) for x in (pins, off, srcc, w)]


def partition_device(dev, prm, n, e, bufs):
    part = np.zeros(n, np.uint32)
    st = H.RunStats()
    rc = dev.lib.hgp_partition_device(dev.ctx, C.byref(prm), n, e, *[C.c_void_p(b) for b in bufs],
                                      part.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(st))
    return rc, part, st


@pytest.mark.parametrize("mode", ["kway", "incc"])
def test_partition_device_equals_host_model(dev, mode, rng):
    """hgp_partition_device (the benchmark's entry point) gives what hgp_partition gives on the
    same hypergraph, and leaves its inputs unchanged."""
    if mode == "kway":
        hg = H.Hypergraph.generate("netlist", n=60_000, e=64_000, seed=5)
        n, _, _ = hg.sizes()
        prm = G.params(H.HGP_MODE_KWAY, O.kway_omega(n, 16, 0.03), 2**31 - 1, 16, kway=16, eps=0.03)
    else:
        hg = H.Hypergraph.generate("snn", layers=5, per_layer=4096, grid=64, fanout=32, sigma=3.0, seed=9)
        prm = G.params(H.HGP_MODE_INCC, 256, 2048, 1 << 20)
    n, e, bufs = upload_csr(dev, hg)
    rc, part, st = partition_device(dev, prm, n, e, bufs)
    assert rc == 0, dev.lib.hgp_ctx_last_error(dev.ctx)
    exp, st2 = G.partition(dev, hg, prm)
    assert np.array_equal(part, exp)
    assert st.levels == st2.levels and st.total_pins == st2.total_pins
    _, off, pins, _, _ = hg.arrays()
    assert np.array_equal(dev.download(bufs[0], len(pins), np.uint32), pins)
    assert 0 < st.peak_device_bytes <= dev.lib.hgp_job_footprint(hg.h, C.byref(prm))


def test_partition_device_rejects_malformed_input(dev):
    hg = H.Hypergraph.generate("random", n=2000, e=2000, max_size=8, seed=3)
    n, off, pins, srcc, w = hg.arrays()
    prm = G.params(H.HGP_MODE_KWAY, O.kway_omega(n, 2, 0.03), 2**31 - 1, 2, kway=2, eps=0.03)
    bad_off = off.copy()
    bad_off += 4  # offsets that do not start at 0
    bufs = [dev.upload(pins), dev.upload(bad_off), dev.upload(srcc), dev.upload(w)]
    rc, _, _ = partition_device(dev, prm, n, len(srcc), bufs)
    assert rc == 1
    bad_pins = pins.copy()
    bad_pins[len(pins) // 2] = n  # a pin past the last node
    bufs = [dev.upload(bad_pins), dev.upload(off), dev.upload(srcc), dev.upload(w)]
    rc, _, _ = partition_device(dev, prm, n, len(srcc), bufs)
    assert rc == 1
    assert b"num_nodes" in dev.lib.hgp_ctx_last_error(dev.ctx)


def test_kway_max_parts_other_than_k_is_rejected(dev):
    hg = H.Hypergraph.generate("random", n=3000, e=3000, max_size=8, seed=4)
    n, _, _ = hg.sizes()
    prm = G.params(H.HGP_MODE_KWAY, O.kway_omega(n, 4, 0.03), 2**31 - 1, 4096, kway=4, eps=0.03)
    part = np.zeros(n, np.uint32)
    rc = dev.lib.hgp_partition(dev.ctx, C.byref(prm), hg.h, part.ctypes.data_as(C.POINTER(C.c_uint32)), None)
    assert rc == 1 and b"max_parts" in dev.lib.hgp_ctx_last_error(dev.ctx)


def test_netlist_k64_matches_oracle(dev):
    """The benchmark's generator and k = 64 at 50K nodes: the GPU partition equals the
    oracle's bit for bit."""
    hg = H.Hypergraph.generate("netlist", n=50_000, e=53_125, alpha=2.5, max_size=1000, radius=8, frac_local=0.8,
                               seed=1)
    n, off, pins, srcc, w = hg.arrays()
    prod, orc = hg_pair(n, off, pins, srcc, w)
    omega = O.kway_omega(n, 64, 0.03)
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, 64, kway=64, eps=0.03)
    part, st = G.partition(dev, prod, prm)
    exp, levels = O.partition(orc, 1, omega, 2**31 - 1, 64, kway=64, eps=0.03)
    assert np.array_equal(part, exp)
    assert st.levels == levels
    assert O.validity(orc, part, omega, 2**31 - 1, 64) == 1
