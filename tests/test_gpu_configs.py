"""Whole-pipeline runs on the benchmark-shaped workloads, at sizes the oracle cannot follow,
checked through size-independent properties: every constraint holds (the reference's own
validity rule, sources/constr.cpp:90-147), the partition count respects the lower bound
ceil(N / Omega), λ−1 computed on the device in fixed point equals the host metric, and two
runs give the same partition."""
import ctypes as C

import numpy as np
import pytest

import axoncuda_amd as H
import gpu_stages as G

pytestmark = pytest.mark.gpu


def incc_params(omega, delta, max_parts, ipm):
    prm = G.params(H.HGP_MODE_INCC, omega, delta, max_parts)
    prm.initial_partitions_merge = 1 if ipm else 0
    prm.device_touching_construction = 1
    return prm


def device_km1(dev, hg, part):
    """lambda-1 of `part` on the device (2^18 fixed point) -> float."""
    n, off, pins, srcc, w = hg.arrays()
    d = [dev.upload(x) for x in (pins, off, w, part)]
    out = C.c_int64()
    dev.check(dev.lib.hgp_connectivity(dev.ctx, C.c_void_p(d[0]), C.c_void_p(d[1]), C.c_void_p(d[2]),
                                       C.c_void_p(d[3]), len(srcc), C.byref(out)), "hgp_connectivity")
    for x in d:
        dev.free(x)
    return out.value / 262144.0


@pytest.mark.parametrize("ipm", [False, True])
def test_snn_incc_scaled(dev, ipm):
    """The SNN workload shape (layered receptive fields, one source per hedge) at 8 x 16384
    neurons, fan-out 64, under -m 1024 4096 with and without -ipm."""
    print("case: snn 8x16384 fanout 64, ipm", ipm, flush=True)
    hg = H.Hypergraph.generate("snn", layers=8, per_layer=16384, grid=128, fanout=64, sigma=4.0, seed=3)
    n, e, p = hg.sizes()
    assert p == 7 * 16384 * 65
    prm = incc_params(1024, 4096, 1 << 20, ipm)
    part, st = G.partition(dev, hg, prm)
    assert hg.validity(part, 1024, 4096, 1 << 20) == 1
    k = int(part.max()) + 1
    assert st.num_partitions == k and k >= -(-n // 1024)
    assert np.array_equal(np.unique(part), np.arange(k))  # zero-based, no gaps
    km1 = hg.metrics(part)[0]
    assert device_km1(dev, hg, part) == pytest.approx(km1, rel=0, abs=1e-6)
    part2, _ = G.partition(dev, hg, prm)
    assert np.array_equal(part, part2), "two runs differ"


def test_netlist_kway_scaled(dev):
    """The netlist workload shape (power-law nets with grid locality) at 400K nodes, k = 64."""
    print("case: netlist 400K k=64", flush=True)
    hg = H.Hypergraph.generate("netlist", n=400_000, e=425_000, seed=5)
    n, e, p = hg.sizes()
    omega = int(np.ceil(np.float32(1.03) * np.float32(n) / np.float32(64)))
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, 64, kway=64, eps=0.03)
    part, st = G.partition(dev, hg, prm)
    assert hg.validity(part, omega, 2**31 - 1, 64) == 1
    assert int(part.max()) < 64
    km1 = hg.metrics(part)[0]
    assert device_km1(dev, hg, part) == pytest.approx(km1, rel=0, abs=1e-6)
    part2, _ = G.partition(dev, hg, prm)
    assert np.array_equal(part, part2), "two runs differ"


def test_snn_incc_full_size(dev):
    """The full SNN workload: 8 layers x 131,072 neurons, fan-out 256 from a Gaussian
    receptive field (sigma 16, 363 x 363 grid), 235.8M pins, under -m 1024 4096 with -ipm
    (about 40 s on one MI355X).  Every constraint holds, the partition count respects the
    lower bound, and the device lambda-1 equals the host metric."""
    print("case: snn 8x131072 fanout 256 (full size), ipm", flush=True)
    hg = H.Hypergraph.generate("snn", layers=8, per_layer=131072, grid=363, fanout=256, sigma=16.0, seed=1)
    n, e, p = hg.sizes()
    assert (n, e, p) == (1048576, 917504, 235798528)
    prm = incc_params(1024, 4096, 1 << 20, True)
    part, st = G.partition(dev, hg, prm)
    assert hg.validity(part, 1024, 4096, 1 << 20) == 1
    k = int(part.max()) + 1
    assert st.num_partitions == k and k >= -(-n // 1024)
    km1 = hg.metrics(part)[0]
    assert device_km1(dev, hg, part) == pytest.approx(km1, rel=0, abs=1e-6)


def test_netlist_kway_full_size(dev):
    """The full netlist benchmark workload: 3.2M nodes, 3.4M nets (power law alpha 2.5 on
    [2, 1000], radius 8, 80% local), k = 64, every level resident (-smh 0), as the benchmark
    runs it.  Every constraint holds, the device lambda-1 equals the host metric, and a second
    run gives the same partition over the same number of levels."""
    print("case: netlist 3.2M k=64 (full size)", flush=True)
    hg = H.Hypergraph.generate("netlist", n=3_200_000, e=3_400_000, alpha=2.5, max_size=1000, radius=8,
                               frac_local=0.8, seed=1)
    n, off, pins, srcc, w = hg.arrays()
    hg = H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)
    n, e, p = hg.sizes()
    assert n == 3_200_000 and e == 3_400_000 and p > 15_000_000
    omega = int(np.ceil(np.float32(1.03) * np.float32(n) / np.float32(64)))
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, 64, kway=64, eps=0.03)
    prm.save_memory_up_to_level = 0
    part, st = G.partition(dev, hg, prm)
    assert hg.validity(part, omega, 2**31 - 1, 64) == 1
    assert int(part.max()) < 64 and st.num_partitions == 64
    km1 = hg.metrics(part)[0]
    assert device_km1(dev, hg, part) == pytest.approx(km1, rel=0, abs=1e-6)
    part2, st2 = G.partition(dev, hg, prm)
    assert np.array_equal(part, part2), "two runs differ"
    assert st.levels == st2.levels and st.levels >= 8
    print(f"netlist 3.2M: levels {st.levels}, lambda-1 {km1:.0f}, core {st.core_ms:.0f} ms, ip {st.ip_ms:.0f} ms",
          flush=True)
