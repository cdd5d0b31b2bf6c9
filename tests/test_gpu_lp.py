"""GPU parity of the label-propagation pieces and the initial partitioners: the LP refinement
stage (hgp_refine_lp) against the oracle's Jacobi restatement, and whole k-way pipelines
with the device initial partitioner (default and other options), with the host bisection,
with the reference's pruning, and with LP refinement enabled, each bit-exact against the
oracle."""
import numpy as np
import pytest

import axoncuda_amd as H
import gen
import gpu_stages as G
import oracle as O

pytestmark = pytest.mark.gpu


def hg_pair(n, off, pins, srcc, w):
    prod = H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)
    return prod, O.HG(*prod.arrays())


@pytest.mark.parametrize("k", [3, 8, 70])
@pytest.mark.parametrize("rounds", [1, 6])
def test_lp_stage_matches_oracle(dev, k, rounds, rng):
    """k = 70 takes the partition-list gain kernels, the others the partition masks."""
    n, off, pins, srcc, w = gen.random_hypergraph(3000, 4000, rng, max_size=12, weights="float")
    _, hg = hg_pair(n, off, pins, srcc, w)
    L = G.Level(dev, hg)
    L.build_touching()
    sizes = rng.integers(1, 4, hg.n).astype(np.uint32)
    part = rng.integers(0, k, hg.n).astype(np.uint32)
    psizes = np.bincount(part, weights=sizes, minlength=k).astype(np.uint32)
    omega = int(psizes.max()) + 6
    print(f"case: lp stage k={k} rounds={rounds}", flush=True)
    gp, gps = G.refine_lp(dev, L, sizes, k, omega, rounds, part, psizes)
    ep, eps_ = O.lp_refine(hg, sizes, k, omega, rounds, part, psizes)
    assert np.array_equal(gp, ep)
    assert np.array_equal(gps, eps_)


def kway_case(rng, n=20000, max_size=10):
    n, off, pins, srcc, w = gen.random_hypergraph(n, n, rng, max_size=max_size)
    return hg_pair(n, off, pins, srcc, w)


@pytest.mark.parametrize("k", [4, 16])
def test_pipeline_with_lp_rounds_matches_oracle(dev, k, rng):
    prod, orc = kway_case(rng)
    omega = O.kway_omega(orc.n, k, 0.03)
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=0.03)
    prm.lp_rounds = 4
    part, _ = G.partition(dev, prod, prm)
    exp, _ = O.partition(orc, 1, omega, 2**31 - 1, k, kway=k, eps=0.03, lp_rounds=4)
    assert np.array_equal(part, exp)
    assert O.validity(orc, part, omega, 2**31 - 1, k) == 1


@pytest.mark.parametrize("k", [2, 8])
def test_pipeline_with_host_bisection_matches_oracle(dev, k, rng):
    prod, orc = kway_case(rng)
    omega = O.kway_omega(orc.n, k, 0.03)
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=0.03)
    prm.initial_partitioner = H.HGP_IP_HOST
    part, _ = G.partition(dev, prod, prm)
    exp, _ = O.partition(orc, 1, omega, 2**31 - 1, k, kway=k, eps=0.03, ip="host")
    assert np.array_equal(part, exp)


@pytest.mark.parametrize("k,opts", [(2, (0, 0, 0, 0)), (16, (0, 0, 0, 0)), (8, (4, 16, 8, 8)), (8, (1, 2, 64, 1))])
def test_pipeline_device_ip_matches_oracle(dev, k, opts, rng):
    """The device initial partitioner (forced: on these sizes the automatic choice takes
    the bisection), with its default and other options."""
    prod, orc = kway_case(rng, n=30000)
    omega = O.kway_omega(orc.n, k, 0.03)
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=0.03)
    prm.initial_partitioner = H.HGP_IP_DEVICE
    prm.ip_starts, prm.ip_rounds, prm.ip_nodes_per_part, prm.ip_refine_repeats = opts
    print(f"case: device ip k={k} opts={opts}", flush=True)
    part, _ = G.partition(dev, prod, prm)
    exp, _ = O.partition(orc, 1, omega, 2**31 - 1, k, kway=k, eps=0.03, ip="device", ip_options=opts)
    assert np.array_equal(part, exp)
    assert O.validity(orc, part, omega, 2**31 - 1, k) == 1


@pytest.mark.parametrize("ip", ["host", "device"])
def test_pipeline_with_pruning_matches_oracle(dev, ip, rng):
    """The reference's hedge pruning (keep target 1x) before either built-in partitioner."""
    prod, orc = kway_case(rng)
    k = 4
    omega = O.kway_omega(orc.n, k, 0.03)
    prm = G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=0.03)
    prm.initial_prune = 1
    prm.initial_partitioner = H.HGP_IP_DEVICE if ip == "device" else H.HGP_IP_HOST
    part, _ = G.partition(dev, prod, prm)
    exp, _ = O.partition(orc, 1, omega, 2**31 - 1, k, kway=k, eps=0.03, prune=1, ip=ip)
    assert np.array_equal(part, exp)
