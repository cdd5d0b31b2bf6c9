"""The continuous-batching service: jobs of mixed modes and sizes submitted together run
concurrently on one GPU (one stream per worker), and each result equals the job run alone."""
import ctypes as C

import numpy as np
import pytest

import axoncuda_amd as H
import gpu_stages as G

pytestmark = pytest.mark.gpu


def job_mix():
    jobs = []
    for i, k in enumerate((2, 8, 32)):
        hg = H.Hypergraph.generate("random", n=30000 + 7000 * i, e=30000 + 5000 * i, max_size=64, seed=11 + i)
        n, _, _ = hg.sizes()
        omega = int(np.ceil(np.float32(1.03) * np.float32(n) / np.float32(k)))
        jobs.append((hg, G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=0.03)))
    for ipm in (0, 1):
        hg = H.Hypergraph.generate("snn", layers=5, per_layer=4096, grid=64, fanout=32, sigma=3.0, seed=21 + ipm)
        prm = G.params(H.HGP_MODE_INCC, 1024, 4096, 1 << 20)
        prm.initial_partitions_merge = ipm
        jobs.append((hg, prm))
    hg = H.Hypergraph.generate("netlist", n=120000, e=127500, seed=31)
    n, _, _ = hg.sizes()
    omega = int(np.ceil(np.float32(1.03) * np.float32(n) / np.float32(32)))
    jobs.append((hg, G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, 32, kway=32, eps=0.03)))
    return jobs


def test_results_alone_equal_results_in_the_mix(dev):
    jobs = job_mix()
    alone = [G.partition(dev, hg, prm)[0] for hg, prm in jobs]
    srv = H.Server(0, workers=4)
    try:
        ids = [srv.submit(hg, prm) for hg, prm in jobs]
        ids += [srv.submit(hg, prm) for hg, prm in jobs]  # every job twice, interleaved
        got = [srv.wait(j) for j in ids]
    finally:
        srv.close()
    for i, (part, st, lat, q) in enumerate(got):
        assert np.array_equal(part, alone[i % len(jobs)]), f"job {i} differs from its solo run"
        assert lat >= q >= 0.0
        hg, prm = jobs[i % len(jobs)]
        assert 0 < st.peak_device_bytes <= dev.lib.hgp_job_footprint(hg.h, C.byref(prm)), "admission under-reserves"


def test_full_size_range_alone_equals_mix(dev):
    """Jobs spanning the serving stream's whole size range, 10^4 to 2 x 10^7 pins (random
    hypergraphs of the 20K-node kind and layered SNNs, k-way and inbound-constrained), all
    submitted at once: each result equals the job run alone, and each job's measured device
    peak stays within what admission reserved for it."""
    jobs = []
    for i, pins in enumerate((1e4, 2e5, 3e6, 2e7)):
        e = int(pins / 4)
        hg = H.Hypergraph.generate("random", n=e, e=e, max_size=64, seed=40 + i)
        n, _, _ = hg.sizes()
        k = (2, 8, 32, 8)[i]
        omega = int(np.ceil(np.float32(1.03) * np.float32(n) / np.float32(k)))
        jobs.append((hg, G.params(H.HGP_MODE_KWAY, omega, 2**31 - 1, k, kway=k, eps=0.03)))
    for i, pins in enumerate((5e4, 4e6, 2e7)):
        per_layer = int(pins / (7 * 65))
        grid = int(np.ceil(np.sqrt(per_layer)))
        hg = H.Hypergraph.generate("snn", layers=8, per_layer=per_layer, grid=grid, fanout=64, sigma=4.0, seed=50 + i)
        jobs.append((hg, G.params(H.HGP_MODE_INCC, 1024, 4096, 1 << 20)))
    for hg, prm in jobs:
        prm.save_memory_up_to_level = 0
    print("case: serving full size range, %d jobs" % len(jobs), flush=True)
    alone = [G.partition(dev, hg, prm)[0] for hg, prm in jobs]
    srv = H.Server(0, workers=4)
    try:
        ids = [srv.submit(hg, prm) for hg, prm in jobs]
        got = [srv.wait(j) for j in ids]
    finally:
        srv.close()
    for i, (part, st, lat, q) in enumerate(got):
        assert np.array_equal(part, alone[i]), f"job {i} differs from its solo run"
        hg, prm = jobs[i]
        assert 0 < st.peak_device_bytes <= dev.lib.hgp_job_footprint(hg.h, C.byref(prm)), "admission under-reserves"
