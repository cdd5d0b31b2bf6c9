"""The continuous-batching service: jobs of mixed modes and sizes submitted together run
concurrently on one GPU (one stream per worker), and each result equals the job run alone."""
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
