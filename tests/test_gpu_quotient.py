"""hgp_quotient (getPartitionsHypergraph(part, 2, true) on the device) against the oracle's
restatement, which the reference's own saved quotient files pin (tests/test_golden.py):
the same merged hedges in the same order, pins, source counts and float weights."""
import ctypes as C
import os

import numpy as np
import pytest

import axoncuda_amd as H
import gen
import gpu_stages as G
import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def device_quotient(dev, hg, part):
    L = G.Level(dev, hg)
    d_part = dev.upload(np.asarray(part, np.uint32))
    k, e, p = C.c_uint32(), C.c_uint32(), C.c_uint64()
    dp, do, ds, dw = G.V(), G.V(), G.V(), G.V()
    st = dev.lib.hgp_quotient(dev.ctx, G.V(L.pins), G.V(L.off), G.V(L.srcc), G.V(L.w), G.V(d_part), hg.n, hg.e,
                              C.byref(k), C.byref(e), C.byref(p), C.byref(dp), C.byref(do), C.byref(ds), C.byref(dw))
    out = None
    if st == 0:
        out = (k.value, dev.download(do.value, e.value + 1, np.uint64), dev.download(dp.value, p.value, np.uint32),
               dev.download(ds.value, e.value, np.uint32), dev.download(dw.value, e.value, np.float32))
        for x in (dp, do, ds, dw):
            dev.free(x.value)
    for x in (L.pins, L.off, L.srcc, L.w, d_part):
        dev.free(x)
    return st, out


def check(dev, hg, part):
    st, got = device_quotient(dev, hg, part)
    assert st == 0
    k, q = O.quotient(hg, part)
    gk, goff, gpins, gsrcc, gw = got
    assert gk == k
    assert np.array_equal(goff, q.off)
    assert np.array_equal(gpins, q.pins)
    assert np.array_equal(gsrcc, q.srcc)
    assert np.array_equal(gw.view(np.uint32), q.w.view(np.uint32))  # float sums in hedge order


@pytest.mark.parametrize("src_prob,k", [(0.0, 2), (0.0, 64), (1.0, 16), (0.6, 300)])
def test_quotient_matches_oracle(dev, src_prob, k, rng):
    hg = O.HG(*gen.random_hypergraph(30000, 40000, rng, max_size=12, src_prob=src_prob, max_src=3, dup_prob=0.2,
                                     weights="float"))
    part = rng.integers(0, k, hg.n).astype(np.uint32)
    part[:k] = np.arange(k)  # every id used
    print(f"case: quotient src_prob={src_prob} k={k}", flush=True)
    check(dev, hg, part)


def test_quotient_of_a_partition_from_the_pipeline(dev, rng):
    n, off, pins, srcc, w = gen.snn_like(6, 800, 12, rng)
    prod = H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)
    orc = O.HG(*prod.arrays())
    prm = G.params(H.HGP_MODE_INCC, 64, 600, 4800)
    part, _ = G.partition(dev, prod, prm)
    check(dev, orc, part)


@pytest.mark.parametrize("name", ["hgr_fmt1", "hgr_fmt10", "snn", "axh"])
def test_quotient_of_golden_inputs(dev, name, rng):
    import json
    data = json.load(open(os.path.join(GOLD, "golden.json")))
    hg = O.HG(*H.Hypergraph.load(os.path.join(GOLD, data[name]["file"]), dedup=True).arrays())
    part = np.array(data["saved"][name]["part"], np.uint32)
    check(dev, hg, part)


def test_quotient_refuses_partition_ids_with_gaps(dev, rng):
    hg = O.HG(*gen.random_hypergraph(500, 600, rng, max_size=6))
    part = np.zeros(hg.n, np.uint32)
    part[::3] = 2  # ids 0 and 2 only
    st, _ = device_quotient(dev, hg, part)
    assert st == 1  # HGP_E_INVALID


def test_quotient_with_every_hedge_inside_one_part(dev, rng):
    """All pins destinations and one partition: every hedge maps to the same {0}; all merge
    into one hedge whose weight is the in-order float sum."""
    hg = O.HG(*gen.random_hypergraph(2000, 5000, rng, max_size=8, weights="float"))
    check(dev, hg, np.zeros(hg.n, np.uint32))
