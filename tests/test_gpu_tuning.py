"""The autotuner: it runs on the device and picks values in range, and the launch parameters
it chooses change speed only: partitions and candidate lists are identical under the
defaults, the tuned values and deliberately extreme ones."""
import ctypes as C

import numpy as np
import pytest

import axoncuda_amd as H
import gen
import gpu_stages as G
import oracle as O

pytestmark = pytest.mark.gpu


def tuning(dev):
    t = H.Tuning()
    dev.check(dev.lib.hgp_tuning_get(dev.ctx, C.byref(t)), "hgp_tuning_get")
    return t


def extremes():
    """(name, setter) pairs of valid but far-from-default settings."""
    def small(t):
        for i in range(40):
            t.sort_rounds[i] = 1
        t.cascade_block = 64
        t.heavy_table_slots = 1 << 13
        t.heavy_table_fill16 = 4
        t.scan_lookback_tiles = 0
        t.mask_wide4_degree = 0  # every level's mask gains 16 waves per node
        t.mask_wide16_degree = 0
        t.side_streams = 0
        t.coop_max_n = 1 << 22  # every sort above 16 K keys in one launch of 64 workgroups
        t.coop_tile = 256

    def large(t):
        for i in range(40):
            t.sort_rounds[i] = 16
        t.cascade_block = 1024
        t.heavy_table_slots = 1 << 20
        t.heavy_table_fill16 = 14
        t.scan_lookback_tiles = 1 << 20
        t.mask_wide4_degree = 1 << 31  # a wave per node on every level
        t.mask_wide16_degree = 1 << 31
        t.side_streams = 3
        t.coop_max_n = 0  # no sort in one launch

    def four_waves(t):
        t.mask_wide4_degree = 0  # 4 waves per node on every level
        t.mask_wide16_degree = 1 << 31

    return [("small", small), ("large", large), ("four_waves", four_waves)]


def with_tuning(dev, setter, fn):
    saved = tuning(dev)
    t = H.Tuning()
    dev.lib.hgp_tuning_defaults(C.byref(t))
    setter(t)
    dev.check(dev.lib.hgp_tuning_set(dev.ctx, C.byref(t)), "hgp_tuning_set")
    try:
        return fn()
    finally:
        dev.check(dev.lib.hgp_tuning_set(dev.ctx, C.byref(saved)), "restore tuning")


def test_autotune_picks_values_in_range(dev):
    saved = tuning(dev)
    out = H.Tuning()
    print("case: autotune max_log2=20", flush=True)
    dev.check(dev.lib.hgp_autotune(dev.ctx, 20, C.byref(out)), "hgp_autotune")
    try:
        got = tuning(dev)
        assert got.as_dict() == out.as_dict()
        assert out.tuned == 1
        assert all(out.sort_rounds[c] in (1, 2, 4, 8, 16) for c in range(15, 21))
        assert out.cascade_block in (64, 128, 256, 512, 1024)
        assert out.heavy_table_slots in (1 << 15, 1 << 16, 1 << 17, 1 << 18)
        assert out.heavy_table_fill16 in (10, 12, 14)
        assert out.scan_lookback_tiles >= 16
    finally:
        dev.check(dev.lib.hgp_tuning_set(dev.ctx, C.byref(saved)), "restore tuning")


def test_invalid_tuning_is_refused(dev):
    t = tuning(dev)
    before = t.as_dict()
    t.cascade_block = 96
    assert dev.lib.hgp_tuning_set(dev.ctx, C.byref(t)) == 1
    assert tuning(dev).as_dict() == before


@pytest.mark.parametrize("name,setter", extremes(), ids=[n for n, _ in extremes()])
def test_partitions_do_not_depend_on_tuning(dev, name, setter, rng):
    cases = []
    n, off, pins, srcc, w = gen.random_hypergraph(40000, 40000, rng, max_size=24)
    cases.append((H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True), H.HGP_MODE_KWAY, 8))
    hg = H.Hypergraph.generate("snn", layers=5, per_layer=4096, grid=64, fanout=48, sigma=3.0, seed=3)
    cases.append((hg, H.HGP_MODE_INCC, 0))
    for hg, mode, k in cases:
        nn, _, _ = hg.sizes()
        if mode == H.HGP_MODE_KWAY:
            prm = G.params(mode, O.kway_omega(nn, k, 0.03), 2**31 - 1, k, kway=k, eps=0.03)
        else:
            prm = G.params(mode, 256, 2048, 1 << 20)
        base, _ = G.partition(dev, hg, prm)
        print(f"case: tuning {name} mode={mode} n={nn}", flush=True)
        got, _ = with_tuning(dev, setter, lambda: G.partition(dev, hg, prm))
        assert np.array_equal(base, got)


def hub_level(rng, n=1 << 18, hubs=8, per_hub=256, size=512):
    """Hubs whose distinct neighbours (about 100 K) exceed every LDS tier of candidate
    scoring, so they are scored in the HBM hash tables that the tuning sizes."""
    e = hubs * per_hub
    pins = np.empty(e * size, np.uint32)
    others = np.arange(hubs, n, dtype=np.uint32)
    for x in range(e):
        pins[x * size] = x % hubs
        pins[x * size + 1:(x + 1) * size] = rng.choice(others, size - 1, replace=False)
    off = np.arange(e + 1, dtype=np.uint64) * size
    return O.HG(n, off, pins, np.zeros(e, np.uint32), np.ones(e, np.float32))


@pytest.mark.parametrize("name,setter", extremes(), ids=[n for n, _ in extremes()])
def test_heavy_candidates_do_not_depend_on_tuning(dev, name, setter, rng):
    hg = hub_level(rng)
    L = G.Level(dev, hg)
    L.build_touching()
    sizes = np.ones(hg.n, np.uint32)
    a, b, *_ = G.candidates(dev, L, sizes, 4, 2**31 - 1, 2**31 - 1)
    print(f"case: heavy candidates tuning {name}", flush=True)
    c, d, *_ = with_tuning(dev, setter, lambda: G.candidates(dev, L, sizes, 4, 2**31 - 1, 2**31 - 1))
    assert np.array_equal(a, c) and np.array_equal(b, d)
