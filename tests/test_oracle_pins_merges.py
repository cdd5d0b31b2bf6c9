"""Pins the oracle's grouping, packing and initial-partition input steps by brute force (CPU only).

Each expectation is written from the reference's definition, with Python sets and plain
sequential loops (not the oracle's sorts, prefix sums and binary searches):

* orphans (sources/chaining.cu:260-329, kernels/chaining.cu:237-270): nodes whose first
  candidate is empty, sorted by (size, id); the k-th smallest pairs with the k-th largest
  when the summed sizes and summed inbound counts fit; the group id is the smaller id;
* greedy packing (sources/postprocess.cu:16-245 for -ipm, :247-420 for small partitions):
  items sorted by (size, inbound, id) are cut into consecutive runs, each as long as its
  summed sizes and inbound counts fit (at least one item); a run takes its smallest id;
* final relabel (sources/main.cu:817-830): ids made zero-based in ascending order;
* the initial partitioner's input (sources/init_part.cu:749-888): pruning by degree score
  (the build's integer rule in place of cuRAND Philox), then hedges of >= 2 pins with
  identical pin sets merged and their weights summed.
"""
import zlib

import numpy as np
import pytest

import gen
import oracle as O

NONE = 0xFFFFFFFF
M32 = 0xFFFFFFFF
M64 = (1 << 64) - 1


def seeded(name):
    return np.random.default_rng(zlib.crc32(name.encode()))


def fixed(w):
    v = float(np.float32(w) * np.float32(262144.0))
    if not v > 0:
        return 0
    return M32 if v >= 4294967296.0 else int(v)


# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_orphan_pairing_brute_force(seed):
    rng = seeded(f"orph{seed}")
    n = int(rng.integers(6, 40))
    cnc = 2
    orphan = rng.random(n) < 0.6
    nodes = np.arange(n)
    pairs = np.full(n * cnc, NONE, np.uint32)
    scores = np.zeros(n * cnc, np.uint32)
    others = nodes[~orphan]
    for v in others:  # candidates only among non-orphans, so the matching never reaches an orphan
        cands = [int(x) for x in others if x != v]
        if cands:
            pairs[v * cnc] = cands[int(rng.integers(0, len(cands)))]
            scores[v * cnc] = int(rng.integers(1, 100))
    sizes = rng.integers(1, 6, n).astype(np.uint32)
    inb = rng.integers(0, 8, n).astype(np.uint32)
    omega, delta = 7, 9
    for v in others:  # a non-orphan without candidates would be an orphan too
        if pairs[v * cnc] == NONE:
            orphan[v] = True
    k, groups, gs, ung, uoff, raw = O.group_nodes(n, inb, pairs, scores, sizes, cnc, omega, delta)
    orph = sorted((int(sizes[v]), v) for v in range(n) if orphan[v])
    expect = {}
    m = len(orph)
    for i in range(m // 2):
        a, b = orph[i][1], orph[m - 1 - i][1]
        if sizes[a] + sizes[b] <= omega and inb[a] + inb[b] <= delta:
            expect[a] = expect[b] = min(a, b)
    for v in range(n):
        if orphan[v]:
            assert raw[v] == expect.get(v, v), v
    assert any(orphan) and len(expect) >= 2 or m < 2


def packs_sequential(items, omega, delta):
    """items: [(size, inbound, id)] already in packing order -> list of runs (lists of ids)."""
    runs, i = [], 0
    while i < len(items):
        s, q, run = items[i][0], items[i][1], [items[i][2]]
        j = i + 1
        while j < len(items) and s + items[j][0] <= omega and q + items[j][1] <= delta:
            s += items[j][0]
            q += items[j][1]
            run.append(items[j][2])
            j += 1
        runs.append(run)
        i = j
    return runs


@pytest.mark.parametrize("seed", range(6))
def test_greedy_merge_groups_brute_force(seed):
    rng = seeded(f"gm{seed}")
    n = int(rng.integers(10, 80))
    ng = int(rng.integers(2, n))
    groups = np.concatenate([np.arange(ng), rng.integers(0, ng, n - ng)]).astype(np.uint32)
    rng.shuffle(groups)
    sizes = rng.integers(1, 4, n).astype(np.uint32)
    inb = rng.integers(0, 5, n).astype(np.uint32)
    gsizes = np.bincount(groups, weights=sizes, minlength=ng).astype(np.uint32)
    omega, delta = int(rng.integers(4, 12)), int(rng.integers(4, 16))
    k, got, gsz = O.greedy_merge_groups(n, ng, sizes, inb, omega, delta, groups, gsizes)
    gin = np.bincount(groups, weights=inb, minlength=ng).astype(np.int64)
    items = sorted((int(gsizes[g]), int(gin[g]), g) for g in range(ng))
    runs = packs_sequential(items, omega, delta)
    for r in runs:  # each run fits, and the next item would not have (the runs are maximal)
        assert len(r) == 1 or (sum(gsizes[g] for g in r) <= omega and sum(gin[g] for g in r) <= delta)
    rep = {g: min(r) for r in runs for g in r}
    reps = sorted(set(rep.values()))
    new_id = {x: i for i, x in enumerate(reps)}
    expect = np.array([new_id[rep[int(g)]] for g in groups], np.uint32)
    assert k == len(runs) and np.array_equal(got, expect)
    assert np.array_equal(gsz, np.bincount(expect, weights=sizes, minlength=k).astype(np.uint32))


@pytest.mark.parametrize("seed", range(6))
def test_merge_small_partitions_brute_force(seed):
    rng = seeded(f"ms{seed}")
    k = int(rng.integers(5, 60))
    n = 400
    part = rng.integers(0, k, n).astype(np.uint32)
    psizes = rng.integers(1, 40, k).astype(np.uint32)  # about a third below the threshold of 16
    pinb = rng.integers(0, 30, k).astype(np.uint32)
    omega, delta = 40, int(rng.integers(20, 80))
    got = O.merge_small_partitions(n, k, psizes, pinb, omega, delta, part)
    items = sorted((int(psizes[p]), int(pinb[p]), p) for p in range(k) if psizes[p] < 16)
    runs = packs_sequential(items, omega, delta)
    pmap = {p: p for p in range(k)}
    for r in runs:
        for p in r:
            pmap[p] = min(r)
    expect = np.array([pmap[int(p)] for p in part], np.uint32)
    assert np.array_equal(got, expect)
    assert any(len(r) > 1 for r in runs) or len(items) < 2


@pytest.mark.parametrize("seed", range(4))
def test_relabel_is_dense_and_order_preserving(seed):
    rng = seeded(f"rl{seed}")
    part = rng.choice(rng.integers(0, 1 << 31, 50), 500).astype(np.uint32)
    k, got = O.relabel(part)
    ids = sorted(set(int(x) for x in part))
    assert k == len(ids)
    assert np.array_equal(got, np.array([ids.index(int(x)) for x in part], np.uint32))


# ---------------------------------------------------------------------------------------
def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def prune_rule(hs, w, n, mult=1):
    """Keep draw and scaled weight per hedge: 0 = dropped.  The keep target is mult times the
    reference's N * floor(P / E)."""
    e = len(hs)
    P = sum(len(h) for h in hs)
    deg = [0] * n
    for h in hs:
        for v in h:
            deg[v] += 1
    cap = M64 // e
    ratio = []
    for x, h in enumerate(hs):
        inv = sum((1 << 24) // deg[v] for v in h)
        ratio.append(min((fixed(w[x]) << 24) // inv, cap) if inv else 0)
    S = sum(ratio)
    assert S <= M64
    T = mult * n * (P // e)
    out, probs = [], []
    for x in range(e):
        p = 1.0 if S == 0 else min(1.0, float(T) * float(ratio[x]) / float(S))
        probs.append(p)
        u = splitmix64(86 ^ ((x * 0x9E3779B97F4A7C15) & M64)) >> 11
        if float(u) < p * 2.0 ** 53:
            out.append(min(int(float(fixed(w[x])) / p), 1 << 40))
        else:
            out.append(0)
    return out, probs, T


def ip_input_sets(hs, w, n, prune):
    weights = prune_rule(hs, w, n, int(prune))[0] if prune else [fixed(x) for x in w]
    order, total = [], {}
    for x, h in enumerate(hs):
        if prune and weights[x] == 0:
            continue
        s = tuple(sorted(set(h)))
        if len(s) < 2:
            continue
        if s not in total:
            order.append(s)
            total[s] = 0
        total[s] += weights[x]
    return order, [total[s] for s in order]


def hedge_lists(g):
    return [[int(x) for x in g.pins[g.off[i]:g.off[i + 1]]] for i in range(g.e)]


@pytest.mark.parametrize("prune", [0, 1, 3])
@pytest.mark.parametrize("seed,n,e", [(0, 40, 300), (1, 200, 2000), (2, 12, 400)])
def test_ip_input_is_the_merged_set_family(seed, n, e, prune):
    rng = seeded(f"ipin{seed}")
    hs = []
    for _ in range(e):
        if hs and rng.random() < 0.3:
            h = list(hs[int(rng.integers(0, len(hs)))])
            rng.shuffle(h)
        else:
            h = [int(x) for x in rng.choice(n, size=int(rng.integers(1, 6)), replace=False)]
            if rng.random() < 0.1:
                h.append(h[0])
        hs.append(h)
    off = np.zeros(e + 1, np.uint64)
    off[1:] = np.cumsum([len(h) for h in hs])
    w = (rng.random(e) * 3).astype(np.float32)
    g = O.HG(n, off, np.array([x for h in hs for x in h], np.uint32), np.zeros(e, np.uint32), w)
    sizes = rng.integers(1, 5, n).astype(np.uint32)
    goff, gpins, gw, gvw = O.ip_input(g, sizes, prune=prune)
    sets, weights = ip_input_sets(hs, w, n, prune)
    got = [tuple(int(x) for x in gpins[goff[i]:goff[i + 1]]) for i in range(len(goff) - 1)]
    assert got == sets
    assert gw.tolist() == weights
    assert np.array_equal(gvw, sizes.astype(np.uint64))


@pytest.mark.parametrize("mult", [1, 4])
def test_pruning_keeps_about_the_target_and_preserves_total_weight(mult):
    """The reference's intent (kernels/init_part.cu:221-281): about mult * N * floor(P / E)
    hedges are kept, hedges whose probability reaches one always stay with their weight, and
    the scaled weights estimate the total weight without bias."""
    rng = seeded("prune-stats")
    n, e = 2000, 60000
    g = O.HG(*gen.random_hypergraph(n, e, rng, max_size=8, weights="float"))
    hs = hedge_lists(g)
    out, probs, T = prune_rule(hs, g.w, n, mult)
    kept = sum(1 for x in out if x)
    expect = sum(probs)
    assert T == mult * n * (len(g.pins) // e)
    assert abs(kept - expect) < 5 * np.sqrt(expect)
    assert expect <= T * 1.0001
    W = [fixed(x) for x in g.w]
    for x, p in enumerate(probs):
        if p == 1.0:
            assert out[x] == W[x]
    tot, est = sum(W), sum(out)
    assert abs(est - tot) / tot < 0.05
    # the oracle applies the same rule
    sizes = np.ones(n, np.uint32)
    goff, gpins, gw, _ = O.ip_input(g, sizes, prune=mult)
    sets, weights = ip_input_sets(hs, g.w, n, mult)
    assert gw.tolist() == weights and len(goff) - 1 == len(sets)


def test_initial_partitioner_invariants():
    """The built-in partitioner (unpinned by nature: Mt-KaHyPar is absent) places every node
    in [0, k), keeps every part within omega and reports the part weights."""
    rng = seeded("ip-inv")
    g = O.dedup(O.HG(*gen.random_hypergraph(3000, 3000, rng, max_size=10, weights="int")))
    sizes = rng.integers(1, 4, g.n).astype(np.uint32)
    for k in (2, 7, 16):
        omega = O.kway_omega(int(sizes.sum()), k, 0.03)
        for prune in (False, True):
            part, ps = O.initial_partition(g, sizes, k, 0.03, omega, prune=prune)
            assert part.max() < k
            assert np.array_equal(ps, np.bincount(part, weights=sizes, minlength=k).astype(np.uint32))
            assert ps.max() <= omega


def test_initial_partitioner_rejects_infeasible_limits():
    rng = seeded("ip-inf")
    g = O.dedup(O.HG(*gen.random_hypergraph(50, 60, rng, max_size=5)))
    sizes = np.full(g.n, 3, np.uint32)
    with pytest.raises(O.OracleError) as ex:
        O.initial_partition(g, sizes, 4, 0.03, 1)  # every node is heavier than the limit
    assert ex.value.status == 6
    with pytest.raises(O.OracleError) as ex:
        O.initial_partition(g, sizes, 2, 0.03, 60)  # 150 units of weight, 2 parts of 60
    assert ex.value.status == 6


def test_kway_partition_rejects_max_parts_other_than_k():
    rng = seeded("kway-mp")
    g = O.dedup(O.HG(*gen.random_hypergraph(100, 100, rng, max_size=5)))
    with pytest.raises(O.OracleError) as ex:
        O.partition(g, 1, O.kway_omega(g.n, 4, 0.03), 2**31 - 1, 8, kway=4, eps=0.03)
    assert ex.value.status == 1


def test_initial_partitioner_is_independent_of_threads():
    """The same partition on one thread and on up to 16 (tries and halves in parallel)."""
    rng = seeded("ip-threads")
    g = O.dedup(O.HG(*gen.random_hypergraph(20000, 24000, rng, max_size=12, weights="int")))
    sizes = np.ones(g.n, np.uint32)
    omega = O.kway_omega(g.n, 16, 0.03)
    try:
        O.set_ip_threads(1)
        one, ps1 = O.initial_partition(g, sizes, 16, 0.03, omega, prune=False)
    finally:
        O.set_ip_threads(0)
    many, ps2 = O.initial_partition(g, sizes, 16, 0.03, omega, prune=False)
    assert np.array_equal(one, many) and np.array_equal(ps1, ps2)
