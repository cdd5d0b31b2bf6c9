"""GPU parity: every stage through the C ABI against the CPU oracle, bit-exact.

Inputs are seeded; sizes are chosen so the oracle finishes in seconds while still covering
the kernels' special paths (LDS vs global tables, small vs large hedges, cycles in pairing)."""
import ctypes as C

import numpy as np
import pytest

import gen
import gpu_stages as G
import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    # name, n, e, max_size, src_prob, max_src, dup, weights
    ("hgr_small", 300, 400, 6, 0.0, 1, 0.0, "unit"),
    ("hgr_wide", 2000, 1500, 40, 0.0, 1, 0.0, "int"),
    ("srcs", 1500, 1800, 12, 0.9, 3, 0.3, "int"),
    ("float_w", 800, 900, 9, 0.5, 2, 0.2, "float"),
    ("big_hedges", 3000, 120, 1200, 0.3, 2, 0.0, "unit"),
]


def make(case, rng):
    _, n, e, ms, sp, mxs, dup, w = case
    raw = O.HG(*gen.random_hypergraph(n, e, rng, max_size=ms, src_prob=sp, max_src=mxs, dup_prob=dup, weights=w))
    return O.dedup(raw)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_touching(dev, case, rng):
    hg = make(case, rng)
    L = G.Level(dev, hg)
    toff, tch, inb = L.build_touching()
    ro, rt, ri = O.touching(hg)
    assert np.array_equal(toff, ro) and np.array_equal(tch, rt) and np.array_equal(inb, ri)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_neighbors(dev, case, rng):
    hg = make(case, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    noff, nbr, _, _ = G.neighbors(dev, L)
    ro, rt, _ = O.touching(hg)
    eo, en = O.neighbors(hg, ro, rt)
    assert np.array_equal(noff, eo) and np.array_equal(nbr, en)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("cnc", [1, 4, 16])
def test_candidates(dev, case, cnc, rng):
    hg = make(case, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    sizes = rng.integers(1, 4, size=hg.n).astype(np.uint32)
    for omega, delta in ((6, 2**31 - 1), (4, 40)):
        pairs, scores, *_ = G.candidates(dev, L, sizes, cnc, omega, delta)
        ep, es = O.candidates(hg, ro, rt, ri, sizes, cnc, omega, delta)
        assert np.array_equal(pairs, ep), f"pairs differ (omega={omega}, delta={delta})"
        assert np.array_equal(scores, es)


def tier_graph(n, big_sizes, n_small, rng):
    """Mostly small hedges plus a few very large ones, so that their pins need candidate
    tables of every size class (LDS hash, LDS indexed by node id, HBM hash, HBM indexed)."""
    off, pins, srcc = [0], [], []
    for sz in big_sizes:
        nodes = rng.choice(n, size=sz, replace=False)
        pins.extend(int(x) for x in nodes)
        off.append(len(pins))
        srcc.append(1)
    for _ in range(n_small):
        sz = int(rng.integers(2, 6))
        nodes = rng.choice(n, size=sz, replace=False)
        pins.extend(int(x) for x in nodes)
        off.append(len(pins))
        srcc.append(int(rng.integers(0, 2)))
    e = len(srcc)
    w = (rng.random(e) * 3 + 0.5).astype(np.float32)
    return O.dedup(O.HG(n, np.array(off, np.uint64), np.array(pins, np.uint32), np.array(srcc, np.uint32), w))


@pytest.mark.parametrize("n,big,deltas", [
    (40000, (2500, 5000, 9000), (2**31 - 1, 30)),  # LDS hash (big), HBM hash (heavy)
    (20000, (9000,), (2**31 - 1, 30)),             # LDS indexed by id; HBM indexed by id with inbound counts
])
def test_candidates_table_tiers(dev, n, big, deltas, rng):
    hg = tier_graph(n, big, 3000, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    sizes = rng.integers(1, 3, size=hg.n).astype(np.uint32)
    for delta in deltas:
        pairs, scores, *_ = G.candidates(dev, L, sizes, 16, 6, delta)
        ep, es = O.candidates(hg, ro, rt, ri, sizes, 16, 6, delta)
        assert np.array_equal(pairs, ep), f"pairs differ (delta={delta})"
        assert np.array_equal(scores, es), f"scores differ (delta={delta})"


def hub_graph(n, hubs, n_small, rng):
    """Hub nodes, each on `count` hedges of `size` pins drawn from a pool of `pool` nodes, so
    their visit bounds (and worst-case tables) are far above their distinct neighbours."""
    off, pins, srcc = [0], [], []
    for hub, count, size, pool in hubs:
        for _ in range(count):
            nodes = rng.choice(pool, size=size, replace=False)
            nodes = nodes[nodes != hub][: size - 1]
            pins.append(hub)
            pins.extend(int(x) for x in nodes)
            off.append(len(pins))
            srcc.append(1)
    for _ in range(n_small):
        sz = int(rng.integers(2, 6))
        pins.extend(int(x) for x in rng.choice(n, size=sz, replace=False))
        off.append(len(pins))
        srcc.append(int(rng.integers(0, 2)))
    w = (rng.random(len(srcc)) * 3 + 0.5).astype(np.float32)
    return O.dedup(O.HG(n, np.array(off, np.uint64), np.array(pins, np.uint32), np.array(srcc, np.uint32), w))


@pytest.mark.parametrize("n,hubs,expect", [
    # (hub, hedges, pins per hedge, pool): distinct neighbours ~20K, ~45K and ~75K
    (120000, [(7, 75, 400, 30000), (11, 141, 400, 120000), (13, 295, 400, 120000)],
     [(15000, 25000), (40400, 49000), (70000, 80000)]),
    # n = 50000: the hub's worst-case table is the whole level, indexed by node id
    (50000, [(5, 229, 400, 50000)], [(40400, 50000)]),
])
def test_candidates_heavy_tables(dev, n, hubs, expect, rng):
    """Nodes whose visit bound is far above their distinct neighbours, so they skip the
    guaranteed tiers: they are scored in the one-pass LDS table, the four-pass LDS table
    (one id class per pass), a 65536-slot HBM table at up to 3/4 load, or their worst-case
    HBM table (hashed, or indexed by node id), depending on the distinct count and on
    whether shared-inbound counts are kept (they shrink the LDS tables)."""
    hg = hub_graph(n, hubs, 4000, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    for (h, *_), (lo, hi) in zip(hubs, expect):
        idx = np.concatenate([np.arange(int(hg.off[e]), int(hg.off[e + 1])) for e in rt[ro[h]:ro[h + 1]]])
        nb = len(np.unique(hg.pins[idx]))
        assert lo < nb < hi, (h, nb)
    sizes = rng.integers(1, 3, size=hg.n).astype(np.uint32)
    for delta in (2**31 - 1, 30):
        pairs, scores, *_ = G.candidates(dev, L, sizes, 16, 6, delta)
        ep, es = O.candidates(hg, ro, rt, ri, sizes, 16, 6, delta)
        assert np.array_equal(pairs, ep), f"pairs differ (delta={delta})"
        assert np.array_equal(scores, es), f"scores differ (delta={delta})"


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("cnc", [1, 3, 16])
def test_group_nodes(dev, case, cnc, rng):
    hg = make(case, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    sizes = np.ones(hg.n, np.uint32)
    omega, delta = 8, 2**31 - 1
    pairs, scores, d_sizes, d_pairs, d_scores = G.candidates(dev, L, sizes, cnc, omega, delta)
    k, groups, gs, ung, uoff, *_ = G.group_nodes(dev, hg.n, L.inb, d_pairs, d_scores, d_sizes, cnc, omega, delta)
    ek, eg, egs, eung, euoff, _ = O.group_nodes(hg.n, ri, pairs, scores, sizes, cnc, omega, delta)
    assert k == ek
    assert np.array_equal(groups, eg) and np.array_equal(gs, egs)
    assert np.array_equal(ung, eung) and np.array_equal(uoff, euoff)
    # invariants: groups of at most two nodes, each pair a candidate edge or an orphan pair
    counts = np.bincount(groups, minlength=k)
    assert counts.max() <= 2


def test_group_nodes_cycles(dev, rng):
    """Hand-made pointer rounds with 2-cycles and a 3-cycle in round 1."""
    n, cnc = 9, 2
    pairs = np.full(n * cnc, 0xFFFFFFFF, np.uint32)
    scores = np.zeros(n * cnc, np.uint32)

    def setp(v, i, t, s):
        pairs[v * cnc + i] = t
        scores[v * cnc + i] = s

    # round 0: chain 0->1->2, mutual 3<->4, 5 alone
    setp(0, 0, 1, 5); setp(1, 0, 2, 7); setp(2, 0, 1, 7); setp(3, 0, 4, 9); setp(4, 0, 3, 9)
    # round 1: 3-cycle 6->7->8->6 plus 5->6
    setp(6, 1, 7, 3); setp(7, 1, 8, 4); setp(8, 1, 6, 2); setp(5, 1, 6, 1)
    setp(6, 0, 0xFFFFFFFF, 0)
    inb = np.zeros(n, np.uint32)
    sizes = np.ones(n, np.uint32)
    d_inb, d_pairs, d_scores, d_sizes = (dev.upload(x) for x in (inb, pairs, scores, sizes))
    k, groups, *_ = G.group_nodes(dev, n, d_inb, d_pairs, d_scores, d_sizes, cnc, 10, 100)
    ek, eg, *_ = O.group_nodes(n, inb, pairs, scores, sizes, cnc, 10, 100)
    assert k == ek and np.array_equal(groups, eg)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_coarsening(dev, case, rng):
    """coarse hedges, coarse touching, coarse neighbours, uncoarsen"""
    hg = make(case, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    sizes = np.ones(hg.n, np.uint32)
    cnc = 4
    pairs, scores, d_sizes, d_pairs, d_scores = G.candidates(dev, L, sizes, cnc, 8, 2**31 - 1)
    k, groups, gs, ung, uoff, d_groups, *_ = G.group_nodes(dev, hg.n, L.inb, d_pairs, d_scores, d_sizes, cnc, 8,
                                                           2**31 - 1)
    choff, cp, csrcc, d_ch, d_co, d_cc = G.coarsen_hedges(dev, L, d_groups)
    eo, ep, es = O.coarsen_hedges(hg, groups)
    assert np.array_equal(choff, eo) and np.array_equal(cp, ep) and np.array_equal(csrcc, es)
    cto, ctt, cti = G.coarsen_touching(dev, d_ch, d_co, d_cc, k, hg.e)
    chg = O.HG(k, eo, ep, es, hg.w)
    xo, xt, xi = O.touching(chg)
    assert np.array_equal(cto, xo) and np.array_equal(ctt, xt) and np.array_equal(cti, xi)
    noff, nbr, d_nbr, d_noff = G.neighbors(dev, L)
    cno, cnb = G.coarsen_neighbors(dev, d_nbr, d_noff, d_groups, hg.n, k)
    eno, enb = O.coarsen_neighbors(hg.n, noff, nbr, groups, k)
    assert np.array_equal(cno, eno) and np.array_equal(cnb, enb)
    # the coarse neighbourhoods equal the neighbourhoods of the coarse hypergraph
    yo, yn = O.neighbors(chg, xo, xt)
    assert np.array_equal(cno, yo) and np.array_equal(cnb, yn)
    # uncoarsen
    cpart = rng.integers(0, 5, size=k).astype(np.uint32)
    d_cpart = dev.upload(cpart)
    d_part = dev.alloc(hg.n * 4)
    dev.check(dev.lib.hgp_uncoarsen(dev.ctx, G.V(d_groups), G.V(d_cpart), hg.n, G.V(d_part)), "uncoarsen")
    assert np.array_equal(dev.download(d_part, hg.n, np.uint32), cpart[groups])


REFINE_CASES = [
    ("kway_k2", 1, 2, "hgr_small"),
    ("kway_k8", 1, 8, "hgr_wide"),
    ("kway_k5_srcs", 1, 5, "srcs"),
    ("incc_k12", 0, 12, "srcs"),
    ("incc_k40_float", 0, 40, "float_w"),
    ("kway_big", 1, 4, "big_hedges"),
    # k-way up to 64 partitions scores from partition masks (both 32-bit words in use here),
    # above 64 from partition lists
    ("kway_k64_wide", 1, 64, "hgr_wide"),
    ("kway_k48_float", 1, 48, "float_w"),
    ("kway_k100_lists", 1, 100, "hgr_wide"),
]


def refine_setup(case_name, k, mode, rng):
    case = [c for c in CASES if c[0] == case_name][0]
    hg = make(case, rng)
    sizes = rng.integers(1, 3, size=hg.n).astype(np.uint32)
    part = rng.integers(0, k, size=hg.n).astype(np.uint32)
    psizes = np.bincount(part, weights=sizes, minlength=k).astype(np.uint32)
    if mode == 1:
        omega = int(np.ceil(1.08 * sizes.sum() / k))
        delta = 2**31 - 1
    else:
        omega = int(psizes.max()) + 3
        inbs = O.inbound_sizes(hg, part, k)
        delta = int(inbs.max()) + 2
    return hg, sizes, part, psizes, omega, delta


@pytest.mark.parametrize("rc", REFINE_CASES, ids=[r[0] for r in REFINE_CASES])
@pytest.mark.parametrize("chainup", [False, True])
@pytest.mark.parametrize("fm_repeat", [0, 6, 13])
def test_refine_analyze(dev, rc, chainup, fm_repeat, rng):
    name, mode, k, case_name = rc
    hg, sizes, part, psizes, omega, delta = refine_setup(case_name, k, mode, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    got = G.refine_analyze(dev, L, sizes, k, mode, omega, delta, 16, fm_repeat, chainup, part, psizes)
    exp = O.refine_analyze(mode, omega, delta, 16, hg, ro, rt, ri, sizes, k, fm_repeat, chainup, part, psizes)
    for key in ("moves", "gains", "ranks", "seq", "sv", "iv"):
        assert np.array_equal(got[key], exp[key]), f"{key} differs"
    assert got["best"] == exp["best"]


@pytest.mark.parametrize("mode,k,fm_repeat,chainup,n,e,max_size", [
    (1, 4096, 0, False, 4500, 10, 5000),   # gains table in LDS indexed by partition
    (0, 3000, 6, True, 4500, 10, 5000),
    (1, 22000, 0, True, 24000, 8, 9000),   # hashed LDS gains tables, some overflowing to HBM
    (0, 22000, 13, False, 24000, 8, 9000),
])
def test_refine_analyze_huge_lists(dev, mode, k, fm_repeat, chainup, n, e, max_size, rng):
    """Hedges of thousands of pins over thousands of partitions: partition lists longer than
    the LDS tables take the HBM-table path, and nodes adjacent to many partitions use the
    workgroup gain kernel (LDS tables indexed or hashed, HBM tables past 3/4 load)."""
    raw = O.HG(*gen.random_hypergraph(n, e, rng, max_size=max_size, src_prob=0.3, max_src=2, weights="unit"))
    hg = O.dedup(raw)
    sizes = rng.integers(1, 3, size=hg.n).astype(np.uint32)
    part = rng.integers(0, k, size=hg.n).astype(np.uint32)
    psizes = np.bincount(part, weights=sizes, minlength=k).astype(np.uint32)
    lam = [len(np.unique(part[hg.pins[hg.off[e]:hg.off[e + 1]]])) for e in range(len(hg.off) - 1)]
    assert max(lam) > 2016, "the case must reach the HBM-table path"
    if mode == 1:
        omega, delta = int(np.ceil(1.08 * sizes.sum() / k)), 2**31 - 1
    else:
        omega, delta = int(psizes.max()) + 3, int(O.inbound_sizes(hg, part, k).max()) + 2
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    got = G.refine_analyze(dev, L, sizes, k, mode, omega, delta, 16, fm_repeat, chainup, part, psizes)
    exp = O.refine_analyze(mode, omega, delta, 16, hg, ro, rt, ri, sizes, k, fm_repeat, chainup, part, psizes)
    for key in ("moves", "gains", "ranks", "seq", "sv", "iv"):
        assert np.array_equal(got[key], exp[key]), f"{key} differs"
    assert got["best"] == exp["best"]


@pytest.mark.parametrize("rc", REFINE_CASES, ids=[r[0] for r in REFINE_CASES])
def test_refine_loop(dev, rc, rng):
    name, mode, k, case_name = rc
    hg, sizes, part, psizes, omega, delta = refine_setup(case_name, k, mode, rng)
    L = G.Level(dev, hg)
    L.build_touching()
    ro, rt, ri = O.touching(hg)
    gp, gps, gpin = G.refine(dev, L, sizes, k, mode, omega, delta, 16, True, part, psizes)
    ep, eps_, epin = O.refine(mode, omega, delta, 16, hg, ro, rt, ri, sizes, k, True, part, psizes)
    assert np.array_equal(gp, ep) and np.array_equal(gps, eps_) and np.array_equal(gpin, epin)
    # the refined state is consistent and no worse than the start
    assert np.array_equal(np.bincount(gp, weights=sizes, minlength=k).astype(np.uint32), gps)
    assert O.connectivity_fixed(hg, gp) <= O.connectivity_fixed(hg, part)


def test_postprocess(dev, rng):
    n = 500
    ng = 120
    groups = rng.integers(0, ng, size=n).astype(np.uint32)
    groups = O.relabel(groups)[1]
    ng = int(groups.max()) + 1
    sizes = rng.integers(1, 4, size=n).astype(np.uint32)
    inb = rng.integers(0, 20, size=n).astype(np.uint32)
    gsizes = np.bincount(groups, weights=sizes, minlength=ng).astype(np.uint32)
    omega, delta = 60, 300
    d_sizes, d_inb, d_groups, d_gs = (dev.upload(x) for x in (sizes, inb, groups, gsizes))
    new = C.c_uint32()
    dev.check(dev.lib.hgp_greedy_merge_groups(dev.ctx, G.V(d_sizes), G.V(d_inb), n, ng, omega, delta, G.V(d_groups),
                                              G.V(d_gs), C.byref(new)), "greedy merge")
    en, eg, egs = O.greedy_merge_groups(n, ng, sizes, inb, omega, delta, groups, gsizes)
    assert new.value == en
    assert np.array_equal(dev.download(d_groups, n, np.uint32), eg)
    assert np.array_equal(dev.download(d_gs, en, np.uint32), egs)
    # small partitions merge + relabel
    k = 80
    part = rng.integers(0, k, size=n).astype(np.uint32)
    psz = np.bincount(part, minlength=k).astype(np.uint32)
    pinb = rng.integers(0, 50, size=k).astype(np.uint32)
    d_part, d_psz, d_pinb = (dev.upload(x) for x in (part, psz, pinb))
    dev.check(dev.lib.hgp_merge_small_partitions(dev.ctx, G.V(d_psz), G.V(d_pinb), n, k, 40, 200, G.V(d_part)), "msp")
    ep = O.merge_small_partitions(n, k, psz, pinb, 40, 200, part)
    gp = dev.download(d_part, n, np.uint32)
    assert np.array_equal(gp, ep)
    nk = C.c_uint32()
    dev.check(dev.lib.hgp_relabel_partitions(dev.ctx, n, G.V(d_part), C.byref(nk)), "relabel")
    ek, er = O.relabel(ep)
    assert nk.value == ek and np.array_equal(dev.download(d_part, n, np.uint32), er)


@pytest.mark.parametrize("case", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_inbound_and_connectivity(dev, case, rng):
    hg = make(case, rng)
    L = G.Level(dev, hg)
    k = 7
    part = rng.integers(0, k, size=hg.n).astype(np.uint32)
    d_part = dev.upload(part)
    d_out = dev.alloc(k * 4)
    dev.check(dev.lib.hgp_inbound_sizes(dev.ctx, G.V(L.pins), G.V(L.off), G.V(L.srcc), G.V(d_part), hg.e, k,
                                        G.V(d_out)), "inbound sizes")
    assert np.array_equal(dev.download(d_out, k, np.uint32), O.inbound_sizes(hg, part, k))
    km1 = C.c_int64()
    dev.check(dev.lib.hgp_connectivity(dev.ctx, G.V(L.pins), G.V(L.off), G.V(L.w), G.V(d_part), hg.e, C.byref(km1)),
              "connectivity")
    assert km1.value == O.connectivity_fixed(hg, part)


def ip_input_hypergraph(n, e, max_size, copies, rng):
    """Hedges with exact copies, copies with their pins permuted, repeated pins, single-pin
    hedges and weights that round to zero in fixed point."""
    hedges = []
    for _ in range(e):
        if hedges and rng.random() < copies:
            h = list(hedges[int(rng.integers(0, len(hedges)))])
            rng.shuffle(h)
        else:
            h = [int(x) for x in rng.choice(n, size=int(rng.integers(1, max_size + 1)), replace=False)]
            if rng.random() < 0.1:
                h.append(h[0])  # a repeated pin
        hedges.append(h)
    off = np.zeros(e + 1, np.uint64)
    off[1:] = np.cumsum([len(h) for h in hedges])
    pins = np.array([x for h in hedges for x in h], np.uint32)
    w = (rng.random(e) * 3).astype(np.float32)
    w[rng.random(e) < 0.05] = 1e-7
    return O.HG(n, off, pins, np.zeros(e, np.uint32), w)


def device_ip_input(dev, L, d_sizes, n, e, prune):
    m, p = C.c_uint32(), C.c_uint64()
    args = (dev.ctx, G.V(L.pins), G.V(L.off), G.V(L.w), G.V(d_sizes), n, e, int(prune), C.byref(m), C.byref(p))
    dev.check(dev.lib.hgp_ip_input(*args, None, None, None), "ip input sizes")
    off = np.zeros(m.value + 1, np.uint64)
    pins = np.zeros(p.value, np.uint32)
    w = np.zeros(m.value, np.uint64)
    dev.check(dev.lib.hgp_ip_input(*args, off.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   pins.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   w.ctypes.data_as(C.POINTER(C.c_uint64))), "ip input")
    return off, pins, w


@pytest.mark.parametrize("prune", [0, 1, 4])
@pytest.mark.parametrize("n,e,max_size,copies", [(300, 2000, 5, 0.3), (5000, 40000, 12, 0.2), (64, 3000, 4, 0.6)])
def test_ip_input_device_equals_oracle(dev, n, e, max_size, copies, prune, rng):
    """The initial partitioner's input built on the device (hedge pruning, sort, merge of
    identical pin sets, summed weights) equals the oracle's restatement exactly."""
    hg = ip_input_hypergraph(n, e, max_size, copies, rng)
    L = G.Level(dev, hg)
    sizes = rng.integers(1, 5, size=n).astype(np.uint32)
    d_sizes = dev.upload(sizes)
    off, pins, w = device_ip_input(dev, L, d_sizes, n, e, prune)
    eo, ep, ew, _ = O.ip_input(hg, sizes, prune=prune)
    assert np.array_equal(off, eo) and np.array_equal(pins, ep) and np.array_equal(w, ew)
    assert 0 < len(w) < e


def test_initial_partition(dev, rng):
    hg = make(CASES[1], rng)
    L = G.Level(dev, hg)
    sizes = rng.integers(1, 5, size=hg.n).astype(np.uint32)
    k, eps = 8, 0.03
    omega = O.kway_omega(int(sizes.sum()), k, eps)
    d_sizes = dev.upload(sizes)
    pp, ps = G.V(), G.V()
    dev.check(dev.lib.hgp_initial_partition(dev.ctx, G.V(L.pins), G.V(L.off), G.V(L.w), hg.n, hg.e, L.p, G.V(d_sizes),
                                            k, eps, omega, 1, C.byref(pp), C.byref(ps)), "initial partition")
    gp = dev.download(pp.value, hg.n, np.uint32)
    ep, eps_ = O.initial_partition(hg, sizes, k, eps, omega)
    assert np.array_equal(gp, ep)
    assert np.array_equal(dev.download(ps.value, k, np.uint32), eps_)
    assert eps_.max() <= omega
