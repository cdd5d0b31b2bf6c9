"""Pins the CPU oracle against brute force and closed forms on tiny inputs (CPU only).

The reference has no tests; these checks fix what the mathematics fixes, each written from
the definition (Python sets and direct simulation) rather than from the oracle's code:

* touching lists = transpose of the incidence, inbound then outbound, each ascending;
* neighbours = union of the touching hedges' pins minus the node itself;
* candidate scores = noise + sum of floor(W_e / |e|) over shared hedges (u32), filtered by
  size and inbound-union limits, top-cnc by (score desc, id desc);
* round-0 pairing = a maximum-weight matching of the candidate forest;
* coarse hedges = set of source groups minus destination groups (ascending), then the set of
  destination groups (descending);
* k-way gains = exact change of sum W_e (lambda_e - 1) for the single move;
* cascade prefix sums = exact change of that sum when the moves are applied in rank order;
* size / inbound violation prefixes = direct recount after applying moves 0..r;
* best prefix = the ordering (inbound ok, fewer size violations, larger gain, earlier rank);
* noise symmetric and in [0, 64), the gain tie mixer a bijection, fixed point = floor(w * 2^18).
"""
import itertools
import zlib

import numpy as np
import pytest

import gen
import oracle as O

NONE = 0xFFFFFFFF
M32 = 0xFFFFFFFF


def seeded(name):
    return np.random.default_rng(zlib.crc32(name.encode()))


def make(n, e, seed, src_prob=0.5, max_size=6, weights="float"):
    rng = seeded(f"hg{seed}")
    g = O.HG(*gen.random_hypergraph(n, e, rng, max_size=max_size, src_prob=src_prob, max_src=2, dup_prob=0.3,
                                    weights=weights))
    return O.dedup(g)


def hedges(g):
    out = []
    for i in range(g.e):
        seg = [int(x) for x in g.pins[g.off[i]:g.off[i + 1]]]
        s = int(g.srcc[i])
        out.append((seg[:s], seg[s:]))
    return out


def fixed(w):
    v = float(np.float32(w) * np.float32(262144.0))
    if not v > 0:
        return 0
    return M32 if v >= 4294967296.0 else int(v)


def conn_fixed(g, part):
    tot = 0
    for i, (s, d) in enumerate(hedges(g)):
        lam = len({int(part[x]) for x in s + d})
        if lam:
            tot += fixed(g.w[i]) * (lam - 1)
    return tot


# ---------------------------------------------------------------------------------------
def test_fixed_point_noise_and_mixer():
    for w in (0.0, -1.0, 1.0, 0.5, 3.75, 1e-7, 1e6, 2e4, np.float32(16383.99)):
        assert O.lib().orc_fixed_weight(float(w)) == fixed(w)
    rng = seeded("noise")
    a = rng.integers(0, 2**32, 2000, dtype=np.uint64)
    b = rng.integers(0, 2**32, 2000, dtype=np.uint64)
    for x, y in zip(a, b):
        x, y = int(x), int(y)
        v = O.lib().orc_noise(x, y)
        assert v == O.lib().orc_noise(y, x) and 0 <= v < 64
    vals = {O.lib().orc_hash_uint32(x) for x in range(1 << 16)}
    assert len(vals) == 1 << 16  # xor-shift / odd-multiply steps are invertible


@pytest.mark.parametrize("seed", range(4))
def test_touching_is_the_transpose(seed):
    g = make(30, 50, seed)
    toff, tch, inb = O.touching(g)
    hs = hedges(g)
    for v in range(g.n):
        ins = sorted(i for i, (s, d) in enumerate(hs) if v in d)
        outs = sorted(i for i, (s, d) in enumerate(hs) if v in s)
        assert list(tch[toff[v]:toff[v + 1]]) == ins + outs
        assert inb[v] == len(ins)


@pytest.mark.parametrize("seed", range(4))
def test_neighbours_are_the_union(seed):
    g = make(30, 40, seed)
    toff, tch, inb = O.touching(g)
    noff, nbr = O.neighbors(g, toff, tch)
    hs = hedges(g)
    for v in range(g.n):
        exp = set()
        for s, d in hs:
            if v in s or v in d:
                exp |= set(s + d)
        exp.discard(v)
        assert sorted(nbr[noff[v]:noff[v + 1]]) == sorted(exp)


@pytest.mark.parametrize("seed,cnc", [(0, 1), (1, 4), (2, 16), (3, 3)])
def test_candidates_brute_force(seed, cnc):
    g = make(24, 40, seed, weights="float")
    toff, tch, inb = O.touching(g)
    rng = seeded(f"sizes{seed}")
    sizes = rng.integers(1, 4, g.n).astype(np.uint32)
    omega, delta = 5, 6
    pairs, scores = O.candidates(g, toff, tch, inb, sizes, cnc, omega, delta)
    hs = hedges(g)
    inset = [{i for i, (s, d) in enumerate(hs) if v in d} for v in range(g.n)]
    for u in range(g.n):
        cand = []
        for v in range(g.n):
            if v == u:
                continue
            shared = [i for i, (s, d) in enumerate(hs) if u in s + d and v in s + d]
            if not shared:
                continue
            sc = sum(fixed(g.w[i]) // len(hs[i][0] + hs[i][1]) for i in shared) & M32
            if sizes[u] + sizes[v] > omega or len(inset[u] | inset[v]) > delta:
                continue
            sc = (sc + O.lib().orc_noise(u, v)) & M32
            if sc:
                cand.append((sc, v))
        cand.sort(key=lambda t: (-t[0], -t[1]))
        cand = cand[:cnc] + [(0, NONE)] * (cnc - len(cand[:cnc]))
        assert list(pairs[u * cnc:(u + 1) * cnc]) == [v for _, v in cand]
        assert list(scores[u * cnc:(u + 1) * cnc]) == [s for s, _ in cand]


def forest_edges(t, w):
    """Parent edges of round 0 after the 2-cycle and longer-cycle rules: {child: (parent, weight)}."""
    n = len(t)
    par = {}
    for v in range(n):
        if t[v] != NONE and not (t[t[v]] == v and v < t[v]):
            par[v] = t[v]
    # longer cycles: the minimum id on the cycle loses its parent edge
    seen = set()
    for s in range(n):
        path, v = [], s
        while v in par and v not in seen and v not in path:
            path.append(v)
            v = par[v]
        if v in path:
            cyc = path[path.index(v):]
            del par[min(cyc)]
        seen |= set(path)
    return {v: (p, int(w[v])) for v, p in par.items()}


@pytest.mark.parametrize("seed", range(12))
def test_round0_pairing_is_a_maximum_weight_matching(seed):
    rng = seeded(f"pair{seed}")
    n = int(rng.integers(4, 11))
    t = [int((v + 1 + rng.integers(0, n - 1)) % n) for v in range(n)]  # every node has a candidate
    w = rng.integers(1, 20, n).astype(np.uint32)
    pairs = np.array(t, np.uint32)
    scores = w.copy()
    sizes = np.ones(n, np.uint32)
    inb = np.zeros(n, np.uint32)
    k, groups, gs, ung, uoff, raw = O.group_nodes(n, inb, pairs, scores, sizes, 1, 1 << 20, 1 << 20)
    edges = forest_edges(t, w)
    members = {}
    for v in range(n):
        members.setdefault(int(groups[v]), []).append(v)
    got = 0
    for m in members.values():
        assert len(m) <= 2
        if len(m) == 2:
            a, b = m
            e = edges.get(a) if edges.get(a, (None,))[0] == b else edges.get(b)
            assert e is not None and e[0] in (a, b), "pair is not a forest edge"
            got += e[1]
    best = 0
    el = list(edges.items())
    for r in range(len(el) + 1):
        for sub in itertools.combinations(el, r):
            used = set()
            ok = True
            for c, (p, _) in sub:
                if c in used or p in used:
                    ok = False
                    break
                used |= {c, p}
            if ok:
                best = max(best, sum(x[1][1] for x in sub))
    assert got == best


@pytest.mark.parametrize("seed", range(4))
def test_coarse_hedges_are_group_sets(seed):
    g = make(30, 50, seed, src_prob=0.7)
    rng = seeded(f"groups{seed}")
    ng = 9
    groups = rng.integers(0, ng, g.n).astype(np.uint32)
    choff, cp, csrcc = O.coarsen_hedges(g, groups)
    for i, (s, d) in enumerate(hedges(g)):
        gs_ = {int(groups[x]) for x in s}
        gd = {int(groups[x]) for x in d}
        exp_s = sorted(gs_ - gd)
        exp_d = sorted(gd, reverse=True)
        seg = list(cp[choff[i]:choff[i + 1]])
        assert seg == exp_s + exp_d
        assert csrcc[i] == len(exp_s)


def inbound_counts(g, part, k):
    cnt = [0] * k
    for s, d in hedges(g):
        for q in {int(part[x]) for x in d}:
            cnt[q] += 1
    return cnt


def analyze(g, mode, part, k, omega, delta, fm_repeat=0, chainup=False, sizes=None, repeats=16):
    toff, tch, inb = O.touching(g)
    sizes = np.ones(g.n, np.uint32) if sizes is None else sizes
    psizes = np.bincount(part, weights=sizes, minlength=k).astype(np.uint32)
    r = O.refine_analyze(mode, omega, delta, repeats, g, toff, tch, inb, sizes, k, fm_repeat, chainup, part, psizes)
    return r, sizes, psizes


@pytest.mark.parametrize("seed,chainup", [(0, False), (1, True), (2, False), (3, True)])
def test_kway_gains_and_cascade_are_exact_connectivity_changes(seed, chainup):
    g = make(40, 70, seed, src_prob=0.0, weights="float")
    rng = seeded(f"part{seed}")
    k = 4
    part = rng.integers(0, k, g.n).astype(np.uint32)
    r, sizes, psizes = analyze(g, 1, part, k, 1 << 20, 2**31 - 1, chainup=chainup)
    base = conn_fixed(g, part)
    for v in range(g.n):
        q = int(r["moves"][v])
        if q == NONE:
            continue
        p2 = part.copy()
        p2[v] = q
        assert r["gains"][v] == base - conn_fixed(g, p2), v
    order = sorted((int(r["ranks"][v]), v) for v in range(g.n) if r["moves"][v] != NONE)
    cur = part.copy()
    for i, (rank, v) in enumerate(order):
        assert rank == i
        cur[v] = r["moves"][v]
        assert r["seq"][rank] == base - conn_fixed(g, cur), rank


def choose_b(a, b, iv, sv, seq):
    """best_move_functor (headers/data_types.cuh:54-77): True when b beats a."""
    aok, bok = iv[a] == 0, iv[b] == 0
    if aok != bok:
        return bok
    if not aok:
        return a > b
    if sv[a] != sv[b]:
        return sv[a] > sv[b]
    if seq[a] != seq[b]:
        return seq[a] < seq[b]
    return a > b


@pytest.mark.parametrize("seed,mode,slack", [(0, 1, 0), (1, 0, 0), (2, 0, 0), (3, 1, 0), (4, 0, -1), (5, 0, -2)])
def test_violation_prefixes_and_best_are_direct_recounts(seed, mode, slack):
    g = make(40, 60, seed, src_prob=0.8 if mode == 0 else 0.0)
    rng = seeded(f"viol{seed}")
    k = 5
    part = rng.integers(0, k, g.n).astype(np.uint32)
    sizes = rng.integers(1, 3, g.n).astype(np.uint32)
    psz = np.bincount(part, weights=sizes, minlength=k)
    omega = int(psz.max())  # tight: some moves overflow
    inb0 = inbound_counts(g, part, k)
    # slack < 0 starts from partitions already over the inbound limit
    delta = max(inb0) + slack if mode == 0 else 2**31 - 1
    r, _, psizes = analyze(g, mode, part, k, omega, delta, sizes=sizes, fm_repeat=7)
    ex0 = sum(max(0, int(x) - omega) for x in psizes)
    order = sorted((int(r["ranks"][v]), v) for v in range(g.n) if r["moves"][v] != NONE)
    cur = part.copy()
    touched = set()
    prev_cnt = list(inb0)
    sv_all, iv_all = {}, {}
    for rank, v in order:
        cur[v] = r["moves"][v]
        ps = np.bincount(cur, weights=sizes, minlength=k)
        sv = sum(max(0, int(x) - omega) for x in ps) - ex0
        assert r["sv"][rank] == sv, rank
        iv = 0
        if mode == 0:
            cnt = inbound_counts(g, cur, k)
            touched |= {q for q in range(k) if cnt[q] != prev_cnt[q]}
            prev_cnt = cnt
            # each partition counts as valid before its first inbound-size change
            iv = sum(1 for q in touched if cnt[q] > delta)
        assert r["iv"][rank] == iv, rank
        sv_all[rank], iv_all[rank] = sv, iv
    best = NONE
    for rank, _ in order:
        if best == NONE or choose_b(best, rank, iv_all, sv_all, r["seq"]):
            best = rank
    assert r["best"] == best
