"""Pins the oracle's refinement move rules by brute force on tiny inputs (CPU only).

Each expectation is computed from the definition with Python sets and direct simulation:

* target selection (kernels/refinement.cu:142-152,184-200; sources/refinement.cu:86): every
  q != p is scored exactly (k-way: the change of sum W_e (lambda_e - 1) when v moves; INCC:
  sum floor(W_e / c_e(p)^2) over v's hedges minus the weights of the hedges with no pin in
  q), filtered by size(q) + s - floor(s / discount) <= omega with the discount of the
  repeat's third, and the best taken by (gain, mix32(q + repeat));
* the INCC cascade (kernels/refinement.cu:216-322): the same score of each move recomputed
  after applying every better-ranked move, then prefix-summed;
* chaining (kernels/chaining.cu:31-230, sources/chaining.cu:19-254): links join a move into
  partition q to a move out of q among the first 256 of q's moves (sorted by gain), with
  size buckets |floor(log2 a) - floor(log2 b)| <= 1, at most one predecessor and one
  successor per move; ranks list whole components, by summed gain descending then head,
  and follow each chain from its head.
"""
import zlib

import numpy as np
import pytest

import gen
import oracle as O

NONE = 0xFFFFFFFF
M32 = 0xFFFFFFFF
R = 16  # refine repeats: discount 1 on repeats 0-4, 2 on 5-9, none from 10


def seeded(name):
    return np.random.default_rng(zlib.crc32(name.encode()))


def make(n, e, seed, src_prob, max_size=6, weights="float"):
    rng = seeded(f"mv{seed}")
    g = O.HG(*gen.random_hypergraph(n, e, rng, max_size=max_size, src_prob=src_prob, max_src=2, dup_prob=0.3,
                                    weights=weights))
    return O.dedup(g)


def hedges(g):
    out = []
    for i in range(g.e):
        seg = [int(x) for x in g.pins[g.off[i]:g.off[i + 1]]]
        out.append(seg)
    return out


def fixed(w):
    v = float(np.float32(w) * np.float32(262144.0))
    if not v > 0:
        return 0
    return M32 if v >= 4294967296.0 else int(v)


def km1(hs, W, part):
    return sum(W[i] * (len({int(part[x]) for x in pins}) - 1) for i, pins in enumerate(hs) if pins)


def incc_score(hs, W, part, v, q):
    """INCC gain of moving v from its partition to q, from the reference's definition."""
    p = int(part[v])
    saving = loss = 0
    for i, pins in enumerate(hs):
        if v not in pins:
            continue
        cp = sum(1 for x in pins if part[x] == p)
        cq = sum(1 for x in pins if part[x] == q)
        saving += W[i] // (cp * cp)
        if cq == 0:
            loss += W[i]
    return saving - loss


def discount_of(rep):
    return 1 if rep < R // 3 else (2 if rep < 2 * R // 3 else None)


def analyze(g, mode, part, k, omega, delta, sizes, fm_repeat, chainup=False):
    toff, tch, inb = O.touching(g)
    psizes = np.bincount(part, weights=sizes, minlength=k).astype(np.uint32)
    r = O.refine_analyze(mode, omega, delta, R, g, toff, tch, inb, sizes, k, fm_repeat, chainup, part, psizes)
    return r, psizes


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("fm_repeat", [0, 6, 13])
@pytest.mark.parametrize("seed", range(3))
def test_gain_target_selection_brute_force(mode, fm_repeat, seed):
    g = make(36, 60, seed, src_prob=0.0 if mode == 1 else 0.7)
    hs = hedges(g)
    W = [fixed(x) for x in g.w]
    rng = seeded(f"sel{mode}{fm_repeat}{seed}")
    k = 6
    part = rng.integers(0, k - 1, g.n).astype(np.uint32)  # partition k-1 starts empty: non-adjacent targets
    sizes = rng.integers(1, 5, g.n).astype(np.uint32)
    psizes = np.bincount(part, weights=sizes, minlength=k).astype(np.int64)
    omega = int(np.median(psizes[:k - 1])) + 2  # binds for some targets, not for others
    r, _ = analyze(g, mode, part, k, omega, 2**31 - 1 if mode == 1 else 1 << 20, sizes, fm_repeat)
    disc = discount_of(fm_repeat)
    base = km1(hs, W, part)
    n_none = n_moves = n_ties = n_filtered = 0
    for v in range(g.n):
        p, s = int(part[v]), int(sizes[v])
        best = None
        gains = []
        for q in range(k):
            if q == p:
                continue
            extra = s - (s // disc if disc else 0)
            if (int(psizes[q]) + extra) & M32 > omega:
                n_filtered += 1
                continue
            if mode == 1:
                p2 = part.copy()
                p2[v] = q
                gain = base - km1(hs, W, p2)
            else:
                gain = incc_score(hs, W, part, v, q)
            gains.append(gain)
            key = (gain, O.lib().orc_hash_uint32((q + fm_repeat) & M32))
            if best is None or key > best[0]:
                best = (key, q)
        n_ties += len(gains) != len(set(gains))
        if best is None:
            assert r["moves"][v] == NONE and r["gains"][v] == 0, v
            n_none += 1
        else:
            assert r["moves"][v] == best[1], (v, r["moves"][v], best)
            assert r["gains"][v] == best[0][0], v
            n_moves += 1
    assert n_moves > 0 and n_ties > 0 and n_filtered > 0  # the cases exercise the tie rule and the filter


@pytest.mark.parametrize("seed,chainup,fm_repeat", [(0, False, 0), (1, True, 6), (2, False, 13), (3, True, 0)])
def test_incc_cascade_is_direct_application_in_rank_order(seed, chainup, fm_repeat):
    g = make(40, 70, 10 + seed, src_prob=0.8)
    hs = hedges(g)
    W = [fixed(x) for x in g.w]
    rng = seeded(f"casc{seed}")
    k = 5
    part = rng.integers(0, k, g.n).astype(np.uint32)
    sizes = np.ones(g.n, np.uint32)
    r, _ = analyze(g, 0, part, k, 1 << 20, 1 << 20, sizes, fm_repeat, chainup=chainup)
    order = sorted((int(r["ranks"][v]), v) for v in range(g.n) if r["moves"][v] != NONE)
    assert len(order) > 10
    cur = part.copy()
    total = 0
    for i, (rank, v) in enumerate(order):
        assert rank == i
        total += incc_score(hs, W, cur, v, int(r["moves"][v]))  # v itself has not moved yet
        assert r["seq"][rank] == total, rank
        cur[v] = r["moves"][v]


def bucket(s):
    return 0 if s == 0 else int(s).bit_length() - 1


@pytest.mark.parametrize("seed,n,k", [(0, 60, 4), (1, 200, 3), (2, 700, 2), (3, 400, 8), (4, 30, 2)])
def test_chaining_rule_properties(seed, n, k):
    rng = seeded(f"chain{seed}")
    part = rng.integers(0, k, n).astype(np.uint32)
    moves = np.array([NONE if rng.random() < 0.15 else int((p + 1 + rng.integers(0, k - 1)) % k) for p in part],
                     np.uint32)
    gains = rng.integers(-5, 6, n).astype(np.int64)  # many ties
    gains[moves == NONE] = 0
    sizes = rng.choice([1, 2, 3, 5, 8, 16, 40], n).astype(np.uint32)
    ranks, prev, comp, pos = O.chain_links(part, moves, gains, sizes)
    valid = moves != NONE
    V = int(valid.sum())
    assert sorted(ranks.tolist()) == list(range(n))
    # the sorted space: by source partition, valid moves first, gain descending, id ascending
    sidx_order = sorted(range(n), key=lambda v: (int(part[v]), not valid[v], -int(gains[v]) if valid[v] else 0, v))
    sidx = {v: i for i, v in enumerate(sidx_order)}
    within = {}
    for v in sidx_order:
        lst = within.setdefault(int(part[v]), [])
        lst.append(v)
    wpos = {v: i for lst in within.values() for i, v in enumerate(lst)}
    succ = {}
    n_links = 0
    for s in range(n):
        i = int(prev[s])
        if i == NONE:
            continue
        n_links += 1
        assert i != s and valid[i] and valid[s]
        assert moves[i] == part[s], "a link joins a move into q to a move out of q"
        assert wpos[s] < 256, "successor outside the 256-entry window"
        assert abs(bucket(sizes[i]) - bucket(sizes[s])) <= 1, "size buckets differ by more than one"
        assert i not in succ, "two successors"
        succ[i] = s
    assert n_links > 0
    for v in range(n):
        if not valid[v]:
            assert prev[v] == NONE and v not in succ
            assert ranks[v] >= V, "invalid moves rank last"
    # components and positions from the links
    comps = {}
    for v in range(n):
        chain = [v]
        cur = v
        while prev[cur] != NONE and prev[cur] not in chain and len(chain) <= 256:
            cur = int(prev[cur])
            chain.append(cur)
        if prev[cur] == NONE:
            head, p_ = cur, len(chain) - 1
        else:  # on a cycle: the member first in the sorted space heads it, every position 0
            cyc = chain[chain.index(int(prev[cur])):] if int(prev[cur]) in chain else chain
            head, p_ = min(cyc, key=lambda x: sidx[x]), 0
        assert comp[v] == head and pos[v] == p_, v
        comps.setdefault(head, []).append(v)
    key = {}
    for h, mem in comps.items():
        bad = any(not valid[x] for x in mem)
        key[h] = (bad, -sum(int(gains[x]) for x in mem if valid[x]) if not bad else 0, sidx[h])
    expect = []
    for h in sorted(comps, key=lambda h: key[h]):
        expect += sorted(comps[h], key=lambda x: (int(pos[x]), sidx[x]))
    assert [int(ranks[v]) for v in expect] == list(range(n))
