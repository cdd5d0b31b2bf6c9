"""Pins the oracle's label-propagation pieces and the device initial partitioner's building
blocks by brute force on tiny inputs (CPU only).

* A Jacobi round (modelled on the reference's legacy jacobi step, kernels/init_part.cu:89-151):
  each node's proposal is its best strictly-fitting target by the exact change of
  sum W_e (lambda_e - 1) when it moves alone, ties by mix32(q + round); positive proposals
  are taken by (target, gain desc, node) and applied while the target's size before the
  round plus everything proposed into it so far fits omega.
* Label-propagation refinement never raises lambda-1 and keeps sizes consistent.
* Best-fit decreasing with distinct sizes (where the seeded shuffle cannot matter) equals
  the textbook rule: sizes descending, the fullest bin that fits below the average load,
  else below omega (ties: lower), otherwise the lightest bin (ties: lower).
* The device initial partitioner's restatement places every node, reports its part
  weights, and is a pure function of its input.
"""
import zlib

import numpy as np
import pytest

import gen
import oracle as O

M32 = 0xFFFFFFFF


def seeded(name):
    return np.random.default_rng(zlib.crc32(name.encode()))


def fixed(w):
    v = float(np.float32(w) * np.float32(262144.0))
    if not v > 0:
        return 0
    return M32 if v >= 4294967296.0 else int(v)


def km1(hs, W, part):
    return sum(W[i] * (len({int(part[x]) for x in pins}) - 1) for i, pins in enumerate(hs) if pins)


def graph(n, e, seed, max_size=5):
    rng = seeded(f"lp{seed}")
    g = O.HG(*gen.random_hypergraph(n, e, rng, max_size=max_size, weights="float"))
    g = O.dedup(g)
    hs = [[int(x) for x in g.pins[g.off[i]:g.off[i + 1]]] for i in range(g.e)]
    return g, hs, [fixed(x) for x in g.w], rng


def brute_jacobi(hs, W, sizes, part, psizes, k, omega, rnd):
    part = [int(x) for x in part]
    base = km1(hs, W, part)
    props = []
    for v in range(len(part)):
        p, s = part[v], int(sizes[v])
        best = None
        for q in range(k):
            if q == p or int(psizes[q]) + s > omega:
                continue
            trial = list(part)
            trial[v] = q
            gain = base - km1(hs, W, trial)
            key = (gain, O.lib().orc_hash_uint32((q + rnd) & M32))
            if best is None or key > best[0]:
                best = (key, q)
        if best is not None and best[0][0] > 0:
            props.append((best[1], -best[0][0], v))
    props.sort()
    before = [int(x) for x in psizes]
    ps = list(before)
    entering = [0] * k
    applied = 0
    for q, _, v in props:
        s = int(sizes[v])
        entering[q] += s
        if before[q] + entering[q] > omega:
            continue
        ps[part[v]] -= s
        ps[q] += s
        part[v] = q
        applied += 1
    return part, ps, applied


@pytest.mark.parametrize("seed,n,e,k", [(0, 12, 20, 2), (1, 20, 30, 3), (2, 16, 40, 5), (3, 24, 24, 4)])
@pytest.mark.parametrize("rnd", [0, 3])
def test_jacobi_round_is_the_brute_force_rule(seed, n, e, k, rnd):
    g, hs, W, rng = graph(n, e, seed)
    sizes = rng.integers(1, 4, n).astype(np.uint32)
    part = rng.integers(0, k, n).astype(np.uint32)
    psizes = np.bincount(part, weights=sizes, minlength=k).astype(np.uint32)
    omega = int(psizes.max()) + 2  # some moves fit, some do not
    ep, eps_, ea = brute_jacobi(hs, W, sizes, part, psizes, k, omega, rnd)
    gp, gps, ga = O.jacobi_round(g, sizes, k, omega, rnd, part, psizes)
    assert ga == ea
    assert gp.tolist() == ep
    assert gps.tolist() == eps_


@pytest.mark.parametrize("seed", range(4))
def test_lp_refine_never_raises_lambda(seed):
    g, hs, W, rng = graph(60, 90, 10 + seed, max_size=6)
    k = 4
    sizes = np.ones(g.n, np.uint32)
    part = rng.integers(0, k, g.n).astype(np.uint32)
    psizes = np.bincount(part, minlength=k).astype(np.uint32)
    omega = int(np.ceil(1.1 * g.n / k))
    gp, gps = O.lp_refine(g, sizes, k, omega, 8, part, psizes)
    assert km1(hs, W, gp) <= km1(hs, W, part)
    assert gps.tolist() == np.bincount(gp, minlength=k).tolist()
    assert int(gps.max()) <= max(omega, int(psizes.max()))


def textbook_bfd(sizes, k, omega):
    """Best fit decreasing against the average load first, then against omega."""
    order = sorted(range(len(sizes)), key=lambda v: -int(sizes[v]))
    loads = [0] * k
    part = [0] * len(sizes)
    avg = min(omega, -(-int(np.sum(sizes)) // k))
    for v in order:
        s = int(sizes[v])
        fit = [b for b in range(k) if loads[b] + s <= avg] or [b for b in range(k) if loads[b] + s <= omega]
        b = min(fit, key=lambda b: (-loads[b], b)) if fit else min(range(k), key=lambda b: (loads[b], b))
        part[v] = b
        loads[b] += s
    return part, loads


@pytest.mark.parametrize("seed", range(6))
def test_bfd_with_distinct_sizes_is_textbook(seed):
    rng = seeded(f"bfd{seed}")
    n = int(rng.integers(5, 40))
    sizes = rng.permutation(np.arange(1, n + 1) * 3).astype(np.uint32)  # distinct
    k = int(rng.integers(2, 6))
    omega = int(sizes.sum() // k + rng.integers(0, 20))
    ep, el = textbook_bfd(sizes, k, omega)
    for s in (0, 1, 7):  # the shuffle seed cannot matter with distinct sizes
        gp, gl = O.bfd_start(sizes, k, omega, s)
        assert gp.tolist() == ep and gl.tolist() == el


def test_bfd_shuffle_orders_equal_sizes_only():
    sizes = np.array([2] * 10 + [5, 9], np.uint32)
    a, la = O.bfd_start(sizes, 3, 12, 0)
    b, lb = O.bfd_start(sizes, 3, 12, 5)
    # the two largest items go first whatever the shuffle: 9 into bin 0, 5 into bin 1 (the
    # average load is ceil(34 / 3) = 12 and 9 + 5 > 12)
    assert a[11] == b[11] == 0 and a[10] == b[10] == 1
    for p, l in ((a, la), (b, lb)):
        assert np.bincount(p, weights=sizes, minlength=3).tolist() == l.tolist()


@pytest.mark.parametrize("k", [2, 4, 16])
def test_device_initial_partitioner_invariants(k):
    rng = seeded(f"dip{k}")
    g = O.dedup(O.HG(*gen.random_hypergraph(3000, 3200, rng, max_size=8, weights="float")))
    sizes = rng.integers(1, 3, g.n).astype(np.uint32)
    omega = int(np.ceil(1.05 * sizes.sum() / k))
    part, ps = O.device_initial_partition(g, sizes, k, omega)
    assert part.max() < k
    assert ps.tolist() == np.bincount(part, weights=sizes, minlength=k).astype(np.uint32).tolist()
    part2, ps2 = O.device_initial_partition(g, sizes, k, omega)
    assert np.array_equal(part, part2) and np.array_equal(ps, ps2)
    # small node weights and 5% slack: the packing fits, and the rounds after it keep sizes
    assert int(ps.max()) <= omega


def test_bfd_uses_every_bin_when_slack_is_small():
    """With 3% slack, packing against omega alone fills 63 of 64 bins; against the average
    load every bin gets nodes."""
    sizes = np.full(6400, 5, np.uint32)
    k = 64
    omega = int(np.ceil(1.03 * sizes.sum() / k))
    part, loads = O.bfd_start(sizes, k, omega, 0)
    assert (loads > 0).all() and int(loads.max()) <= omega
