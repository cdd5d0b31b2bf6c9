"""hgp_connectivity_noalloc: lambda-1 with no device allocation, through the C ABI.  Each call
is checked against hgp_connectivity (which builds partition lists in allocated memory) on the
same inputs, bit for bit, on two calls in a row, the second with a smaller input; the
context's allocator must not move during the call."""
import ctypes as C

import numpy as np
import pytest

import gen
import gpu_stages as G
import oracle as O

pytestmark = pytest.mark.gpu


def mixed_sizes(n, e, rng):
    """Hedges of 2..8 pins, with every 50th hedge between 65 and 400 pins (the bitmap path)."""
    sizes = rng.integers(2, 9, size=e)
    sizes[::50] = rng.integers(65, 401, size=len(sizes[::50]))
    off = np.zeros(e + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    pins = np.concatenate([rng.choice(n, size=int(s), replace=False) for s in sizes]).astype(np.uint32)
    w = rng.integers(1, 5, size=e).astype(np.float32)
    return O.HG(n, off, pins, np.zeros(e, np.uint32), w)


def memory(dev):
    a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
    dev.check(dev.lib.hgp_ctx_memory(dev.ctx, C.byref(a), C.byref(b), C.byref(c)), "hgp_ctx_memory")
    return a.value, b.value, c.value


def both(dev, hg, part, k, d_out):
    L = G.Level(dev, hg)
    d_part = dev.upload(part)
    ref = C.c_int64()
    dev.check(dev.lib.hgp_connectivity(dev.ctx, G.V(L.pins), G.V(L.off), G.V(L.w), G.V(d_part), hg.e, C.byref(ref)),
              "hgp_connectivity")
    before = memory(dev)
    dev.check(dev.lib.hgp_connectivity_noalloc(dev.ctx, G.V(L.pins), G.V(L.off), G.V(L.w), G.V(d_part), hg.e, k,
                                               G.V(d_out)), "hgp_connectivity_noalloc")
    dev.check(dev.lib.hgp_synchronize(dev.ctx), "sync")
    assert memory(dev) == before, "the no-allocation entry point moved the allocator"
    got = int(dev.download(d_out, 1, np.int64)[0])
    for p in (L.pins, L.off, L.srcc, L.w, d_part):
        dev.free(p)
    return got, ref.value


@pytest.mark.parametrize("k", [2, 64, 5000, 65536])
def test_noalloc_equals_allocating_entry_point_twice(dev, k, rng):
    d_out = dev.alloc(8)
    big = mixed_sizes(60000, 150000, rng)
    small = mixed_sizes(3000, 4000, rng)
    for hg in (big, small):  # the second call is smaller, in the same scratch
        part = rng.integers(0, k, size=hg.n).astype(np.uint32)
        got, ref = both(dev, hg, part, k, d_out)
        print(f"case: noalloc k={k} n={hg.n} e={hg.e} km1={ref}", flush=True)
        assert got == ref
    # against the oracle on the small input as well
    part = rng.integers(0, k, size=small.n).astype(np.uint32)
    got, _ = both(dev, small, part, k, d_out)
    assert got == O.connectivity_fixed(small, part)
    dev.free(d_out)


def test_noalloc_edge_cases(dev, rng):
    d_out = dev.alloc(8)
    # no hedges at all
    empty = O.HG(10, np.zeros(1, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.float32))
    d_part = dev.upload(np.zeros(10, np.uint32))
    d_off = dev.upload(empty.off)
    dev.check(dev.lib.hgp_connectivity_noalloc(dev.ctx, G.V(d_off), G.V(d_off), G.V(d_off), G.V(d_part), 0, 4,
                                               G.V(d_out)), "empty")
    assert int(dev.download(d_out, 1, np.int64)[0]) == 0
    # one partition: lambda-1 is zero everywhere
    hg = O.HG(*gen.random_hypergraph(500, 800, rng, max_size=90, weights="float"))
    got, ref = both(dev, hg, np.zeros(hg.n, np.uint32), 1, d_out)
    assert got == ref == 0
    # a partition id beyond num_partitions (short and long hedges) is reported as -1
    hg = mixed_sizes(2000, 3000, rng)
    part = rng.integers(0, 8, size=hg.n).astype(np.uint32)
    part[hg.pins[int(hg.off[0])]] = 9  # a long hedge holds this pin
    got, _ = both(dev, hg, part, 8, d_out)
    assert got == -1
    # and the next call is clean again
    part[part >= 8] = 0
    got, ref = both(dev, hg, part, 8, d_out)
    assert got == ref
    # more partitions than the bitmap holds: refused before any launch
    st = dev.lib.hgp_connectivity_noalloc(dev.ctx, G.V(d_off), G.V(d_off), G.V(d_off), G.V(d_part), 0, 65537,
                                          G.V(d_out))
    assert st == 4  # HGP_E_CAPACITY
    for p in (d_part, d_off, d_out):
        dev.free(p)
