"""CPU checks of the C ABI boundary: the library loads, exports every symbol declared in
include/hgp.h, and its host-side model (loaders, dedup, metrics, validity) agrees with the
oracle.  No device compute is called here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import axoncuda_amd as H
import gen
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hgp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hgp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = H.load_library()
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    bound = {name for name, _, _ in H.SIGNATURES}
    assert set(syms) <= bound, set(syms) - bound


def test_no_device_is_reported(monkeypatch):
    lib = H.load_library()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    ctx = C.c_void_p()
    assert lib.hgp_ctx_create(0, 0, C.byref(ctx)) == 7  # HGP_E_NODEVICE


@pytest.mark.parametrize("fmt", ["hgr", "snn", "axh"])
def test_product_loader_matches_oracle(tmp_path, fmt, rng):
    if fmt == "hgr":
        n, off, pins, srcc, w = gen.random_hypergraph(60, 90, rng, max_size=7, weights="int")
        p = str(tmp_path / "x.hgr")
        gen.write_hgr(p, n, off, pins, w, singletons=2)
    elif fmt == "snn":
        n, off, pins, srcc, w = gen.random_hypergraph(60, 90, rng, max_size=7, src_prob=1, max_src=1, dup_prob=0.5,
                                                      weights="float")
        p = str(tmp_path / "x.snn")
        gen.write_snn(p, n, off, pins, w)
    else:
        n, off, pins, srcc, w = gen.random_hypergraph(60, 90, rng, max_size=7, src_prob=1, max_src=3, dup_prob=0.6,
                                                      weights="float")
        p = str(tmp_path / "x.axh")
        gen.write_axh(p, n, off, pins, srcc, w)
    for dedup in (False, True):
        got = H.Hypergraph.load(p, dedup=dedup).arrays()
        exp = O.load(p, dedup)
        assert got[0] == exp.n
        for a, b in zip(got[1:], (exp.off, exp.pins, exp.srcc, exp.w)):
            assert np.array_equal(a, b)


def write_big_hgr(path, rng, n=60000, e=450000, fmt=11, bad_at=None):
    """A .hgr above the threaded parser's 4 MB threshold: weights and vertex weights (fmt 11),
    comment and blank lines, CRLF endings on some lines, repeated pins and singleton hedges.
    With bad_at, hyperedge lines at those indexes name node n + 1."""
    sizes = np.minimum(rng.geometric(0.3, e) + 1, 40)
    sizes[rng.random(e) < 0.05] = 1  # singletons, skipped by the loader
    lines = ["% threaded-parser test", f"{e} {n} {fmt}"]
    bad = set(bad_at or [])
    for i in range(e):
        ps = rng.integers(1, n + 1, sizes[i])
        if i in bad:
            ps[-1] = n + 1
        body = " ".join(map(str, ps))
        if fmt in (1, 11):
            body = f"{int(rng.integers(1, 9))} {body}"
        lines.append(body + ("\r" if i % 7 == 3 else ""))
        if i % 1000 == 17:
            lines.append("   % comment inside the hyperedges")
        if i % 5003 == 11:
            lines.append("")
    if fmt in (10, 11):
        lines += [str(int(x)) for x in rng.integers(1, 5, n)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@pytest.mark.parametrize("fmt", [0, 11])
def test_threaded_hgr_parser_matches_oracle(tmp_path, fmt, rng):
    p = str(tmp_path / "big.hgr")
    write_big_hgr(p, rng, fmt=fmt)
    assert os.path.getsize(p) > 4 << 20
    for dedup in (False, True):
        got = H.Hypergraph.load(p, dedup=dedup).arrays()
        exp = O.load(p, dedup)
        assert got[0] == exp.n
        for a, b in zip(got[1:], (exp.off, exp.pins, exp.srcc, exp.w)):
            assert np.array_equal(a, b)


def test_threaded_hgr_parser_reports_errors(tmp_path, rng):
    p = str(tmp_path / "bad.hgr")
    write_big_hgr(p, rng, bad_at=[400000, 9])
    hg = C.c_void_p()
    assert H.load_library().hgp_hgraph_load(p.encode(), 0, C.byref(hg)) != 0
    short = str(tmp_path / "short.hgr")
    write_big_hgr(short, rng, fmt=0)
    text = open(short).read().split("\n")
    text[1] = text[1].replace(" 0", "").split()[0]
    head = f"{int(text[1]) + 10} 60000"  # ten more hyperedges than the file holds
    open(short, "w").write("\n".join([text[0], head] + text[2:]))
    assert H.load_library().hgp_hgraph_load(short.encode(), 0, C.byref(hg)) != 0


def test_threaded_metrics_and_validity_match_oracle(rng):
    """Above 2^20 pins the graders count per hedge on several threads and sum in hedge order."""
    n, off, pins, srcc, w = gen.random_hypergraph(150000, 300000, rng, max_size=12, src_prob=0.6, max_src=3,
                                                  weights="float")
    hg = H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)
    assert hg.sizes()[2] > 1 << 20
    orc = O.HG(*hg.arrays())
    for k in (2, 64, 5000):
        part = O.relabel(rng.integers(0, k, size=n).astype(np.uint32))[1]
        assert hg.metrics(part) == O.metrics(orc, part)  # the same sums in the same order
        for cap in ((n, 10**9, k), (n // k + 40, 400, k), (5, 10**9, k)):
            assert hg.validity(part, *cap) == O.validity(orc, part, *cap)


def test_product_metrics_and_validity_match_oracle(rng):
    n, off, pins, srcc, w = gen.random_hypergraph(200, 300, rng, max_size=8, src_prob=0.7, max_src=2, weights="int")
    hg = H.Hypergraph.from_csr(n, off, pins, srcc, w, dedup=True)
    orc = O.HG(*hg.arrays())
    for k in (2, 7, 30):
        part = O.relabel(rng.integers(0, k, size=n).astype(np.uint32))[1]
        assert np.allclose(hg.metrics(part), O.metrics(orc, part))
        for cap in ((n, 10**9, k), (n // k + 2, 60, k), (5, 10**9, k)):
            assert hg.validity(part, *cap) == O.validity(orc, part, *cap)


def test_generators_are_deterministic():
    a = H.Hypergraph.generate("netlist", n=5000, e=5200, seed=7).arrays()
    b = H.Hypergraph.generate("netlist", n=5000, e=5200, seed=7).arrays()
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
    s = H.Hypergraph.generate("snn", layers=3, per_layer=400, grid=20, fanout=16, sigma=3.0, seed=2).arrays()
    n, off, pins, srcc, w = s
    assert n == 1200 and len(srcc) == 800 and np.all(srcc == 1)
    assert np.all(np.diff(off) == 17)


def test_cli_help_and_errors():
    exe = os.path.join(H.BIN_DIR, "hgraph_gpu")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and "Usage:" in r.stdout
    r = subprocess.run([exe, "-m", "1", "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "three constraints" in r.stderr
    r = subprocess.run([exe, "-cnc", "17"], capture_output=True, text=True)
    assert r.returncode == 1
    r = subprocess.run([exe, "-bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown option" in r.stderr
    r = subprocess.run([exe, "-ip", "gpu"], capture_output=True, text=True)
    assert r.returncode == 1 and "'host' or 'device'" in r.stderr


def test_tuning_file_round_trip_and_validation(tmp_path):
    """hgp_tuning as a text file: defaults, save/load round trip, malformed files refused."""
    lib = H.load_library()
    t = H.Tuning()
    lib.hgp_tuning_defaults(C.byref(t))
    assert t.heavy_table_slots == 65536 and t.heavy_table_fill16 == 12 and t.cascade_block == 256
    assert t.scan_lookback_tiles == 64 and t.tuned == 0 and not any(t.sort_rounds)
    t.sort_rounds[20] = 4
    t.sort_rounds[26] = 16
    t.cascade_block = 512
    t.heavy_table_slots = 1 << 17
    t.heavy_table_fill16 = 10
    t.tuned = 1
    p = str(tmp_path / "tune.txt")
    assert lib.hgp_tuning_save(C.byref(t), p.encode()) == 0
    u = H.Tuning()
    assert lib.hgp_tuning_load(p.encode(), C.byref(u)) == 0
    assert u.as_dict() == t.as_dict()
    assert lib.hgp_tuning_load(str(tmp_path / "missing.txt").encode(), C.byref(u)) == 5  # HGP_E_IO
    for bad in ("cascade_block 96\n", "heavy_table_slots 5000\n", "heavy_table_fill16 15\n", "sort_rounds 3 17\n",
                "sort_rounds 40 1\n", "unknown_key 1\n", "cascade_block\n"):
        q = tmp_path / "bad.txt"
        q.write_text(bad)
        assert lib.hgp_tuning_load(str(q).encode(), C.byref(u)) == 1, bad  # HGP_E_INVALID
    bad = H.Tuning()
    lib.hgp_tuning_defaults(C.byref(bad))
    bad.cascade_block = 100
    assert lib.hgp_tuning_set_process_default(C.byref(bad)) == 1
    assert lib.hgp_tuning_set_process_default(None) == 0
