"""Host-side pieces of the C-ABI that run without a GPU: exported symbols, generators,
edge-list files, status codes and argument validation."""
import json
import os
import re

import numpy as np
import pytest

import eegraph_amd as eg
from eegraph_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_functions():
    with open(os.path.join(ROOT, "include", "eeg_capi.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eeg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L.lib, n)]
    assert not missing, missing
    assert sorted(L.EXPORTS) == names


def test_status_strings():
    assert L.lib.eeg_strerror(0) == b"ok"
    for s in range(1, 11):
        assert L.lib.eeg_strerror(s) not in (b"", b"unknown status")


def test_rmat_deterministic_and_thread_independent():
    a = eg.gen_rmat(16, 8, seed=7, threads=1)
    b = eg.gen_rmat(16, 8, seed=7, threads=8)
    c = eg.gen_rmat(16, 8, seed=8, threads=8)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])


def test_rmat_quadrant_statistics():
    """Top-level bit of (src, dst) follows the a/b/c/d quadrant probabilities."""
    s, d, w = eg.gen_rmat(16, 16, seed=3)
    top = 1 << 15
    hs, hd = s >= top, d >= top
    m = len(s)
    frac = [np.mean(~hs & ~hd), np.mean(~hs & hd), np.mean(hs & ~hd), np.mean(hs & hd)]
    for f, p in zip(frac, (0.57, 0.19, 0.19, 0.05)):
        assert abs(f - p) < 5 * np.sqrt(p * (1 - p) / m)
    assert w.min() >= 1 and w.max() <= 64
    counts = np.bincount(w, minlength=65)[1:]
    assert counts.min() > 0.9 * m / 64 and counts.max() < 1.1 * m / 64


def test_grid_shape_and_shortcuts():
    s, d, w = eg.gen_grid(7, 5, 0.1, seed=3, wmin=1, wmax=1000)
    undirected = 5 * 6 + 4 * 7
    ns = round(0.1 * undirected)
    assert len(s) == 2 * (undirected + ns)
    # every pair is present in both directions
    assert np.array_equal(s[0::2], d[1::2]) and np.array_equal(d[0::2], s[1::2])
    # lattice edges connect neighbours
    lat = slice(0, 2 * undirected)
    diff = np.abs(s[lat].astype(np.int64) - d[lat].astype(np.int64))
    assert set(np.unique(diff)) <= {1, 7}
    assert w.min() >= 1 and w.max() <= 1000
    # the 4096^2 config: 335,462 shortcut pairs (1% of 33,546,240 undirected lattice edges)
    import ctypes as C

    m = C.c_uint64()
    assert L.lib.eeg_gen_grid_size(4096, 4096, 0.01, C.byref(m)) == 0
    assert m.value == 2 * (33546240 + 335462)


def test_loader_matches_reference_parser():
    """Expected values come from the reference's own Graph::ReadGraph (tests/golden)."""
    with open(os.path.join(GOLD, "loader_expected.json")) as f:
        exp = json.load(f)
    for name, e in exp.items():
        n, s, d, w, hz = eg.load_edgelist(os.path.join(GOLD, name), weighted=True)
        assert n == e["n"] and hz == e["has_zero_id"] and len(s) == e["m"], name
        shift = 0 if e["has_zero_id"] else 1  # 1-based files shift down (virtual_graph.cpp:5-12)
        assert s.tolist() == [x - shift for x in e["src"]], name
        assert d.tolist() == [x - shift for x in e["dst"]], name
        assert w.tolist() == e["w"], name


def test_binary_csr_round_trip(tmp_path):
    s, d, w = eg.gen_rmat(10, 4, seed=2)
    n = 1 << 10
    for ext, weighted in (("bwcsr", True), ("bcsr", False)):
        p = tmp_path / f"g.{ext}"
        eg.save_edgelist(p, n, s, d, w if weighted else None)
        n2, s2, d2, w2, _ = eg.load_edgelist(p)
        order = np.argsort(s, kind="stable")  # the file is in CSR order, stable by input order
        assert n2 == n
        assert np.array_equal(s2, s[order]) and np.array_equal(d2, d[order])
        if weighted:
            assert np.array_equal(w2, w[order])
        else:
            assert w2 is None


def test_text_round_trip(tmp_path):
    s, d, w = eg.gen_rmat(8, 4, seed=4)
    s[0] = 0  # make sure id 0 occurs so the file reads back 0-based
    p = tmp_path / "g.wel"
    eg.save_edgelist(p, 256, s, d, w)
    n2, s2, d2, w2, hz = eg.load_edgelist(p)
    assert hz and np.array_equal(s2, s) and np.array_equal(d2, d) and np.array_equal(w2, w)


def test_loader_errors(tmp_path):
    with pytest.raises(eg.EEGError) as e:
        eg.load_edgelist(tmp_path / "missing.el")
    assert e.value.status == L.ERR_IO
    bad = tmp_path / "g.xyz"
    bad.write_text("0 1\n")
    with pytest.raises(eg.EEGError) as e:
        eg.load_edgelist(bad)
    assert e.value.status == L.ERR_IO
    trunc = tmp_path / "g.bcsr"
    trunc.write_bytes(b"\x05\x00\x00\x00")
    with pytest.raises(eg.EEGError) as e:
        eg.load_edgelist(trunc)
    assert e.value.status == L.ERR_IO


def test_generator_argument_errors():
    with pytest.raises(eg.EEGError) as e:
        eg.gen_rmat(14, 16, a=0.6, b=0.3, c=0.3)
    assert e.value.status == L.ERR_INVALID_ARG
    with pytest.raises(eg.EEGError):
        eg.gen_rmat(14, 16, wmin=10, wmax=5)


def test_no_gpu_is_reported_not_fatal():
    """Without a GPU the device entry points return EEG_ERR_NO_DEVICE instead of exiting."""
    if eg.device_count() > 0:
        pytest.skip("a GPU is present")
    s = np.array([0], np.uint32)
    with pytest.raises(eg.EEGError) as e:
        eg.Graph.from_edges(s, s, None, n=1)
    assert e.value.status == L.ERR_NO_DEVICE


def test_energy_counter_optional():
    """amd-smi may be absent or not permitted; the call then reports unsupported."""
    v = eg.energy_read(0)
    assert v is None or v >= 0


def test_entry_points_validate_before_touching_a_device():
    """Argument errors are reported without a GPU (no device call is made first)."""
    assert L.lib.eeg_run_packed(None, 0, L.BFS, None, None, None, None) == L.ERR_INVALID_ARG
    assert L.lib.eeg_autotune(None, L.BFS, None, 0, None, None) == L.ERR_INVALID_ARG
    assert L.lib.eeg_run_batch(None, L.BFS, None, 0, None, None, None) == L.ERR_INVALID_ARG


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 (or a WORLD_SIZE above 1) exits with status 2 and a
    message, before anything touches a device."""
    import subprocess
    import sys
    bench = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")
    p = subprocess.run([sys.executable, bench, "--gpus", "2"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
    env = dict(os.environ, WORLD_SIZE="4")
    p = subprocess.run([sys.executable, bench], capture_output=True, text=True, timeout=60, env=env)
    assert p.returncode == 2 and "WORLD_SIZE 4" in p.stderr
