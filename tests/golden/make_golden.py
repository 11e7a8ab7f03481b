"""Writes the committed golden fixtures under tests/golden/.

  loader_*.el / loader_*.wel     hand-written edge-list files covering the reference parser's
                                 rules (comment and blank lines, 2/3 columns, leading
                                 whitespace, files with and without vertex 0)
  loader_expected.json           what the reference's own Graph::ReadGraph (src/graph.cpp,
                                 compiled unmodified into oracle/_ref) reads from each file
  rmat14_seed1.json              the config-1 graph (RMAT scale 14, EF 16, seed 1, w in
                                 [1,64]): SHA-256 of the edge arrays, its first 64 edges, the
                                 argmax-degree source and SHA-256 of the textbook-oracle labels
  tiny_bruteforce.json           random graphs with <= 24 vertices and their Floyd-Warshall
                                 closures for BFS/SSSP/SSWP/CC

Run from the repository root after `make -C oracle ref` (needs the reference tree).
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
HERE = os.path.dirname(os.path.abspath(__file__))

import eegraph_amd as eg  # noqa: E402
import oracle_py as O  # noqa: E402

LOADER_FILES = {
    "loader_zero.el": "# comment line\n0 1\n1 2\n\n2 0\n% another comment\n3 3\n",
    "loader_onebased.el": "1 2\n2 3\n3 1\n5 4\n",
    "loader_weighted.wel": "0 1 5\n1 2\n2 3 7\n 4 0 9\n3 4 11\nx 1 2\n",
    "loader_spaces.txt": "0\t1\t3\n1  2  4\n2 0 1\r\n",
}


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def loader_fixtures():
    ref = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
    ref.ref_read_graph.restype = C.c_longlong
    ref.ref_read_graph.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_int),
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong]
    expected = {}
    for name, text in LOADER_FILES.items():
        path = os.path.join(HERE, name)
        with open(path, "w", newline="") as f:
            f.write(text)
        weighted = 1
        cap = 64
        s = np.zeros(cap, np.uint32)
        d = np.zeros(cap, np.uint32)
        w = np.zeros(cap, np.uint32)
        n = C.c_uint32(0)
        hz = C.c_int(0)
        m = ref.ref_read_graph(path.encode(), weighted, C.byref(n), C.byref(hz), s.ctypes.data,
                               d.ctypes.data, w.ctypes.data, cap)
        expected[name] = {"n": n.value, "has_zero_id": bool(hz.value), "m": int(m),
                          "src": s[:m].tolist(), "dst": d[:m].tolist(), "w": w[:m].tolist()}
    with open(os.path.join(HERE, "loader_expected.json"), "w") as f:
        json.dump(expected, f, indent=1)


def rmat_fixture():
    s, d, w = eg.gen_rmat(14, 16, seed=1, wmin=1, wmax=64)
    n = 1 << 14
    src = int(np.argmax(np.bincount(s, minlength=n)))
    labels = {name: sha(O.reference_labels(code, n, s, d, w, src))
              for name, code in (("bfs", O.BFS), ("sssp", O.SSSP), ("sswp", O.SSWP), ("cc", O.CC))}
    out = {"scale": 14, "edge_factor": 16, "seed": 1, "wmin": 1, "wmax": 64, "m": len(s),
           "sha256_src_dst_w": sha(s, d, w), "head": [[int(a), int(b), int(c)] for a, b, c in
                                                     zip(s[:64], d[:64], w[:64])],
           "argmax_source": src, "labels_sha256": labels}
    with open(os.path.join(HERE, "rmat14_seed1.json"), "w") as f:
        json.dump(out, f, indent=1)


def tiny_fixture():
    rng = np.random.default_rng(2024)
    cases = []
    for _ in range(24):
        n = int(rng.integers(1, 25))
        m = int(rng.integers(0, 3 * n + 1))
        s = rng.integers(0, n, m).astype(np.uint32)
        d = rng.integers(0, n, m).astype(np.uint32)
        w = rng.integers(1, 65, m).astype(np.uint32)
        src = int(rng.integers(0, n))
        case = {"n": n, "src": s.tolist(), "dst": d.tolist(), "w": w.tolist(), "source": src}
        for name, code in (("bfs", O.BFS), ("sssp", O.SSSP), ("sswp", O.SSWP), ("cc", O.CC)):
            case[name] = O.brute(code, n, s, d, w, src).tolist()
        cases.append(case)
    with open(os.path.join(HERE, "tiny_bruteforce.json"), "w") as f:
        json.dump(cases, f)


if __name__ == "__main__":
    loader_fixtures()
    rmat_fixture()
    tiny_fixture()
    print("fixtures written to", HERE)
