"""Serving loop with continuous batching: every answer must be the same alone (eeg_run) and in
the mix, bit for bit for BFS/SSSP/SSWP/CC; PPR and PR within the stated L1 tolerance of the
single run and of the FP64 oracle."""
import threading
import time

import numpy as np
import pytest

import eegraph_amd as eg
import eegraph_amd._lib as L
import oracle_py as O

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6  # L1 between ranks (FP32 compensated) and the FP64 oracle / single run


@pytest.fixture(scope="module")
def rmat12():
    s, d, w = eg.gen_rmat(12, 16, seed=6, wmin=1, wmax=64)
    n = 1 << 12
    g = eg.Graph.from_edges(s, d, w, n=n)
    row, col, _ = O.csr(n, s, d, w)
    yield n, s, d, w, g, row, col
    g.close()


def _single(g, alg, src):
    return g.run(alg, src, variant="sync_push_dd", K=8)[0]


def _check(g, alg, src, lab, n, row, col, singles):
    key = (alg, src)
    if key not in singles:
        singles[key] = _single(g, alg, src)
    if alg in ("ppr", "pr"):
        assert np.abs(lab.astype(np.float64) - singles[key].astype(np.float64)).sum() < PR_L1_TOL
        ref, _, _ = O.pagerank(n, row, col, 0.85, 1e-12, personalised=(alg == "ppr"), source=src)
        assert np.abs(lab.astype(np.float64) - ref).sum() < PR_L1_TOL
    else:
        assert np.array_equal(lab, singles[key]), key


MODES = {"default": {}, "planes": {"plane_engines": True},
         "lanes": {"bfs_engine": False, "ppr_engine": False, "lanes": 3}}


@pytest.mark.parametrize("mode", list(MODES))
def test_mixed_stream_alone_equals_in_mix(rmat12, mode):
    """Two client threads submit a 40/30/20/10 BFS/SSSP/SSWP/PPR mix with random gaps; small
    engines (24 columns) force queries to wait for columns and to join running traversals;
    the lanes run several queries at once."""
    n, s, d, w, g, row, col = rmat12
    rng = np.random.default_rng(7)
    mix = rng.choice(["bfs", "sssp", "sswp", "ppr"], size=240, p=[0.4, 0.3, 0.2, 0.1])
    srcs = rng.integers(0, n, size=240)
    srcs[::7] = g.info()["argmax_out_degree"]
    results = {}
    lock = threading.Lock()
    with eg.Server(g, max_batch=24, **MODES[mode]) as srv:
        def client(part):
            for i in range(part, 240, 2):
                t = srv.submit(mix[i], int(srcs[i]))
                with lock:
                    results[i] = t
                time.sleep(float(rng.exponential(2e-4)) if part == 0 else 0.0)
        ths = [threading.Thread(target=client, args=(p,)) for p in (0, 1)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        answers = {}
        for i in range(240):
            lab, r = srv.wait(results[i], algo=mix[i])
            assert r["status"] == L.OK
            assert r["t_submit_ms"] <= r["t_start_ms"] <= r["t_done_ms"]
            answers[i] = (lab, r)
    singles = {}
    shared = 0
    for i in range(240):
        lab, r = answers[i]
        _check(g, mix[i], int(srcs[i]), lab, n, row, col, singles)
        if r["batch_size"] > 1:
            shared += 1
    if mode != "lanes":
        assert shared > 0, "no query shared an iteration with another one"


def test_late_arrivals_join_running_traversal(rmat12):
    """A query submitted while another is running starts before the first one finishes and
    still gets the single-query answer."""
    n, s, d, w, g, row, col = rmat12
    src = g.info()["argmax_out_degree"]
    with eg.Server(g, max_batch=64, plane_engines=True) as srv:
        t1 = srv.submit("sssp", src)
        time.sleep(0.0005)
        others = [srv.submit("sssp", int(x)) for x in (1, 2, 3, src)]
        lab1, r1 = srv.wait(t1)
        res = [srv.wait(t) for t in others]
    singles = {}
    _check(g, "sssp", src, lab1, n, row, col, singles)
    for (lab, r), x in zip(res, (1, 2, 3, src)):
        _check(g, "sssp", x, lab, n, row, col, singles)


def test_column_reuse_many_rounds(rmat12):
    """Far more queries than columns: columns are given back and reset between occupants."""
    n, s, d, w, g, row, col = rmat12
    rng = np.random.default_rng(11)
    srcs = rng.integers(0, n, size=150)
    singles = {}
    with eg.Server(g, max_batch=5, plane_engines=True) as srv:
        for alg in ("bfs", "sswp"):
            tk = [srv.submit(alg, int(x)) for x in srcs[:50]]
            for t, x in zip(tk, srcs[:50]):
                lab, r = srv.wait(t)
                assert r["batch_size"] <= 5
                _check(g, alg, int(x), lab, n, row, col, singles)


def test_whole_graph_queries_and_cancel(rmat12):
    n, s, d, w, g, row, col = rmat12
    with eg.Server(g, max_batch=8) as srv:
        tc = srv.submit("cc")
        tp = srv.submit("pr")
        tk = [srv.submit("bfs", i) for i in range(20)]
        srv.cancel(tk[-1])
        cc, _ = srv.wait(tc)
        pr, _ = srv.wait(tp, algo="pr")
        for t, i in zip(tk[:-1], range(19)):
            lab, _ = srv.wait(t)
            assert np.array_equal(lab, _single(g, "bfs", i))
        with pytest.raises(eg.EEGError):
            srv.wait(tk[-1])  # cancelled or forgotten
    assert np.array_equal(cc, _single(g, "cc", 0))
    ref, _, _ = O.pagerank(n, row, col, 0.85, 1e-12)
    assert np.abs(pr.astype(np.float64) - ref).sum() < PR_L1_TOL


def test_discard_mode_and_errors(rmat12):
    n, s, d, w, g, row, col = rmat12
    with eg.Server(g, max_batch=16, discard_labels=True) as srv:
        tk = [srv.submit("sssp", i) for i in range(40)]
        for t in tk:
            lab, r = srv.wait(t)
            assert lab is None and r["status"] == L.OK
        with pytest.raises(eg.EEGError) as e:
            srv.submit("bfs", n)
        assert e.value.status == L.ERR_INVALID_ARG
        with pytest.raises(eg.EEGError):
            srv.wait(123456)
    # a budget too small for one column rejects traversal queries
    with eg.Server(g, label_budget_bytes=1024) as srv:
        with pytest.raises(eg.EEGError) as e:
            srv.submit("bfs", 0)
        assert e.value.status == L.ERR_REJECTED


def test_engine_counter_restarts_near_its_limit(rmat12):
    """Engines whose iteration counter starts just below the restart limit (0xf0000000): the
    first queries run past it, further admissions wait until the engine drains, then every
    column is reset and the counter restarts at 1.  Answers stay equal to eeg_run's across the
    restart, for the BFS engine (levels stored as counter + depth) and the plane engines."""
    n, s, d, w, g, row, col = rmat12
    rng = np.random.default_rng(13)
    srcs = rng.integers(0, n, size=48)
    srcs[::5] = g.info()["argmax_out_degree"]
    singles = {}
    with eg.Server(g, max_batch=8, plane_engines=True, first_iteration=0xf0000000 - 3) as srv:
        for alg in ("bfs", "sssp", "sswp"):
            tk = [srv.submit(alg, int(x)) for x in srcs]
            for t, x in zip(tk, srcs):
                lab, r = srv.wait(t)
                assert r["status"] == L.OK
                _check(g, alg, int(x), lab, n, row, col, singles)


@pytest.mark.parametrize("columns", [8, 16, 32])
def test_ppr_engine_columns_join_and_leave(rmat12, columns):
    """More PPR queries than the engine has columns, submitted while others run: each answer
    within L1 1e-6 of its single-query run and of the FP64 oracle; queries share iterations.
    columns = 0 is the default width (16)."""
    n, s, d, w, g, row, col = rmat12
    rng = np.random.default_rng(17)
    srcs = [int(x) for x in rng.integers(0, n, size=70)]
    srcs[::9] = [g.info()["argmax_out_degree"]] * len(srcs[::9])
    singles = {}
    with eg.Server(g, ppr_columns=columns) as srv:
        tk = []
        for i, x in enumerate(srcs):
            tk.append(srv.submit("ppr", x))
            if i % 10 == 0:
                time.sleep(0.002)
        res = [srv.wait(t, algo="ppr") for t in tk]
    shared = 0
    for (lab, r), x in zip(res, srcs):
        assert r["status"] == L.OK
        _check(g, "ppr", x, lab, n, row, col, singles)
        shared += r["batch_size"] > 1
    assert shared > 0


@pytest.mark.parametrize("lanes", [1, 2])
def test_lanes_batch_waiting_sssp_sswp(rmat12, lanes):
    """A lane that takes an SSSP or SSWP query while 31 or more of the same algorithm wait
    answers up to 256 of them in one multi-source pass: 150 SSSP and 100 SSWP queries (and some
    BFS on lanes between them) submitted at once, every answer equal to its single run, and the
    batch size reported.  The same stream with lane batching off answers one at a time."""
    n, s, d, w, g, row, col = rmat12
    rng = np.random.default_rng(31 + lanes)
    mix = ["sssp"] * 150 + ["sswp"] * 100 + ["bfs"] * 10
    rng.shuffle(mix)
    srcs = rng.integers(0, n, size=len(mix))
    srcs[::9] = g.info()["argmax_out_degree"]
    singles = {}
    for batch in (True, False):
        with eg.Server(g, lanes=lanes, bfs_engine=False, ppr_engine=False, lane_batch=batch) as srv:
            tickets = [srv.submit(a, int(x)) for a, x in zip(mix, srcs)]
            sizes = []
            for a, x, t in zip(mix, srcs, tickets):
                lab, r = srv.wait(t, algo=a)
                assert r["status"] == L.OK
                _check(g, a, int(x), lab, n, row, col, singles)
                if a != "bfs":
                    sizes.append(r["batch_size"])
        if batch:
            assert max(sizes) >= 32 and max(sizes) <= 256, max(sizes)
        else:
            assert max(sizes) == 1


def test_lane_batches_with_cancelled_queries(rmat12):
    """Queries cancelled while queued never join a lane batch (one a lane already took is
    answered); the rest of a long SSSP queue is answered in batches, each answer equal to its
    single run."""
    n, s, d, w, g, row, col = rmat12
    srcs = np.random.default_rng(41).integers(0, n, size=120)
    singles = {}
    with eg.Server(g, lanes=1, bfs_engine=False, ppr_engine=False) as srv:
        tk = [srv.submit("sssp", int(x)) for x in srcs]
        gone = set(range(3, 120, 5))
        for i in gone:
            srv.cancel(tk[i])
        sizes = []
        for i, (t, x) in enumerate(zip(tk, srcs)):
            if i in gone:
                try:
                    lab, r = srv.wait(t, algo="sssp")
                except eg.EEGError:
                    continue  # cancelled in the queue
            else:
                lab, r = srv.wait(t, algo="sssp")
            assert r["status"] == L.OK
            _check(g, "sssp", int(x), lab, n, row, col, singles)
            sizes.append(r["batch_size"])
        assert max(sizes) >= 32
