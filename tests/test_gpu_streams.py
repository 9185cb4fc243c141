"""GPU: joins on caller-owned HIP streams and joins over a query range.

* Two collections joined at the same time from two host threads, each call on the
  thread's own stream (fgss_join_opts.stream), several times: every join equals the
  oracle.  A batch run on a caller stream equals its run on the batch's own stream.
* A query range [a, b) joins those queries against every later record: its pairs are
  the whole join's pairs of those queries, its counters the oracle's over them; ranges
  that partition the records add up to the whole join.
"""
import ctypes as C
import sys
import threading

import numpy as np
import pytest

from _util import assert_join_equal, hip_runtime, oracle_sets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tri(p):
    return np.stack([p.q, p.s, p.overlap], axis=1) if len(p) else np.zeros((0, 3), np.int32)


def test_two_collections_on_two_caller_streams(fg, oracle_mod):
    hip = hip_runtime()
    streams = [C.c_void_p(), C.c_void_p()]
    for s in streams:
        assert hip.hipStreamCreate(C.byref(s)) == 0
    colls = [fg.Collection.generate(fg.GEN_TITLES, 30000, 61), fg.Collection.generate(fg.GEN_LONG, 2000, 62)]
    ts = [0.6, 0.7]
    sets = [oracle_sets(oracle_mod, c) for c in colls]
    want = [oracle_mod.selfjoin(s, t) for s, t in zip(sets, ts)]
    for c in colls:
        c.upload()
    got = [[], []]
    errors = []

    def worker(k):
        try:
            for _ in range(4):
                got[k].append(colls[k].selfjoin(ts[k], stream=streams[k].value))
        except Exception as e:   # reported by the main thread
            errors.append(e)

    _say("two threads, two caller streams, 4 joins each")
    th = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    for k in range(2):
        assert len(got[k]) == 4
        for p in got[k]:
            assert_join_equal(oracle_mod, sets[k], p, ts[k], counters=want[k])
    # a batch on a caller stream, replayed graph, equals its own-stream run
    b = fg.Batch(colls, ts)
    b.run(graph=True)
    own = b.all_pairs()
    b.run(graph=True, stream=streams[0].value)
    for p, q in zip(b.all_pairs(), own):
        assert np.array_equal(_tri(p), _tri(q))
    b.close()
    for s in streams:
        hip.hipStreamDestroy(s)


@pytest.mark.parametrize("t", [0.6, 0.9])
def test_query_ranges(fg, oracle_mod, t):
    coll = fg.Collection.generate(fg.GEN_TITLES, 100000, 2)
    n = coll.info.n_sets
    full = coll.selfjoin(t)
    ft = _tri(full)
    sets = oracle_sets(oracle_mod, coll)
    cuts = [0, 1, 777, 40000, 99999, n]
    parts = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        _say(f"titles 100k t={t} queries [{a}, {b})")
        p = coll.selfjoin(t, queries=(a, b))
        tri, ctr = oracle_mod.selfjoin(sets, t, queries=np.arange(a, b), checks=False)
        assert np.array_equal(_tri(p), tri)
        for k, ok in (("probe_tokens", "p_q"), ("postings_scanned", "e_scan"), ("candidates", "candidates"),
                      ("intersections", "intersections"), ("results", "results"), ("index_entries", "e_idx")):
            assert p.stats[k] == ctr[ok], (a, b, k)
        parts.append(p)
    assert np.array_equal(np.concatenate([_tri(p) for p in parts]), ft)
    for k in ("probe_tokens", "postings_scanned", "candidates", "intersections", "results"):
        assert sum(p.stats[k] for p in parts) == full.stats[k], k
    empty = coll.selfjoin(t, queries=(5, 5))   # an empty range: no queries ((0, 0) means all)
    assert len(empty) == 0 and empty.stats["probe_tokens"] == 0
    with pytest.raises(fg.FgssError):
        coll.selfjoin(t, queries=(10, n + 1))
