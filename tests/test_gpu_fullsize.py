"""GPU parity at the benchmark configurations' full sizes, against the oracle.

* 1M long records (200-1200 chars, seed 3), q=3, t=0.8: the whole join on the GPU;
  its index and probe counts against their closed forms; the joins of the two halves
  of the query range add up to it (pairs and every counter); three query ranges of
  300 (longest, middle, shortest records), each joined against the whole collection,
  equal the oracle pair for pair and counter for counter, and equal the whole join's
  pairs of those queries.
* 512 batched self-joins (sizes log-uniform in 500..50k, seed 4000+j, t in
  {0.50, ..., 0.95}), one graph-replayed batch run: every job equals the oracle.
* Serving against a 2M-title resident index (seed 5, t_min = 0.7): requests of up to
  4,096 probes (20% edited copies of indexed titles), alone and in one mixed run,
  each equal to the oracle's R-S join (pairs, similarities, counters).

The oracle verifies each candidate with the reference's merge alone (checks=False,
pinned to the checked join by tests/test_oracle.py), run on the host's cores.
"""
import concurrent.futures as cf
import os
import sys

import numpy as np
import pytest

from _util import oracle_sets

pytestmark = pytest.mark.gpu

HOST_THREADS = max(1, min(16, os.cpu_count() or 1))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tri(p):
    return np.stack([p.q, p.s, p.overlap], axis=1) if len(p) else np.zeros((0, 3), np.int32)


COUNTERS = (("probe_tokens", "p_q"), ("postings_scanned", "e_scan"), ("candidates", "candidates"),
            ("intersections", "intersections"), ("results", "results"))


def test_long_records_1m(fg, oracle_mod):
    t = float(np.float32(0.8))
    _say("C3: generating 1M long records (seed 3)")
    text = fg.generate_text(fg.GEN_LONG, 1_000_000, 3)
    coll = fg.Collection.from_text_bytes(text, 3, device=True)
    del text
    n = coll.info.n_sets
    assert n == 1_000_000
    _say(f"C3: {n} records, {coll.info.n_tokens} tokens; whole join")
    full = coll.selfjoin(t)
    st = full.stats
    _say(f"C3: {len(full)} pairs, {st['candidates']} candidates, {st['ms_total']:.0f} ms")
    ids, lens, pos, tok = coll.arrays()
    # index entries and probe tokens: closed-form sums over the record lengths
    ul, cnt = np.unique(lens, return_counts=True)
    mid = np.array([min(oracle_mod.bounds("mid_prefix", int(l), t), int(l)) for l in ul], np.int64)
    mx = np.array([min(oracle_mod.bounds("max_prefix", int(l), t), int(l)) for l in ul], np.int64)
    assert st["index_entries"] == int((mid * cnt).sum())
    assert st["probe_tokens"] == int((mx * cnt).sum())
    # the two halves of the query range add up to the whole join
    ft = _tri(full)
    halves = []
    for a, b in ((0, n // 2), (n // 2, n)):
        _say(f"C3: queries [{a}, {b})")
        halves.append(coll.selfjoin(t, queries=(a, b)))
    assert np.array_equal(np.concatenate([_tri(h) for h in halves]), ft)
    for k, _ in COUNTERS:
        assert sum(h.stats[k] for h in halves) == st[k], k
    for h in halves:
        assert h.stats["index_entries"] == st["index_entries"]
    # three ranges of 300 queries against the oracle
    ranges = ((0, 300), (n // 2, n // 2 + 300), (n - 300, n))
    got = [coll.selfjoin(t, queries=r) for r in ranges]
    sets = oracle_mod.Sets.from_arrays(ids, lens, pos, tok, coll.info.n_terms)
    qs = np.concatenate([np.arange(a, b) for a, b in ranges])
    _say(f"C3: oracle on {qs.size} queries ({HOST_THREADS} threads)")
    tri, ctr = oracle_mod.selfjoin(sets, t, nthreads=HOST_THREADS, queries=qs, checks=False)
    mine = np.concatenate([_tri(g) for g in got])
    assert mine.shape == tri.shape and np.array_equal(mine, tri)
    for k, ok in COUNTERS:
        assert sum(g.stats[k] for g in got) == ctr[ok], k
    assert all(g.stats["index_entries"] == ctr["e_idx"] for g in got)
    inr = np.zeros(n, bool)
    inr[qs] = True
    assert np.array_equal(ft[inr[ft[:, 0]]], tri)
    if len(tri):
        sims = np.array([oracle_mod.similarity(int(lens[q]), int(lens[s]), int(o)) for q, s, o in tri], np.float32)
        gs = np.concatenate([g.similarity for g in got])
        assert np.array_equal(gs.view(np.uint32), sims.view(np.uint32))


def _batch_workload(fg, n_jobs=512):
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(500), np.log(50000), n_jobs)).astype(int)
    ts = [float(np.float32(x)) for x in rng.choice(np.round(np.arange(0.5, 0.951, 0.05), 2), n_jobs)]
    colls = [fg.Collection.generate(fg.GEN_TITLES, int(m), 4000 + j, device=True) for j, m in enumerate(sizes)]
    return colls, ts


def test_batch_512_jobs(fg, oracle_mod):
    _say("C4: generating 512 collections")
    colls, ts = _batch_workload(fg)
    b = fg.Batch(colls, ts)
    _say("C4: batch run (graph)")
    st = b.run(graph=True)
    assert st["graph"] == 1
    got = b.all_pairs()

    def job(j):
        tri, ctr = oracle_mod.selfjoin(oracle_sets(oracle_mod, colls[j]), ts[j], nthreads=1, checks=False)
        return j, tri, ctr

    _say(f"C4: oracle on every job ({HOST_THREADS} concurrent)")
    with cf.ThreadPoolExecutor(HOST_THREADS) as ex:
        for j, tri, ctr in ex.map(job, range(len(colls))):
            g = got[j]
            assert np.array_equal(_tri(g), tri), f"job {j}"
            for k, ok in COUNTERS:
                assert g.stats[k] == ctr[ok], (j, k)
            assert g.stats["index_entries"] == ctr["e_idx"], j
    assert sum(len(g) for g in got) == st["n_pairs"]
    b.close()


def _edit(line, rng, n_edits):
    s = bytearray(line)
    for _ in range(n_edits):
        if not s:
            break
        i = int(rng.integers(len(s)))
        op = int(rng.integers(3))
        if op == 0:
            s[i] = int(rng.integers(97, 123))
        elif op == 1:
            del s[i]
        else:
            s.insert(i, int(rng.integers(97, 123)))
    return bytes(s)


def test_served_2m_index(fg, oracle_mod):
    t_min = float(np.float32(0.7))
    _say("C5: 2M-title index (seed 5), GPU tokenizer")
    text = fg.generate_text(fg.GEN_TITLES, 2_000_000, 5)
    idx = fg.Collection.from_text_bytes(text, 3, device=True)
    lines = [ln for ln in text.split(b"\n") if ln]
    del text
    rng = np.random.default_rng(505)
    sizes = [4096, 4096, 2500, 700, 90, 1]
    fresh = [ln for ln in fg.generate_text(fg.GEN_TITLES, sum(sizes), 51).split(b"\n") if ln]
    reqs, ts, f = [], [], 0
    for m in sizes:
        out = []
        for _ in range(m):
            if rng.random() < 0.2:
                out.append(_edit(lines[int(rng.integers(len(lines)))], rng, int(rng.integers(1, 4))))
            else:
                out.append(fresh[f])
                f += 1
        reqs.append(fg.Collection.from_text_dict(b"\n".join(out) + b"\n", idx))
        ts.append(float(np.float32(rng.uniform(0.7, 0.95))))
    del lines, fresh
    srv = fg.Server(idx, t_min, slots=2)
    _say("C5: each request alone")
    alone = [srv.result(srv.submit(r, t))[0] for r, t in zip(reqs, ts)]
    _say("C5: all requests in one mixed burst")
    srv.hold(True)
    rids = [srv.submit(r, t) for r, t in zip(reqs, ts)]
    srv.hold(False)
    mixed = [srv.result(rid) for rid in rids]
    assert max(st["batch_requests"] for _, st in mixed) > 1
    srv.close()
    S = oracle_sets(oracle_mod, idx)
    ids, lens = S.id, S.len
    for k, (r, t) in enumerate(zip(reqs, ts)):
        _say(f"C5: oracle request {k} ({r.info.n_sets} probes, t={t:.4f})")
        P = oracle_sets(oracle_mod, r)
        tri, ctr = oracle_mod.rsjoin(P, S, t, t_min, nthreads=HOST_THREADS)
        for g in (alone[k], mixed[k][0]):
            assert np.array_equal(_tri(g), tri), k
            for key, ok in COUNTERS:
                assert g.stats[key] == ctr[ok], (k, key)
            if len(tri):
                assert np.array_equal(g.rid, P.id[tri[:, 0]]) and np.array_equal(g.sid, ids[tri[:, 1]])
                sims = np.array([oracle_mod.similarity(int(P.len[a]), int(lens[b]), int(o)) for a, b, o in tri],
                                np.float32)
                assert np.array_equal(g.similarity.view(np.uint32), sims.view(np.uint32))
