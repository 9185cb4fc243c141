"""Served R-S joins: every request's pairs equal the oracle's R-S join, bit for bit,
and are the same whether the request ran alone or continuously batched with others."""
import sys

import numpy as np
import pytest

from _util import oracle_sets

pytestmark = pytest.mark.gpu

T_MIN = float(np.float32(0.6))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b):
    return (np.array_equal(a.q, b.q) and np.array_equal(a.s, b.s) and np.array_equal(a.overlap, b.overlap)
            and np.array_equal(a.rid, b.rid) and np.array_equal(a.sid, b.sid)
            and np.array_equal(a.similarity.view(np.uint32), b.similarity.view(np.uint32)))


def _probe_text(index_text: bytes, n: int, seed: int, fg) -> bytes:
    """n lines: 70% copies of indexed lines with 0-3 character edits, 30% fresh titles."""
    rng = np.random.default_rng(seed)
    lines = index_text.split(b"\n")
    lines = [ln for ln in lines if ln]
    fresh = fg.generate_text(fg.GEN_TITLES, n, 10_000 + seed).split(b"\n")
    out = []
    for k in range(n):
        if rng.random() < 0.7:
            s = bytearray(lines[int(rng.integers(len(lines)))])
            for _ in range(int(rng.integers(0, 4))):
                if not s:
                    break
                i = int(rng.integers(len(s)))
                op = rng.integers(3)
                if op == 0:
                    s[i] = int(rng.integers(97, 123))
                elif op == 1:
                    del s[i]
                else:
                    s.insert(i, int(rng.integers(97, 123)))
            out.append(bytes(s))
        else:
            out.append(fresh[k % max(len(fresh) - 1, 1)])
    return b"\n".join(out) + b"\n"


@pytest.fixture(scope="module")
def world(fg, oracle_mod):
    text = fg.generate_text(fg.GEN_TITLES, 30000, 77)
    idx = fg.Collection.from_text_bytes(text, 3)
    rng = np.random.default_rng(3)
    reqs = []
    for k in range(20):
        n = int(np.exp(rng.uniform(0, np.log(600))))
        t = float(np.float32(rng.choice([0.6, 0.65, 0.7, 0.75, 0.8, 0.85, 0.9, 0.95])))
        reqs.append((fg.Collection.from_text_dict(_probe_text(text, n, 100 + k, fg), idx), t))
    # edge requests: no probes at all, probes of q-grams the index never saw, t = t_min
    reqs.append((fg.Collection.from_text_dict(b"ab\n", idx), 0.7))
    reqs.append((fg.Collection.from_text_dict(b"zzqqxxjjvv zzqqxxjj\nqqqqqqqqqqqqqqq\n", idx), 0.6))
    reqs.append((fg.Collection.from_text_dict(_probe_text(text, 50, 999, fg), idx), T_MIN))
    return {"text": text, "idx": idx, "sets": oracle_sets(oracle_mod, idx), "reqs": reqs}


def _check_vs_oracle(oracle_mod, world, probes, t, pairs):
    p = oracle_sets(oracle_mod, probes)
    tri, ctr = oracle_mod.rsjoin(p, world["sets"], t, t_index=T_MIN)
    got = np.stack([pairs.q, pairs.s, pairs.overlap], 1) if len(pairs) else np.zeros((0, 3), np.int32)
    assert np.array_equal(got, tri), f"t={t}: {got.shape[0]} pairs vs oracle {tri.shape[0]}"
    if len(pairs):
        s = world["sets"]
        assert np.array_equal(pairs.rid, p.id[tri[:, 0]])
        assert np.array_equal(pairs.sid, s.id[tri[:, 1]])
        sims = np.array([oracle_mod.similarity(int(p.len[a]), int(s.len[b]), int(o)) for a, b, o in tri], np.float32)
        assert np.array_equal(pairs.similarity.view(np.uint32), sims.view(np.uint32))
    st = pairs.stats
    assert st["postings_scanned"] == ctr["e_scan"]
    assert st["probe_tokens"] == ctr["p_q"]
    assert st["candidates"] == ctr["candidates"]
    assert st["intersections"] == ctr["intersections"]
    assert st["results"] == ctr["results"]


def test_requests_alone_match_oracle(fg, oracle_mod, world):
    srv = fg.Server(world["idx"], T_MIN, slots=1)
    for k, (probes, t) in enumerate(world["reqs"]):
        _say(f"alone request {k} probes={probes.info.n_sets} t={t}")
        rid = srv.submit(probes, t)
        pairs, st = srv.result(rid)
        assert st["batch_requests"] == 1
        _check_vs_oracle(oracle_mod, world, probes, t, pairs)
    srv.close()


@pytest.mark.parametrize("slots,budget", [(2, 0), (3, 1 << 20), (1, 1 << 40)])
def test_batched_equals_alone(fg, world, slots, budget):
    alone_srv = fg.Server(world["idx"], T_MIN, slots=1)
    alone = []
    for probes, t in world["reqs"]:
        alone.append(alone_srv.result(alone_srv.submit(probes, t))[0])
    alone_srv.close()
    _say(f"mixed slots={slots} budget={budget}")
    srv = fg.Server(world["idx"], T_MIN, slots=slots, budget_bytes=budget)
    srv.hold()
    ids = [srv.submit(p, t) for p, t in world["reqs"]]   # all queued before any run
    assert srv.poll() == (len(ids), 0)
    with pytest.raises(fg.FgssError):
        srv.wait(ids[0])                                  # held and nothing running
    srv.hold(False)
    shared = 0
    cap = budget if budget > 0 else 256 << 20
    for k in np.random.default_rng(slots).permutation(len(ids)):   # collect in a shuffled order
        pairs, st = srv.result(ids[k])
        shared = max(shared, st["batch_requests"])
        # admission: the charges in flight fit the budget, or the request ran alone
        assert st["charged_bytes"] > 0
        assert st["inflight_bytes"] >= st["charged_bytes"]
        assert st["inflight_bytes"] <= cap or (
            st["batch_requests"] == 1 and st["inflight_bytes"] == st["charged_bytes"]), (k, st)
        assert _same(pairs, alone[k]), f"request {k}"
        for key in ("candidates", "intersections", "results", "postings_scanned"):
            assert pairs.stats[key] == alone[k].stats[key], (k, key)
    if budget != 1 << 20:
        assert shared > 1   # requests really shared runs
    assert srv.poll() == (0, 0)
    srv.close()


@pytest.mark.parametrize("bounds_cap,stream_parts", [(0, 0), (1, 0), (0, 1)])
def test_overflow_resplit_and_small_tables(fg, oracle_mod, world, bounds_cap, stream_parts):
    # 256-slot tables at full load force the re-split rounds of overflowing items; a
    # one-entry boundary budget or stream_parts=1 makes split queries stream
    srv = fg.Server(world["idx"], T_MIN, slots=2, table_slots=256, load_pct=100, bounds_cap=bounds_cap,
                    stream_parts=stream_parts)
    probes, _ = world["reqs"][0]
    ids = [(srv.submit(p, t), p, t) for p, t in world["reqs"][:8]]
    for rid, p, t in ids:
        pairs, _ = srv.result(rid)
        _check_vs_oracle(oracle_mod, world, p, t, pairs)
    srv.close()


def test_long_records_index(fg, oracle_mod):
    text = fg.generate_text(fg.GEN_LONG, 3000, 5)
    idx = fg.Collection.from_text_bytes(text, 3)
    sets = oracle_sets(oracle_mod, idx)
    lines = [ln for ln in text.split(b"\n") if ln]
    rng = np.random.default_rng(1)
    probe_text = b"\n".join(lines[int(i)][: int(rng.integers(150, 1200))] for i in rng.integers(0, len(lines), 300))
    probes = fg.Collection.from_text_dict(probe_text + b"\n", idx)
    srv = fg.Server(idx, 0.5, slots=2)
    for t in (0.5, 0.8):
        t = float(np.float32(t))
        pairs, _ = srv.result(srv.submit(probes, t))
        tri, ctr = oracle_mod.rsjoin(oracle_sets(oracle_mod, probes), sets, t, t_index=float(np.float32(0.5)))
        got = np.stack([pairs.q, pairs.s, pairs.overlap], 1) if len(pairs) else np.zeros((0, 3), np.int32)
        assert np.array_equal(got, tri), t
        assert pairs.stats["candidates"] == ctr["candidates"]
    srv.close()


def test_rejects_bad_requests(fg, world):
    srv = fg.Server(world["idx"], T_MIN, slots=1)
    probes, _ = world["reqs"][0]
    with pytest.raises(fg.FgssError):
        srv.submit(probes, 0.5)        # below t_min
    with pytest.raises(fg.FgssError):
        srv.submit(probes, 1.0)
    with pytest.raises(fg.FgssError):
        srv.result(12345)              # unknown request
    srv.close()
    with pytest.raises(fg.FgssError):
        fg.Server(world["idx"], 0.0)


def test_empty_index_and_empty_records(fg, oracle_mod, world):
    from _util import records_to_bin
    empty = fg.Collection.from_bin_ints(np.zeros(0, np.int32))
    srv = fg.Server(empty, T_MIN, slots=1)
    probes, t = world["reqs"][0]
    pairs, _ = srv.result(srv.submit(probes, t))
    assert len(pairs) == 0
    srv.close()
    # an index holding empty records: they never join
    recs = [np.arange(12), np.arange(10), np.arange(6), np.zeros(0), np.zeros(0)]
    idx = fg.Collection.from_bin_ints(records_to_bin(recs))
    prb = fg.Collection.from_bin_ints(records_to_bin([np.zeros(0), np.arange(11), np.arange(5), np.array([3, 70])]))
    srv = fg.Server(idx, 0.5, slots=1)
    t = float(np.float32(0.5))
    pairs, _ = srv.result(srv.submit(prb, t))
    tri, _ = oracle_mod.rsjoin(oracle_sets(oracle_mod, prb), oracle_sets(oracle_mod, idx), t, t_index=t)
    got = np.stack([pairs.q, pairs.s, pairs.overlap], 1) if len(pairs) else np.zeros((0, 3), np.int32)
    assert np.array_equal(got, tri) and len(tri) > 0
    srv.close()


@pytest.mark.parametrize("t", [0.6, 0.8, 0.9])
def test_served_against_itself_is_the_selfjoin(fg, t):
    # the longer record of a pair is its query, so the q < s half is the self-join
    coll = fg.Collection.generate(fg.GEN_TITLES, 20000, 12)
    t = float(np.float32(t))
    srv = fg.Server(coll, 0.6, slots=2)
    pairs, _ = srv.result(srv.submit(coll, t))
    srv.close()
    sj = coll.selfjoin(t)
    up = pairs.q < pairs.s
    assert np.array_equal(pairs.q[up], sj.q) and np.array_equal(pairs.s[up], sj.s)
    assert np.array_equal(pairs.overlap[up], sj.overlap)
    assert np.array_equal(pairs.similarity[up].view(np.uint32), sj.similarity.view(np.uint32))
    diag = pairs.q == pairs.s
    assert diag.sum() == int((coll.arrays()[1] > 0).sum())
