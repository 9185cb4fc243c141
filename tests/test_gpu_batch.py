"""Batched self-joins: every job of one batch run (eager or a replayed HIP graph)
equals its own plain run and the oracle, bit for bit."""
import sys

import numpy as np
import pytest

from _util import golden_path, oracle_sets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(a, b):
    return (np.array_equal(a.q, b.q) and np.array_equal(a.s, b.s) and np.array_equal(a.overlap, b.overlap)
            and np.array_equal(a.rid, b.rid) and np.array_equal(a.sid, b.sid)
            and np.array_equal(a.similarity.view(np.uint32), b.similarity.view(np.uint32)))


def _jobs(fg):
    rng = np.random.default_rng(7)
    colls, ts = [], []
    for j in range(14):
        n = int(np.exp(rng.uniform(np.log(300), np.log(6000))))
        colls.append(fg.Collection.generate(fg.GEN_TITLES, n, 4000 + j))
        ts.append(float(rng.choice(np.arange(0.5, 0.951, 0.05))))
    colls.append(fg.Collection.from_bin(golden_path("small.txt.3gram.bin")))
    ts.append(0.25)
    colls.append(fg.Collection.from_bin_ints(np.zeros(0, np.int32)))   # an empty job
    ts.append(0.7)
    return colls, [float(np.float32(t)) for t in ts]


def test_batch_equals_plain_runs(fg, oracle_mod):
    colls, ts = _jobs(fg)
    plain = [c.selfjoin(t) for c, t in zip(colls, ts)]
    b = fg.Batch(colls, ts)
    for graph in (False, True, True, True):   # eager, capture + replay, two more replays
        _say(f"batch run graph={graph}")
        st = b.run(graph=graph)
        colls[1].selfjoin(ts[1])   # eager device work and a host sync between replays
        fg.device_synchronize()
        assert st["graph"] == (1 if graph else 0)
        assert st["n_pairs"] == sum(len(p) for p in plain)
        assert st["candidates"] == sum(p.stats["candidates"] for p in plain)
        every = b.all_pairs()   # all jobs in one call: the same pairs and counters
        assert len(every) == len(plain)
        for j, p in enumerate(plain):
            got = b.pairs(j)
            assert _same(got, p), f"job {j}"
            assert _same(every[j], p), f"job {j} (all_pairs)"
            for k in ("candidates", "intersections", "results", "postings_scanned"):
                assert got.stats[k] == p.stats[k], (j, k)
                assert every[j].stats[k] == p.stats[k], (j, k)
    # spot-check three jobs against the oracle directly
    for j in (0, 5, len(colls) - 2):
        tri, ctr = oracle_mod.selfjoin(oracle_sets(oracle_mod, colls[j]), ts[j])
        got = b.pairs(j)
        assert np.array_equal(np.stack([got.q, got.s, got.overlap], 1) if len(got) else np.zeros((0, 3), np.int32), tri)
        assert got.stats["candidates"] == ctr["candidates"]


def test_batch_new_thresholds_and_table_sizes(fg):
    colls, ts = _jobs(fg)
    b = fg.Batch(colls, ts)
    ts2 = [float(np.float32(min(0.95, t + 0.1))) for t in ts]
    b.set_thresholds(ts2)
    for slots, load in ((512, 65), (1024, 80), (256, 50)):
        _say(f"batch new thresholds slots={slots} load={load}")
        b.run(graph=True, table_slots=slots, load_pct=load)
        for j, c in enumerate(colls):
            assert _same(b.pairs(j), c.selfjoin(ts2[j])), f"job {j}"


def test_graph_replays_follow_growing_capacities(fg):
    """Lower thresholds run after run: item, result and overflow counts grow, and a
    capacity that grows inside a buffer's slack must still reach the replayed launches."""
    colls, ts = _jobs(fg)
    b = fg.Batch(colls, ts)
    for k in range(6):
        tk = [float(np.float32(max(0.3, t - 0.05 * k))) for t in ts]
        b.set_thresholds(tk)
        _say(f"batch graph replay, thresholds lowered {k} times")
        b.run(graph=True)
        for j, c in enumerate(colls):
            assert _same(b.pairs(j), c.selfjoin(tk[j])), (k, j)


def test_autotune_returns_a_valid_choice(fg, tmp_path):
    colls, ts = _jobs(fg)
    b = fg.Batch(colls[:6], ts[:6])
    best = b.autotune(reps=2)
    assert best["table_slots"] in (512, 1024) and best["load_pct"] in (50, 65, 80) and best["ms"] > 0
    assert best["index_groups"] in (0, 2, 4) and best["block_records"] in (0, 4096, 32768)
    b.run(graph=True, table_slots=best["table_slots"], load_pct=best["load_pct"],
          index_groups=best["index_groups"], block_records=best["block_records"])
    for j in range(6):
        assert _same(b.pairs(j), colls[j].selfjoin(ts[j]))
    # the tuning cache: tuned and stored once, then returned untimed; another workload misses
    cache = str(tmp_path / "tune.txt")
    first = b.autotune(reps=1, cache_path=cache)
    assert not first["from_cache"]
    again = b.autotune(reps=1, cache_path=cache)
    assert again["from_cache"]
    for k in ("table_slots", "load_pct", "index_groups", "block_records"):
        assert again[k] == first[k], k
    assert len(open(cache).read().splitlines()) == 1
    other = fg.Batch(colls[:5], ts[:5])
    assert not other.autotune(reps=1, cache_path=cache)["from_cache"]
    assert len(open(cache).read().splitlines()) == 2


@pytest.mark.parametrize("groups,block", [(2, 0), (3, 0), (5, 0), (16, 0), (1, 64), (1, 1000), (4, 4096)])
def test_partitioning_choices_give_identical_pairs(fg, oracle_mod, groups, block):
    """index groups (one index and run per group of jobs) and record blocks (items span at
    most that many sources): every choice yields each job's plain pairs and counters"""
    colls, ts = _jobs(fg)
    plain = [c.selfjoin(t) for c, t in zip(colls, ts)]
    b = fg.Batch(colls, ts)
    for graph in (False, True, True):
        _say(f"batch groups={groups} block_records={block} graph={graph}")
        st = b.run(graph=graph, index_groups=groups, block_records=block)
        assert st["n_pairs"] == sum(len(p) for p in plain)
        every = b.all_pairs()
        for j, p in enumerate(plain):
            assert _same(every[j], p), f"job {j}"
            for k in ("candidates", "intersections", "results", "postings_scanned"):
                assert every[j].stats[k] == p.stats[k], (j, k)
    tri, ctr = oracle_mod.selfjoin(oracle_sets(oracle_mod, colls[3]), ts[3])
    got = every[3]
    assert np.array_equal(np.stack([got.q, got.s, got.overlap], 1) if len(got) else np.zeros((0, 3), np.int32), tri)


@pytest.mark.parametrize("block", [16, 300, 5000])
def test_record_blocks_in_plain_joins(fg, oracle_mod, block):
    from _util import assert_join_equal
    coll = fg.Collection.generate(fg.GEN_TITLES, 20000, 47)
    sets = oracle_sets(oracle_mod, coll)
    _say(f"titles n=20000 block_records={block}")
    assert_join_equal(oracle_mod, sets, coll.selfjoin(0.6, block_records=block), 0.6)
