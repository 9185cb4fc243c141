"""GPU parity: the HIP self-join through the C ABI vs the CPU oracle, bit for bit.

Every case compares the sorted (q, s, overlap) triples, the printed ids and
float32 similarities, and the reference's counters (# Inverted Index Entries,
# Candidates, # Intersections, # Results) plus the probe/posting counts.
"""
import json
import sys

import numpy as np
import pytest

from _util import assert_join_equal, dblp_text, filter_counts, golden_path, oracle_sets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_fixture_known_answers(fg, oracle_mod):
    coll = fg.Collection.from_bin(golden_path("small.txt.3gram.bin"))
    sets = oracle_sets(oracle_mod, coll)
    ka = json.load(open(golden_path("small_known_answers.json")))
    for ts, lines in ka["answers"].items():
        t = fg.parse_threshold(ts)
        _say(f"fixture t={ts}")
        p = coll.selfjoin(t)
        assert sorted(p.lines()) == sorted(lines)
        assert_join_equal(oracle_mod, sets, p, t)


@pytest.mark.parametrize("t", [0.5, 0.7, 0.8, 0.9, 0.95])
def test_titles_4096(fg, oracle_mod, t):
    _say(f"titles n=4096 seed=1 t={t}")
    coll = fg.Collection.generate(fg.GEN_TITLES, 4096, 1)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t)


@pytest.mark.parametrize("slots", [256, 512, 1024, 2048, 4096])
def test_table_sizes_and_splitting(fg, oracle_mod, slots):
    """Tiny LDS tables force source-window splits and overflow retries; results must not change."""
    _say(f"titles n=4096 seed=7 t=0.5 slots={slots}")
    coll = fg.Collection.generate(fg.GEN_TITLES, 4096, 7)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(0.5, table_slots=slots), 0.5)


@pytest.mark.parametrize("t", [0.6, 0.75, 0.9])
def test_titles_100k(fg, oracle_mod, t):
    _say(f"titles n=100000 seed=2 t={t}")
    coll = fg.Collection.generate(fg.GEN_TITLES, 100000, 2)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t)


@pytest.mark.parametrize("t", [0.5, 0.8])
def test_dblp_18k(fg, oracle_mod, t):
    _say(f"dblp_t_18k q=3 t={t}")
    coll = fg.Collection.from_text_bytes(dblp_text(), 3)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t)


@pytest.mark.parametrize("t", [0.5, 0.8])
def test_long_records(fg, oracle_mod, t):
    """200-1200 char records: probe prefixes longer than one staged token batch at t=0.5."""
    _say(f"long n=1500 seed=3 t={t}")
    coll = fg.Collection.generate(fg.GEN_LONG, 1500, 3)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t)


def test_edge_collections(fg, oracle_mod):
    cases = {
        "empty": np.zeros(0, np.int32),
        "single": np.array([7, 3, 1, 2, 3], np.int32),
        "identical": np.array([1, 3, 1, 2, 3, 2, 3, 1, 2, 3], np.int32),
        "len1": np.array([1, 1, 5, 2, 1, 5, 3, 1, 6], np.int32),
        "with_empty_record": np.array([9, 0, 1, 2, 4, 5, 2, 2, 4, 5], np.int32),
        # t=0.8 boundary: |Q|=|S|=9, overlap 8 -> min_overlap = 8.000001f shortens the mid prefix
        "prefix_boundary": np.array([1, 9, 0] + list(range(2, 10)) + [2, 9] + list(range(1, 10)), np.int32),
        "prefix_hit": np.array([1, 9] + list(range(9)) + [2, 9] + list(range(8)) + [9], np.int32),
        # t=0.6 boundary: floor(9/0.6f) = 14 < 15 drops a subset pair
        "size_boundary": np.array([1, 9] + list(range(9)) + [2, 15] + list(range(15)), np.int32),
    }
    for name, ints in cases.items():
        for t in (0.3, 0.6, 0.8, 0.9):
            _say(f"edge {name} t={t}")
            coll = fg.Collection.from_bin_ints(ints)
            sets = oracle_mod.Sets(ints) if ints.size else oracle_sets(oracle_mod, coll)
            assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t)


def test_repeatable(fg):
    coll = fg.Collection.generate(fg.GEN_TITLES, 20000, 11)
    a = coll.selfjoin(0.7)
    b = coll.selfjoin(0.7)
    assert np.array_equal(a.q, b.q) and np.array_equal(a.s, b.s) and np.array_equal(a.overlap, b.overlap)
    assert a.stats["candidates"] == b.stats["candidates"]


def test_repeated_joins_equal_the_oracle(fg, oracle_mod):
    """Joins of one collection that reuse its device workspace, across thresholds and
    with another collection's joins and host syncs in between, each equal the oracle."""
    coll = fg.Collection.generate(fg.GEN_TITLES, 6000, 21)
    other = fg.Collection.generate(fg.GEN_TITLES, 3000, 22)
    sets = oracle_sets(oracle_mod, coll)
    want = {t: oracle_mod.selfjoin(sets, t) for t in (0.6, 0.8)}
    for rep in range(3):
        for t in (0.6, 0.8, 0.6):
            _say(f"graph replay rep={rep} t={t}")
            assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t, counters=want[t])
            other.selfjoin(0.7)
            fg.device_synchronize()


def test_cli_outputs(fg, oracle_mod, tmp_path):
    """The CLI's stdout lines and stderr counters, both input forms."""
    import subprocess
    ka = json.load(open(golden_path("small_known_answers.json")))
    for form in ("bin", "text"):
        for ts in ("0.25", "0.5", "0.75", "0.9"):
            _say(f"cli {form} t={ts}")
            if form == "bin":
                args = [fg.CLI_PATH, golden_path("small.txt.3gram.bin"), ts]
            else:
                out = tmp_path / "o.bin"
                args = [fg.CLI_PATH, "-f", golden_path("small.txt"), "-q", "3", "-t", ts, "-o", str(out)]
            r = subprocess.run(args, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            assert sorted(r.stdout.splitlines()) == sorted(ka["answers"][ts])
            assert f"# Results: {len(ka['answers'][ts])}" in r.stderr
            assert "# Records: 5" in r.stderr and "# Terms: 42" in r.stderr and "# Tokens: 88" in r.stderr
            if form == "text":
                assert out.read_bytes() == open(golden_path("small.txt.3gram.bin"), "rb").read()
    # counters against the oracle on a larger input
    coll = fg.Collection.generate(fg.GEN_TITLES, 3000, 4)
    path = tmp_path / "g.bin"
    coll.write_bin(str(path))
    sets = oracle_mod.Sets(np.fromfile(str(path), "<i4"))
    tri, ctr = oracle_mod.selfjoin(sets, np.float32(0.7))
    r = subprocess.run([fg.CLI_PATH, str(path), "0.7"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert f"# Candidates: {ctr['candidates']}" in r.stderr
    assert f"# Intersections: {ctr['intersections']}" in r.stderr
    assert f"# Inverted Index Entries: {ctr['e_idx']}" in r.stderr
    want = sorted("%d %d %f" % (sets.id[q], sets.id[s], oracle_mod.similarity(int(sets.len[q]), int(sets.len[s]), int(o)))
                  for q, s, o in tri)
    assert sorted(r.stdout.splitlines()) == want


def _expected_index_dump(oracle_mod, sets, t):
    """Per term, the ids of the records whose (clamped) mid prefix holds it, in
    internal order -- built from the oracle's bound expression."""
    lists = [[] for _ in range(sets.n_terms)]
    for s in range(sets.n):
        mp = min(oracle_mod.bounds("mid_prefix", int(sets.len[s]), t), int(sets.len[s]))
        for tok in sets.record(s)[:mp]:
            lists[int(tok)].append(int(sets.id[s]))
    out = []
    for term, ids in enumerate(lists):
        out.append(f"term: [{term}], size: {len(ids)}")
        out.append("".join(f"({i}) " for i in ids))
        out.append("")
    return out


def test_cli_debug_index_dump_and_host_tokenizer(fg, oracle_mod, tmp_path):
    import subprocess
    # the fixture through the debug form: sets on stdout, the index on stderr
    t = oracle_mod.parse_threshold("0.5")
    sets = oracle_mod.Sets(np.fromfile(golden_path("small.txt.3gram.bin"), "<i4"))
    r = subprocess.run([fg.CLI_PATH, golden_path("small.txt.3gram.bin"), "0.5", "debug"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = _expected_index_dump(oracle_mod, sets, t)
    lines = r.stderr.splitlines()
    k = lines.index(want[0])
    assert lines[k:k + len(want)] == want
    assert "id: [4]" in r.stdout and "size: [24]" in r.stdout
    # a bigger collection, and the text form with either tokenizer
    text = dblp_text()[:300000]
    src = tmp_path / "d.txt"
    src.write_bytes(text)
    binp = tmp_path / "d.bin"
    fg.Collection.from_text_bytes(text, 3, device=True).write_bin(str(binp))
    s2 = oracle_mod.Sets(np.fromfile(str(binp), "<i4"))
    r = subprocess.run([fg.CLI_PATH, str(binp), "0.8", "debug"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    want = _expected_index_dump(oracle_mod, s2, oracle_mod.parse_threshold("0.8"))
    lines = r.stderr.splitlines()
    k = lines.index(want[0])
    assert lines[k:k + len(want)] == want
    outs = []
    for extra in ([], ["--host-tokenizer"]):
        r = subprocess.run([fg.CLI_PATH, "-f", str(src), "-q", "3", "-t", "0.8"] + extra, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(sorted(r.stdout.splitlines()))
    assert outs[0] == outs[1] and len(outs[0]) > 0


@pytest.mark.parametrize("bounds_cap", [1, 300, 20000])
def test_split_queries_beyond_the_boundary_budget(fg, oracle_mod, bounds_cap):
    """Split queries whose boundary rows exceed the budget search their posting ranges
    per part instead; results must not change."""
    _say(f"titles n=4096 seed=7 t=0.5 slots=256 bounds_cap={bounds_cap}")
    coll = fg.Collection.generate(fg.GEN_TITLES, 4096, 7)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(0.5, table_slots=256, bounds_cap=bounds_cap), 0.5)
    _say(f"long n=1500 seed=3 t=0.5 bounds_cap={bounds_cap}")
    coll = fg.Collection.generate(fg.GEN_LONG, 1500, 3)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(0.5, bounds_cap=bounds_cap), 0.5)


def test_titles_100k_small_boundary_budget(fg, oracle_mod):
    _say("titles n=100000 seed=2 t=0.6 bounds_cap=5000")
    coll = fg.Collection.generate(fg.GEN_TITLES, 100000, 2, device=True)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(0.6, bounds_cap=5000), 0.6)


@pytest.mark.parametrize("slots,load", [(512, 65), (256, 100)])
def test_streamed_queries(fg, oracle_mod, slots, load):
    """stream_parts=1 walks every split query part by part in one wavefront (posting
    cursors, galloping search); with 256-slot tables at full load some parts overflow
    and are re-split on their own.  Results must not change."""
    for gen, n, seed, ts in ((fg.GEN_TITLES, 4096, 7, (0.5, 0.7)), (fg.GEN_LONG, 1500, 3, (0.5, 0.8))):
        coll = fg.Collection.generate(gen, n, seed)
        sets = oracle_sets(oracle_mod, coll)
        for t in ts:
            _say(f"streamed gen={gen} n={n} t={t} slots={slots} load={load}")
            p = coll.selfjoin(t, table_slots=slots, load_pct=load, stream_parts=1)
            assert_join_equal(oracle_mod, sets, p, t)


def test_titles_100k_streamed(fg, oracle_mod):
    _say("titles n=100000 seed=2 t=0.6 stream_parts=2")
    coll = fg.Collection.generate(fg.GEN_TITLES, 100000, 2, device=True)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(0.6, stream_parts=2), 0.6)


def test_titles_100k_streamed_window_retries(fg, oracle_mod):
    """Streamed items join several consecutive parts per window while the table stays
    sparse; with 256-slot tables some multi-part windows overflow and run again part
    by part (counted in window_retries).  Results must not change."""
    _say("titles n=100000 seed=2 t=0.6 slots=256 stream_parts=1")
    coll = fg.Collection.generate(fg.GEN_TITLES, 100000, 2, device=True)
    sets = oracle_sets(oracle_mod, coll)
    p = coll.selfjoin(0.6, table_slots=256, stream_parts=1)
    assert p.stats["window_retries"] > 0
    assert_join_equal(oracle_mod, sets, p, 0.6)


@pytest.mark.parametrize("load", [3, 20])
def test_streamed_chunks(fg, oracle_mod, load):
    """A few postings per part: heavy queries get many parts, streamed 64 parts per
    item, each item positioning its cursors at its own window start."""
    coll = fg.Collection.generate(fg.GEN_TITLES, 4096, 7)
    sets = oracle_sets(oracle_mod, coll)
    for t in (0.5, 0.8):
        _say(f"streamed chunks t={t} load={load}")
        assert_join_equal(oracle_mod, sets, coll.selfjoin(t, table_slots=256, load_pct=load, stream_parts=1), t)


def test_resplit_round_larger_than_the_result_buffer(fg, oracle_mod):
    """700 identical records among 20,000 unrelated ones of the same length: the
    planned parts of the first 440 queries hold more distinct sources than a 256-slot
    table, so they are re-split, and the re-split round finds about 212k pairs, more
    than a fresh collection's 65,536-triple result buffer.  That round is launched
    again with a larger buffer; the counters of its first launch must not count twice."""
    from _util import records_to_bin
    recs = [list(range(20))] * 700 + [list(range(100 + 20 * i, 120 + 20 * i)) for i in range(20000)]
    ints = records_to_bin(recs)
    sets = oracle_mod.Sets(ints)
    want = oracle_mod.selfjoin(sets, 0.5)
    coll = fg.Collection.from_bin_ints(ints)
    _say("resplit retry: 700 duplicates + 20000 singles, t=0.5, table_slots=256")
    p = coll.selfjoin(0.5, table_slots=256)
    assert p.stats["kernel_launches"] >= 3, p.stats["kernel_launches"]   # main, re-split, its repeat
    assert_join_equal(oracle_mod, sets, p, 0.5, counters=want)
    _say("resplit retry: the same collection again (buffers now sized)")
    assert_join_equal(oracle_mod, sets, coll.selfjoin(0.5, table_slots=256), 0.5, counters=want)


@pytest.mark.parametrize("dense_pct", [-1, 1])
def test_filter_counters(fg, dense_pct):
    """The counters of the filters after the reference's candidates -- positional-filter
    survivors (signature_checks) and those passing the 64-bucket bound
    (signature_checks_128) -- equal their host restatement, in hash-table parts (-1) and
    in dense parts (1)."""
    cases = [(fg.Collection.generate(fg.GEN_TITLES, 4096, 1), 0.6), (fg.Collection.generate(fg.GEN_LONG, 3000, 13), 0.8)]
    for coll, t in cases:
        t = float(np.float32(t))
        cand, keep, pass64 = filter_counts(coll, t)
        _say(f"filter counters dense_pct={dense_pct} n={coll.info.n_sets} t={t}")
        st = coll.selfjoin(t, dense_pct=dense_pct).stats
        assert (st["candidates"], st["signature_checks"], st["signature_checks_128"]) == (cand, keep, pass64)
        assert st["verified"] <= pass64


@pytest.mark.parametrize("dense_pct", [1, 0, -1])
def test_dense_parts(fg, oracle_mod, dense_pct):
    """Windows joined in dense parts (a table indexed by the source, 2 x table-slots
    sources a part) give the oracle's pairs and counters, as hash-table parts do: dense
    parts forced wherever a window spans two parts (1), the default rule (0), none (-1)."""
    cases = [(fg.Collection.generate(fg.GEN_LONG, 20000, 13), (0.5, 0.8), (512, 1024)),
             (fg.Collection.generate(fg.GEN_TITLES, 30000, 14), (0.5, 0.7), (512,))]
    for coll, ts, slots in cases:
        sets = oracle_sets(oracle_mod, coll)
        for t in ts:
            want = oracle_mod.selfjoin(sets, t)
            for cap in slots:
                _say(f"dense_pct={dense_pct} n={coll.info.n_sets} t={t} slots={cap}")
                p = coll.selfjoin(t, dense_pct=dense_pct, table_slots=cap)
                if dense_pct == 1:
                    assert p.stats["dense_parts"] > 0
                if dense_pct == -1:
                    assert p.stats["dense_parts"] == 0
                # the dense parts' postings are part of the postings scanned
                assert 0 <= p.stats["dense_postings"] <= p.stats["postings_scanned"]
                assert (p.stats["dense_postings"] > 0) == (p.stats["dense_parts"] > 0)
                assert_join_equal(oracle_mod, sets, p, t, counters=want)


def _overflow_collection(n_bad=1200, n_clean=1300, seed=23):
    """Records of 300 tokens: the first n_bad hold 260 tokens of one 64-bucket (a count
    above 255: an overflowed signature, which the signature bounds must not use), the
    others spread over the buckets; a tenth of each group are near-copies of another
    record of the group.  Tokens ascending in each record, all records the same length."""
    rng = np.random.default_rng(seed)
    ids = np.arange(60000, dtype=np.int64)
    b64 = ((ids * 0x9E3779B1) & 0xFFFFFFFF) >> 26
    pool_a = ids[b64 == 0][:280]
    other = ids[b64 != 0]
    pool_b, pool_c = other[:300], other[300:700]
    recs = []
    for g, n in ((0, n_bad), (1, n_clean)):
        first = len(recs)
        for i in range(n):
            if i > 10 and rng.random() < 0.1:
                r = recs[first + int(rng.integers(0, i))].copy()
                pool = pool_a if g == 0 else pool_c
                free = np.setdiff1d(pool, r)
                k = int(rng.integers(0, r.size))
                r[k] = free[int(rng.integers(0, free.size))]
                r = np.sort(r)
            elif g == 0:
                r = np.sort(np.concatenate([rng.choice(pool_a, 260, replace=False), rng.choice(pool_b, 40, replace=False)]))
            else:
                r = np.sort(rng.choice(pool_c, 300, replace=False))
            recs.append(r.astype(np.int32))
    return recs


@pytest.mark.parametrize("dense_pct", [1, 0])
def test_dense_signature_overflow(fg, oracle_mod, dense_pct):
    """Sources whose 64-bucket counts overflow a byte (their signature bounds are skipped)
    in grid parts beside parts without any: pairs and counters equal the oracle's, and the
    filter counters their host restatement."""
    from _util import records_to_bin
    coll = fg.Collection.from_bin_ints(records_to_bin(_overflow_collection()))
    sets = oracle_sets(oracle_mod, coll)
    _, lens, pos, tok = coll.arrays()
    b64 = ((tok.astype(np.int64) * 0x9E3779B1) & 0xFFFFFFFF) >> 26
    rec = np.repeat(np.arange(len(lens)), lens)
    per = np.zeros((len(lens), 64), np.int64)
    np.add.at(per, (rec, b64), 1)
    nbad = int((per > 255).any(axis=1).sum())
    assert 0 < nbad < len(lens)
    t = float(np.float32(0.8))
    _say(f"signature overflow dense_pct={dense_pct}: {nbad} of {len(lens)} records overflow")
    p = coll.selfjoin(t, dense_pct=dense_pct)
    if dense_pct == 1:
        assert p.stats["dense_parts"] > 0
    assert len(p) > 0
    assert_join_equal(oracle_mod, sets, p, t)
    cand, keep, pass64 = filter_counts(coll, t)
    assert (p.stats["candidates"], p.stats["signature_checks"], p.stats["signature_checks_128"]) == (cand, keep, pass64)
