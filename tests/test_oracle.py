"""Pinning the CPU oracle (oracle/fgss_oracle.c) before it is trusted as the checker.

* tokenizer + reordering: the reference fixture small.txt -> .bin/.loc bytes, and the
  reference's own tokenizer sources (oracle/_ref/libref_tok.so) on the 18k DBLP titles;
* bound expressions: exact hand-derived values, the reference header compiled as-is
  (oracle/_ref/libref_bounds.so) and an independent float32 evaluation in numpy;
* join: closed form == literal restatement of the reference tile loop for block sizes
  {1, 7, N}; subset of brute force with every difference at a float boundary; the
  known answers on the fixture collection; hand-built float-boundary cases.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from _util import dblp_text, golden_path

REF_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")


def _fixture_bin():
    return np.fromfile(golden_path("small.txt.3gram.bin"), dtype="<i4")


def test_tokenize_matches_fixture_bytes(oracle_mod):
    text = open(golden_path("small.txt"), "rb").read()
    b, l = oracle_mod.tokenize(text, 3)
    assert np.array_equal(b, _fixture_bin())
    assert np.array_equal(l, np.fromfile(golden_path("small.txt.3gram.bin.loc"), dtype="<i4"))


def test_bin_load_reverses_file_order(oracle_mod):
    s = oracle_mod.Sets(_fixture_bin())
    assert s.id.tolist() == [4, 2, 3, 1, 5]
    assert s.len.tolist() == [24, 19, 18, 16, 11]
    assert s.n_terms == 42 and s.n_tokens == 88


def _rank_rule(hashes, rec_len):
    """Independent numpy restatement of the reordering rule over per-record hashes."""
    h = np.asarray(hashes, dtype=np.uint64)
    uniq, first, inv, cnt = np.unique(h, return_index=True, return_inverse=True, return_counts=True)
    order = np.lexsort((first, cnt))          # df ascending, then first occurrence
    rank = np.empty(uniq.size, np.int64)
    rank[order] = np.arange(uniq.size)
    tid = rank[inv]
    offs = np.concatenate([[0], np.cumsum(rec_len)])
    recs = [(i, np.sort(tid[offs[i]:offs[i + 1]])) for i in range(len(rec_len)) if rec_len[i] > 0]
    recs.sort(key=lambda r: r[1].size)        # stable: size ascending, line order
    out = []
    for i, toks in recs:
        out += [i + 1, toks.size] + toks.tolist()
    return np.array(out, np.int32)


def _ref_tok():
    so = os.path.join(REF_DIR, "libref_tok.so")
    if not os.path.exists(so):
        pytest.skip("oracle/_ref/libref_tok.so not built (needs the reference sources)")
    L = C.CDLL(so)
    L.ref_qgram_hashes.restype = C.c_int64
    L.ref_qgram_hashes.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.POINTER(C.c_uint64)),
                                   C.POINTER(C.POINTER(C.c_int64)), C.POINTER(C.c_int64)]
    L.ref_free.argtypes = [C.c_void_p]
    return L


@pytest.mark.parametrize("source,q", [("small", 3), ("small", 2), ("dblp", 3), ("dblp", 4)])
def test_tokenize_matches_reference_tokenizer(oracle_mod, tmp_path, source, q):
    L = _ref_tok()
    text = open(golden_path("small.txt"), "rb").read() if source == "small" else dblp_text()
    path = tmp_path / "in.txt"
    path.write_bytes(text)
    hp, lp, nr = C.POINTER(C.c_uint64)(), C.POINTER(C.c_int64)(), C.c_int64()
    n = L.ref_qgram_hashes(str(path).encode(), q, C.byref(hp), C.byref(lp), C.byref(nr))
    hashes = np.ctypeslib.as_array(hp, (n,)).copy() if n else np.zeros(0, np.uint64)
    lens = np.ctypeslib.as_array(lp, (nr.value,)).copy()
    L.ref_free(C.cast(hp, C.c_void_p))
    L.ref_free(C.cast(lp, C.c_void_p))
    want = _rank_rule(hashes, lens)
    b, _ = oracle_mod.tokenize(text, q)
    assert np.array_equal(b, want)


def test_bounds_known_values(oracle_mod):
    f32 = np.float32
    assert oracle_mod.bounds("max_prefix", 24, f32(0.75)) == 7
    assert oracle_mod.bounds("mid_prefix", 18, f32(0.75)) == 3
    assert f32(oracle_mod.bounds("min_overlap", 24, 18, f32(0.75))) == f32(18.0)
    assert np.floor(f32(oracle_mod.bounds("max_size", 9, f32(0.6)))) == 14
    assert f32(oracle_mod.bounds("min_overlap", 9, 9, f32(0.8))) == f32(8.000001)
    assert f32(oracle_mod.bounds("similarity", 24, 18, 18)) == f32(0.75)


def _grid():
    ts = [np.float32(x) for x in np.arange(0.05, 1.0, 0.05)] + [np.float32(0.8), np.float32(0.9),
                                                               np.float32(0.333), np.float32(0.999)]
    ls = list(range(1, 300)) + list(range(300, 5001, 37))
    return ts, ls


def test_bounds_match_numpy_float32(oracle_mod):
    ts, ls = _grid()
    for t in ts:
        r = np.float32(t / (np.float32(1) + t))
        for l in ls:
            l32 = np.float32(l)
            assert np.float32(oracle_mod.bounds("min_size", l, t)) == l32 * t
            assert np.float32(oracle_mod.bounds("max_size", l, t)) == l32 / t
            assert np.float32(oracle_mod.bounds("min_overlap", l, l, t)) == r * np.float32(2 * l)
            assert oracle_mod.bounds("max_prefix", l, t) == int(l32 - np.ceil(l32 * t) + np.float32(1))
            assert oracle_mod.bounds("mid_prefix", l, t) == int(l32 - np.ceil(r * np.float32(2 * l)) + np.float32(1))


def test_bounds_match_reference_header(oracle_mod):
    so = os.path.join(REF_DIR, "libref_bounds.so")
    if not os.path.exists(so):
        pytest.skip("oracle/_ref/libref_bounds.so not built (needs the reference sources)")
    R = C.CDLL(so)
    for n, res, args in [("min_size", C.c_float, [C.c_uint32, C.c_float]),
                         ("max_size", C.c_float, [C.c_uint32, C.c_float]),
                         ("min_overlap", C.c_float, [C.c_uint32, C.c_uint32, C.c_float]),
                         ("max_prefix", C.c_uint32, [C.c_uint32, C.c_float]),
                         ("mid_prefix", C.c_uint32, [C.c_uint32, C.c_float]),
                         ("similarity", C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32])]:
        getattr(R, "ref_" + n).restype = res
        getattr(R, "ref_" + n).argtypes = args
    ts, ls = _grid()
    for t in ts:
        for l in ls:
            for n in ("min_size", "max_size", "max_prefix", "mid_prefix"):
                assert getattr(R, "ref_" + n)(l, t) == oracle_mod.bounds(n, l, t), (n, l, t)
            for l2 in (1, l // 2 + 1, l, l + 7):
                assert R.ref_min_overlap(l, l2, t) == oracle_mod.bounds("min_overlap", l, l2, t)
                assert R.ref_similarity(l, l2, min(l, l2)) == oracle_mod.bounds("similarity", l, l2, min(l, l2))


def test_known_answers_fixture(oracle_mod):
    s = oracle_mod.Sets(_fixture_bin())
    ka = json.load(open(golden_path("small_known_answers.json")))
    for ts, lines in ka["answers"].items():
        t = oracle_mod.parse_threshold(ts)
        tri, _ = oracle_mod.selfjoin(s, t)
        got = ["%d %d %f" % (s.id[q], s.id[r], oracle_mod.similarity(int(s.len[q]), int(s.len[r]), int(o)))
               for q, r, o in tri]
        assert sorted(got) == sorted(lines), ts


def _titles(n, seed):
    import new_fgssjoin_amd as fg
    return fg.generate_text(fg.GEN_TITLES, n, seed)


@pytest.mark.parametrize("t", [0.4, 0.6, 0.8, 0.9])
def test_closed_form_equals_literal(oracle_mod, t):
    b, _ = oracle_mod.tokenize(_titles(400, 21), 3)
    s = oracle_mod.Sets(b)
    tri, ctr = oracle_mod.selfjoin(s, t)
    for B in (1, 7, s.n):
        lit, lc = oracle_mod.selfjoin(s, t, mode="literal", block_size=B)
        assert np.array_equal(lit, tri), B
        for k in ("e_idx", "p_q", "e_scan", "candidates", "intersections", "results", "m_reads"):
            assert lc[k] == ctr[k], (B, k)


@pytest.mark.parametrize("t", [0.5, 0.7, 0.9])
def test_any_record_order(oracle_mod, t):
    """A .bin in any record order (ppjoin_format, data.cu:31-126, reads records as they
    come; the earlier internal index is the query): the closed form equals the reference's
    tile loop at block size 1 -- whose tile skip (block.cu:121-130) then drops only what the
    4-way size filter drops -- on pairs and every counter, and lies within brute force."""
    from _util import shuffle_bin
    b, _ = oracle_mod.tokenize(_titles(300, 61), 3)
    s = oracle_mod.Sets(shuffle_bin(b, 5))
    assert np.any(np.diff(s.len) > 0)   # lengths really out of order
    tri, ctr = oracle_mod.selfjoin(s, t)
    lit, lc = oracle_mod.selfjoin(s, t, mode="literal", block_size=1)
    assert np.array_equal(lit, tri)
    for k in ("e_idx", "p_q", "e_scan", "candidates", "intersections", "results", "m_reads"):
        assert lc[k] == ctr[k], k
    br, _ = oracle_mod.selfjoin(s, t, mode="brute")
    assert {tuple(x) for x in tri.tolist()} <= {tuple(x) for x in br.tolist()}
    assert len(tri) > 0


def test_tile_skip_drops_pairs_of_unsorted_input(oracle_mod):
    """The reference's tile skip (block.cu:121-130) compares the last query's length with
    the first source's and assumes lengths non-increasing in internal order: on a .bin in
    another order a whole tile is skipped although it holds a pair.  Internal order
    {1,2}, {1,2,3}, {10..19} at t = 0.5: one block of 3 skips everything (10 > floor(2 / 0.5));
    block size 1, the closed form and this build find (0, 1) with overlap 2."""
    from _util import records_to_bin
    s = oracle_mod.Sets(records_to_bin([[1, 2], [1, 2, 3], list(range(10, 20))]))
    t = float(np.float32(0.5))
    tri, _ = oracle_mod.selfjoin(s, t)
    assert tri.tolist() == [[0, 1, 2]]
    assert oracle_mod.selfjoin(s, t, mode="literal", block_size=1)[0].tolist() == [[0, 1, 2]]
    assert oracle_mod.selfjoin(s, t, mode="literal", block_size=3)[0].shape[0] == 0


@pytest.mark.parametrize("t", [0.3, 0.5, 0.7, 0.8, 0.9])
def test_subset_of_brute_force(oracle_mod, t):
    b, _ = oracle_mod.tokenize(_titles(600, 33), 3)
    s = oracle_mod.Sets(b)
    tri, _ = oracle_mod.selfjoin(s, t)
    br, _ = oracle_mod.selfjoin(s, t, mode="brute")
    got = {tuple(x) for x in tri.tolist()}
    allp = {tuple(x) for x in br.tolist()}
    assert got <= allp
    for q, r, o in allp - got:   # only exact-boundary pairs may be missing
        a, c = int(s.len[q]), int(s.len[r])
        assert abs(o / (a + c - o) - float(t)) < 1e-6, (q, r, o)


def test_float_boundary_cases(oracle_mod):
    # |Q|=|S|=9, overlap 8, t=0.8: Jaccard = 0.8 exactly, but min_overlap(9,9) = 8.000001f
    # makes the mid prefix 1 token, so the prefixes {0,2} and {1} share nothing (m = 0)
    pos = np.array([1, 9, 0] + list(range(2, 10)) + [2, 9] + list(range(1, 10)), np.int32)
    # |Q|=15, |S|=9, S subset of Q, t=0.6: Jaccard 0.6 but floor(9/0.6f) = 14 < 15
    size = np.array([1, 9] + list(range(9)) + [2, 15] + list(range(15)), np.int32)
    for ints, t in ((pos, np.float32(0.8)), (size, np.float32(0.6))):
        s = oracle_mod.Sets(ints)
        tri, _ = oracle_mod.selfjoin(s, t)
        lit, _ = oracle_mod.selfjoin(s, t, mode="literal", block_size=1)
        br, _ = oracle_mod.selfjoin(s, t, mode="brute")
        assert tri.shape[0] == 0 and lit.shape[0] == 0 and br.shape[0] == 1


def test_empty_records_never_join(oracle_mod):
    ints = np.array([9, 0, 1, 2, 4, 5, 2, 2, 4, 5], np.int32)
    s = oracle_mod.Sets(ints)
    tri, _ = oracle_mod.selfjoin(s, np.float32(0.5))
    assert tri.tolist() == [[0, 1, 2]]


@pytest.mark.parametrize("t", [0.5, 0.8])
def test_query_subset_equals_full_join_restricted(oracle_mod, t):
    b, _ = oracle_mod.tokenize(_titles(800, 77), 3)
    s = oracle_mod.Sets(b)
    tri, ctr = oracle_mod.selfjoin(s, t)
    qs = np.random.default_rng(5).choice(s.n, 120, replace=False)
    sub, sc = oracle_mod.selfjoin(s, t, queries=qs)
    keep = np.isin(tri[:, 0], qs)
    assert np.array_equal(sub, tri[keep])
    # counters: per-query sums, from single-query runs
    tot = {k: 0 for k in ("p_q", "e_scan", "candidates", "intersections", "results", "m_reads")}
    for q in qs[:20]:
        _, c1 = oracle_mod.selfjoin(s, t, queries=[q])
        for k in tot:
            tot[k] += c1[k]
    _, c20 = oracle_mod.selfjoin(s, t, queries=qs[:20])
    for k in tot:
        assert c20[k] == tot[k], k
    assert sc["e_idx"] == ctr["e_idx"]
    with pytest.raises(ValueError):
        oracle_mod.selfjoin(s, t, queries=[s.n])


@pytest.mark.parametrize("t", [0.5, 0.8])
def test_single_merge_mode_equals_checked_join(oracle_mod, t):
    """checks=False (each candidate verified by the reference's merge alone, the CPU
    baseline's timed pass) gives the checked join's triples and counters, on the whole
    collection, on a query subset and on the literal tile loop's answer."""
    b, _ = oracle_mod.tokenize(_titles(900, 41), 3)
    s = oracle_mod.Sets(b)
    tri, ctr = oracle_mod.selfjoin(s, t)
    fast, fc = oracle_mod.selfjoin(s, t, checks=False)
    assert np.array_equal(fast, tri)
    lit, _ = oracle_mod.selfjoin(s, t, mode="literal", block_size=s.n)
    assert np.array_equal(fast, lit)
    assert fc["v_reads"] == 0
    for k in ("e_idx", "p_q", "e_scan", "m_reads", "candidates", "intersections", "results"):
        assert fc[k] == ctr[k], k
    qs = np.random.default_rng(9).choice(s.n, 150, replace=False)
    sub, sc = oracle_mod.selfjoin(s, t, queries=qs)
    fsub, fsc = oracle_mod.selfjoin(s, t, queries=qs, checks=False)
    assert np.array_equal(fsub, sub)
    for k in ("p_q", "e_scan", "m_reads", "candidates", "intersections", "results"):
        assert fsc[k] == sc[k], k
