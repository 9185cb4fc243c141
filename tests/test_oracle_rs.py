"""Pinning the oracle's R-S join (orc_rsjoin) before it checks the served joins.

The reference is a self-join only, so the R-S rules are pinned by what they must
share with it and by what mathematics fixes:
* every result is a pair with float32 Jaccard >= t (brute force), and every pair
  brute force finds and the closed form misses sits on a float boundary;
* R-S of a collection with itself has the self-join as its q<s half, exactly (the
  longer record of a pair is its query, as in the self-join) -- so on the reference
  fixture it reproduces the fixture's known answers -- holds every non-empty self
  pair, and is symmetric away from float boundaries;
* the index threshold t_index <= t changes no result, only the postings scanned.
"""
import json

import numpy as np
import pytest

from _util import golden_path, records_to_bin


def _fixture_sets(oracle_mod):
    return oracle_mod.Sets(np.fromfile(golden_path("small.txt.3gram.bin"), dtype="<i4"))


def _titles_sets(oracle_mod, n, seed):
    import new_fgssjoin_amd as fg
    b, _ = oracle_mod.tokenize(fg.generate_text(fg.GEN_TITLES, n, seed), 3)
    return oracle_mod.Sets(b)


def _perturbed_probes(oracle_mod, sets, n, seed):
    """Edited copies of indexed records (drop/add tokens, some unseen by the index)
    plus random records, as an oracle Sets whose internal index k is record k."""
    rng = np.random.default_rng(seed)
    recs = []
    for _ in range(n):
        if rng.random() < 0.8 and sets.n:
            r = set(sets.record(int(rng.integers(sets.n))).tolist())
            for _ in range(int(rng.integers(0, 4))):
                if r and rng.random() < 0.5:
                    r.discard(rng.choice(sorted(r)))
                else:
                    r.add(int(rng.integers(0, sets.n_terms + 40)))
        else:
            r = set(rng.integers(0, sets.n_terms + 40, int(rng.integers(0, 30))).tolist())
        recs.append(np.array(sorted(r), np.int32))
    return oracle_mod.Sets(records_to_bin(recs))


def _boundary(o, a, c, t):
    return abs(o / (a + c - o) - float(t)) < 1e-6


@pytest.mark.parametrize("t", [0.5, 0.7, 0.8, 0.9])
def test_rs_subset_of_brute_force(oracle_mod, t):
    s = _titles_sets(oracle_mod, 500, 41)
    p = _perturbed_probes(oracle_mod, s, 300, 5)
    t = float(np.float32(t))
    tri, ctr = oracle_mod.rsjoin(p, s, t)
    br, _ = oracle_mod.rsjoin(p, s, t, mode="brute")
    got, allp = {tuple(x) for x in tri.tolist()}, {tuple(x) for x in br.tolist()}
    assert got <= allp
    assert len(got) > 100
    for r, q, o in allp - got:
        assert _boundary(o, int(p.len[r]), int(s.len[q]), t), (r, q, o)
    assert ctr["results"] == len(got)
    assert ctr["candidates"] >= ctr["intersections"] >= ctr["results"]


@pytest.mark.parametrize("t", [0.4, 0.6, 0.8, 0.95])
def test_rs_with_itself_is_the_selfjoin(oracle_mod, t):
    s = _titles_sets(oracle_mod, 700, 43)
    t = float(np.float32(t))
    tri, _ = oracle_mod.rsjoin(s, s, t)
    got = {tuple(x) for x in tri.tolist()}
    upper = {x for x in got if x[0] < x[1]}
    sj, _ = oracle_mod.selfjoin(s, t)
    assert upper == {tuple(x) for x in sj.tolist()}
    assert {(i, i, int(s.len[i])) for i in range(s.n) if s.len[i] > 0} <= got
    for q, r, o in got ^ {(b, a, o) for a, b, o in got}:              # symmetric but at boundaries
        assert _boundary(o, int(s.len[q]), int(s.len[r]), t), (q, r, o)


def test_rs_known_answers_fixture(oracle_mod):
    s = _fixture_sets(oracle_mod)
    ka = json.load(open(golden_path("small_known_answers.json")))
    for ts, lines in ka["answers"].items():
        t = oracle_mod.parse_threshold(ts)
        tri, _ = oracle_mod.rsjoin(s, s, t)
        got = ["%d %d %f" % (s.id[q], s.id[r], oracle_mod.similarity(int(s.len[q]), int(s.len[r]), int(o)))
               for q, r, o in tri if q < r]
        assert sorted(got) == sorted(lines), ts


@pytest.mark.parametrize("t", [0.6, 0.85])
def test_rs_index_threshold_changes_no_result(oracle_mod, t):
    s = _titles_sets(oracle_mod, 400, 47)
    p = _perturbed_probes(oracle_mod, s, 200, 9)
    t = float(np.float32(t))
    base, c0 = oracle_mod.rsjoin(p, s, t)
    for ti in (0.3, 0.5, t):
        tri, c = oracle_mod.rsjoin(p, s, t, t_index=float(np.float32(ti)))
        assert np.array_equal(tri, base), ti
        for k in ("candidates", "intersections", "results", "p_q"):
            assert c[k] == c0[k], (ti, k)
        assert c["e_scan"] >= c0["e_scan"] and c["e_idx"] >= c0["e_idx"]
    with pytest.raises(ValueError):
        oracle_mod.rsjoin(p, s, t, t_index=float(np.float32(t + 0.01)))


def test_rs_hand_cases(oracle_mod):
    # indexed records (internal order): {0..9}, {0..4}, {}, {20..23}
    s = oracle_mod.Sets(records_to_bin([np.arange(10), np.arange(5), np.zeros(0), np.arange(20, 24)]))
    # probes: a superset (longer than every source), a subset, an empty probe, one
    # with tokens the index never saw, and {20..23} plus unseen tokens
    p = oracle_mod.Sets(records_to_bin([np.arange(12), np.arange(4), np.zeros(0), np.array([50, 51]),
                                        np.array([20, 21, 22, 23, 60])]))
    t = float(np.float32(0.8))
    tri, _ = oracle_mod.rsjoin(p, s, t)
    # 10/12 = 0.833 >= 0.8; 4/5 = 0.8 (float32 0.8 == 0.8f); 4/5 for the last probe
    assert tri.tolist() == [[0, 0, 10], [1, 1, 4], [4, 3, 4]]
