"""GPU parity on .bin files in any record order (ppjoin_format reads records as they come,
src/fgssjoin/data.cu:31-126; the earlier internal index of a pair is its query).

The join of a job whose lengths are not non-increasing runs every query against every
later record of the job, each posting through the 4-way size filter (filtering.cu:88-120).
Every case equals the oracle's closed form -- pinned to the reference's tile loop at block
size 1 on shuffled input by tests/test_oracle.py -- pair for pair and counter for counter;
the small case also equals that tile loop here.
"""
import sys

import numpy as np
import pytest

from _util import assert_join_equal, dblp_text, oracle_sets, shuffle_bin

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _shuffled(fg, coll, seed):
    c = fg.Collection.from_bin_ints(shuffle_bin(coll.bin_ints(), seed))
    assert c.info.sorted_desc == 0
    return c


@pytest.mark.parametrize("t", [0.5, 0.7, 0.9])
def test_shuffled_titles(fg, oracle_mod, t):
    _say(f"shuffled titles n=20000 seed=41 t={t}")
    coll = _shuffled(fg, fg.Collection.generate(fg.GEN_TITLES, 20000, 41), 1)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t)


@pytest.mark.parametrize("opts", [dict(table_slots=256), dict(stream_parts=1), dict(bounds_cap=16),
                                  dict(table_slots=1024, load_pct=90)])
def test_shuffled_titles_every_item_mode(fg, oracle_mod, opts):
    """small tables (splits, overflow re-splits), streamed and searched parts"""
    t = 0.5
    _say(f"shuffled titles n=20000 seed=42 t={t} {opts}")
    coll = _shuffled(fg, fg.Collection.generate(fg.GEN_TITLES, 20000, 42), 2)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t, **opts), t)


@pytest.mark.parametrize("t", [0.6, 0.8])
def test_shuffled_dblp(fg, oracle_mod, t):
    _say(f"shuffled dblp 18k t={t}")
    coll = _shuffled(fg, fg.Collection.from_text_bytes(dblp_text(), 3), 3)
    sets = oracle_sets(oracle_mod, coll)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t)


def test_shuffled_long_records(fg, oracle_mod):
    t = 0.8
    _say(f"shuffled long records n=1500 seed=43 t={t}")
    coll = _shuffled(fg, fg.Collection.generate(fg.GEN_LONG, 1500, 43), 4)
    sets = oracle_sets(oracle_mod, coll)
    p = coll.selfjoin(t)
    assert_join_equal(oracle_mod, sets, p, t)
    assert p.stats["dense_parts"] == 0   # dense parts need a sorted job


@pytest.mark.parametrize("t", [0.5, 0.8])
def test_shuffled_small_equals_tile_loop(fg, oracle_mod, t):
    _say(f"shuffled titles n=300 seed=44 t={t} vs the tile loop")
    coll = _shuffled(fg, fg.Collection.generate(fg.GEN_TITLES, 300, 44), 5)
    sets = oracle_sets(oracle_mod, coll)
    lit = oracle_mod.selfjoin(sets, t, mode="literal", block_size=1)
    assert_join_equal(oracle_mod, sets, coll.selfjoin(t), t, counters=lit)


def test_batch_mixes_record_orders(fg, oracle_mod):
    """jobs of either order in one batch: each equals its own oracle join"""
    colls, ts = [], []
    for j in range(6):
        c = fg.Collection.generate(fg.GEN_TITLES, 1500 + 700 * j, 60 + j)
        colls.append(_shuffled(fg, c, 70 + j) if j % 2 else c)
        ts.append(float(np.float32((0.5, 0.6, 0.7, 0.8, 0.9, 0.55)[j])))
    _say("batch of 6 jobs, every other one shuffled")
    b = fg.Batch(colls, ts)
    b.run(graph=True)
    b.run(graph=True)   # a replay
    got = b.all_pairs()
    for j, c in enumerate(colls):
        assert_join_equal(oracle_mod, oracle_sets(oracle_mod, c), got[j], ts[j])
    b.close()


def test_server_refuses_an_unsorted_index(fg):
    coll = _shuffled(fg, fg.Collection.generate(fg.GEN_TITLES, 500, 45), 6)
    with pytest.raises(fg.FgssError) as e:
        fg.Server(coll, 0.7, slots=1)
    assert e.value.status == fg.FGSS_ERR_UNSUPPORTED
