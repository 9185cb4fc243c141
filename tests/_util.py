"""Shared helpers for the parity tests (product vs oracle on identical inputs)."""
from __future__ import annotations

import gzip
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_path(name: str) -> str:
    return os.path.join(GOLDEN, name)


def dblp_text() -> bytes:
    with gzip.open(golden_path("dblp_t_18k.txt.gz"), "rb") as f:
        return f.read()


def oracle_sets(oracle, coll):
    """Oracle view of a product collection (same internal-order arrays)."""
    ids, lens, pos, tok = coll.arrays()
    return oracle.Sets.from_arrays(ids, lens, pos, tok, coll.info.n_terms)


def assert_join_equal(oracle, sets, pairs, t, counters=None):
    """GPU pairs == oracle triples, bit for bit, with the printed fields and stats."""
    tri, ctr = oracle.selfjoin(sets, t) if counters is None else counters
    got = np.stack([pairs.q, pairs.s, pairs.overlap], axis=1) if len(pairs) else np.zeros((0, 3), np.int32)
    assert got.shape == tri.shape, f"t={t}: {got.shape[0]} pairs vs oracle {tri.shape[0]}"
    assert np.array_equal(got, tri), f"t={t}: pair sets differ"
    if len(pairs):
        assert np.array_equal(pairs.rid, sets.id[tri[:, 0]])
        assert np.array_equal(pairs.sid, sets.id[tri[:, 1]])
        sims = np.array([oracle.similarity(int(sets.len[q]), int(sets.len[s]), int(o)) for q, s, o in tri],
                        np.float32)
        assert np.array_equal(pairs.similarity.view(np.uint32), sims.view(np.uint32))
    st = pairs.stats
    assert st["index_entries"] == ctr["e_idx"]
    assert st["probe_tokens"] == ctr["p_q"]
    assert st["postings_scanned"] == ctr["e_scan"]
    assert st["candidates"] == ctr["candidates"]
    assert st["intersections"] == ctr["intersections"]
    assert st["results"] == ctr["results"]
    return tri, ctr


def records_to_bin(records, ids=None) -> np.ndarray:
    """A .bin int32 stream whose internal order (file order reversed) is `records`
    in list order; record k gets id ids[k] (default k + 1)."""
    out = []
    n = len(records)
    for k in range(n - 1, -1, -1):
        r = np.asarray(records[k], np.int32)
        out += [int(ids[k]) if ids is not None else k + 1, int(r.size)] + r.tolist()
    return np.array(out, np.int32)


def bin_records(bin_ints):
    """(id, tokens) per record of a .bin int32 stream, in file order."""
    v = np.asarray(bin_ints, np.int32)
    out, i = [], 0
    while i < v.size:
        n = int(v[i + 1])
        out.append((int(v[i]), v[i + 2:i + 2 + n].copy()))
        i += 2 + n
    return out


def shuffle_bin(bin_ints, seed: int) -> np.ndarray:
    """The same records as a .bin stream in a seeded random file order (any order is a
    valid .bin: ppjoin_format reads records as they come, src/fgssjoin/data.cu:31-126)."""
    recs = bin_records(bin_ints)
    order = np.random.default_rng(seed).permutation(len(recs))
    out = []
    for k in order:
        rid, tok = recs[int(k)]
        out += [rid, int(tok.size)] + tok.tolist()
    return np.array(out, np.int32)


def hip_runtime():
    """The HIP runtime that libfgssjoin.so links (already mapped in this process), for
    tests that create their own streams or device buffers."""
    import ctypes as C

    import new_fgssjoin_amd as fg
    fg.lib()
    path = next((ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64.so" in ln), None)
    assert path, "libamdhip64 not mapped"
    h = C.CDLL(path)
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    return h
