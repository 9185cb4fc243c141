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


def filter_counts(coll, t):
    """This build's candidate filters restated on the host for a collection in size order
    (lengths non-increasing): (candidates, positional-filter survivors, survivors whose
    64-bucket counts pass the overlap bound).  The survivors are the candidates whose
    count plus the tokens left after the last prefix match can still reach floor(bound);
    the bound (|Q| + |S| - SAD) / 2 over 64 byte counts, bucket (token x 0x9E3779B1) >> 26
    of 32 bits, is skipped (the candidate kept) when a count of either record exceeds 255.
    The join reports them as candidates, signature_checks and signature_checks_128."""
    _, lens, pos, tok = coll.arrays()
    lens = lens.astype(np.int64)
    pos = pos.astype(np.int64)
    tok = tok.astype(np.int64)
    n = len(lens)
    f32 = np.float32
    t = f32(t)
    r = t / (f32(1) + t)
    lf = lens.astype(f32)
    max_pre = np.minimum((lf - np.ceil(lf * t) + f32(1)).astype(np.int64), lens)
    mid_pre = np.minimum((lf - np.ceil(r * (lf + lf)) + f32(1)).astype(np.int64), lens)
    fmin = np.floor(lf * t).astype(np.int64)
    fmax = np.floor(lf / t).astype(np.int64)
    # 64-bucket counts per record
    b64 = ((tok * 0x9E3779B1) & 0xFFFFFFFF) >> 26
    rec = np.repeat(np.arange(n), lens)
    sig = np.zeros((n, 64), np.int64)
    np.add.at(sig, (rec, b64), 1)
    bad = (sig > 255).any(axis=1)
    # index of the mid-prefix tokens: (token, record, position), grouped by token
    src = np.repeat(np.arange(n), mid_pre)
    off = np.arange(len(src)) - np.repeat(np.cumsum(mid_pre) - mid_pre, mid_pre)
    ptok = tok[pos[src] + off]
    order = np.lexsort((src, ptok))
    ptok, psrc, ppos = ptok[order], src[order], off[order]
    starts = np.searchsorted(ptok, np.arange(int(tok.max(initial=0)) + 2))
    cand = keep = pass64 = 0
    for q in range(n):
        lq = int(lens[q])
        qt = tok[pos[q]:pos[q] + max_pre[q]]
        if len(qt) == 0:
            continue
        a, b = starts[qt], starts[qt + 1]
        cnt = b - a
        if cnt.sum() == 0:
            continue
        idx = np.repeat(a - np.cumsum(cnt) + cnt, cnt) + np.arange(cnt.sum())
        js = np.repeat(np.arange(len(qt)), cnt)
        s, ss = psrc[idx], ppos[idx]
        ls = lens[s]
        ok = (s > q) & ~((ls < fmin[q]) | (ls > fmax[q]) | (lq < fmin[s]) | (lq > fmax[s]))
        s, ss, js = s[ok], ss[ok], js[ok]
        if len(s) == 0:
            continue
        o = np.lexsort((js, s))   # per source, its last match last
        s, ss, js = s[o], ss[o], js[o]
        last = np.r_[s[1:] != s[:-1], True]
        m = np.bincount(np.searchsorted(np.unique(s), s))
        s, ss, js = s[last], ss[last], js[last]
        ls = lens[s]
        bound = r * (f32(lq) + ls.astype(f32)).astype(f32)
        x = m + np.minimum(lq - js, ls - ss)
        ps = ~(x.astype(f32) < bound)
        kp = ps & (x - 1 >= np.floor(bound).astype(np.int64))
        cand += int(ps.sum())
        keep += int(kp.sum())
        sad = np.abs((sig[s[kp]] & 255) - (sig[q] & 255)).sum(axis=1)
        fb = np.floor(bound[kp]).astype(np.int64)
        good = ~(bad[q] | bad[s[kp]])
        pass64 += int((good & ((lq + ls[kp] - sad) // 2 >= fb)).sum())
    return cand, keep, pass64
