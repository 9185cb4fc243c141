"""GPU: fgss_selfjoin_noalloc (no allocation, static device scratch) equals fgss_selfjoin.

The records are copied to device memory by the test (through the HIP runtime the
library itself links), then joined by the no-allocation entry point; its triples,
sorted, must equal the allocating entry point's pairs and counters bit for bit.
Two calls in a row share the scratch: the second, on a smaller collection, must
not see anything the first left there.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hip():
    """The HIP runtime libfgssjoin.so is linked against (already in this process)."""
    import new_fgssjoin_amd as fg
    fg.lib()
    path = None
    for ln in open("/proc/self/maps"):
        if "libamdhip64.so" in ln:
            path = ln.split()[-1]
            break
    assert path, "libamdhip64 not mapped"
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


class _Dev:
    """A device copy of a numpy array (freed with the object)."""
    def __init__(self, hip, a=None, nbytes=0):
        self.hip = hip
        self.n = a.nbytes if a is not None else nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.n, 16)) == 0
        if a is not None and a.nbytes:
            assert hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0   # host to device

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        if count:
            assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0   # device to host
        return out

    def __del__(self):
        self.hip.hipFree(self.p)


def _noalloc_join(fg, hip, coll, t, stream, out_cap=None, **kw):
    ids, lens, pos, tok = coll.arrays()
    inf = coll.info
    dt, dl, dp = _Dev(hip, tok.astype(np.int32)), _Dev(hip, lens.astype(np.int32)), _Dev(hip, pos.astype(np.int32))
    cap = out_cap if out_cap is not None else 1 << 20
    dout = _Dev(hip, nbytes=12 * cap)
    hip.hipMemset(dout.p, 0xFF, 12 * cap)
    n, st = fg.selfjoin_noalloc(dt.p.value, dl.p.value, dp.p.value, inf.n_sets, inf.n_tokens, inf.n_terms, t,
                                dout.p.value, cap, stream=stream, **kw)
    tri = dout.get(np.int32, 3 * n).reshape(-1, 3)
    order = np.lexsort((tri[:, 1], tri[:, 0])) if n else np.zeros(0, np.int64)
    return tri[order], st


def _equal(fg, coll, t, tri, st, **kw):
    p = coll.selfjoin(t, **kw)
    want = np.stack([p.q, p.s, p.overlap], axis=1) if len(p) else np.zeros((0, 3), np.int32)
    assert tri.shape == want.shape, f"{tri.shape[0]} triples vs {want.shape[0]}"
    assert np.array_equal(tri, want)
    for k in ("index_entries", "probe_tokens", "postings_scanned", "candidates", "intersections", "results",
              "verify_tokens", "signature_checks", "verified"):
        assert st[k] == p.stats[k], k


def test_noalloc_equals_selfjoin_twice(fg):
    hip = _hip()
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    try:
        big = fg.Collection.generate(fg.GEN_TITLES, 6000, 31)
        small = fg.Collection.generate(fg.GEN_TITLES, 1500, 32)
        for coll, t in ((big, 0.6), (small, 0.8)):
            _say(f"noalloc titles n={coll.info.n_sets} t={t}")
            tri, st = _noalloc_join(fg, hip, coll, t, s.value)
            _equal(fg, coll, t, tri, st)
        _say("noalloc small table (overflow re-split rounds)")
        tri, st = _noalloc_join(fg, hip, big, 0.5, s.value, table_slots=256)
        _equal(fg, big, 0.5, tri, st, table_slots=256)
    finally:
        hip.hipStreamDestroy(s)


def test_noalloc_null_stream_long_records_and_small_output(fg):
    hip = _hip()
    coll = fg.Collection.generate(fg.GEN_LONG, 600, 33)
    _say("noalloc long records n=600 t=0.5, null stream")
    tri, st = _noalloc_join(fg, hip, coll, 0.5, 0)
    _equal(fg, coll, 0.5, tri, st)
    if len(tri) > 1:
        with pytest.raises(fg.FgssError) as e:
            _noalloc_join(fg, hip, coll, 0.5, 0, out_cap=len(tri) - 1)
        assert e.value.status == fg.FGSS_ERR_ARG
    # invalid records are refused by the device-side check, before the join
    bad = fg.Collection.from_bin_ints(np.array([1, 2, 3, 1, 2, 2, 2, 3], np.int32))   # tokens 3, 1: not ascending
    ids, lens, pos, tok = bad.arrays()
    dt, dl, dp = _Dev(hip, tok), _Dev(hip, lens), _Dev(hip, pos.astype(np.int32))
    with pytest.raises(fg.FgssError) as e:
        fg.selfjoin_noalloc(dt.p.value, dl.p.value, dp.p.value, 2, 4, 4, 0.5, 0, 0)
    assert e.value.status == fg.FGSS_ERR_UNSUPPORTED


def test_noalloc_any_record_order(fg):
    """records whose lengths are out of order: the device-side check lets them through and
    the join runs each query against every later record, as fgss_selfjoin does"""
    from _util import shuffle_bin
    hip = _hip()
    coll = fg.Collection.from_bin_ints(shuffle_bin(fg.Collection.generate(fg.GEN_TITLES, 3000, 34).bin_ints(), 9))
    assert coll.info.sorted_desc == 0
    _say("noalloc shuffled titles n=3000 t=0.6")
    tri, st = _noalloc_join(fg, hip, coll, 0.6, 0)
    _equal(fg, coll, 0.6, tri, st)
