"""ctypes wrapper of the CPU oracle (oracle/fgss_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's ``cpu_baseline`` leg.  The product package
``new_fgssjoin_amd`` never imports it.

The library is built by ``oracle/Makefile`` into ``oracle/build/liboracle.so``
(``build()`` in ``__graft_entry__`` runs it).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


class Counters(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "e_idx", "p_q", "e_scan", "m_reads", "v_reads",
        "candidates", "intersections", "results")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile the oracle with make (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        i32p, i64p, f = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.c_float
        L.orc_min_size.restype = f; L.orc_min_size.argtypes = [C.c_uint32, f]
        L.orc_max_size.restype = f; L.orc_max_size.argtypes = [C.c_uint32, f]
        L.orc_min_overlap.restype = f; L.orc_min_overlap.argtypes = [C.c_uint32, C.c_uint32, f]
        L.orc_max_prefix.restype = C.c_uint32; L.orc_max_prefix.argtypes = [C.c_uint32, f]
        L.orc_mid_prefix.restype = C.c_uint32; L.orc_mid_prefix.argtypes = [C.c_uint32, f]
        L.orc_similarity.restype = f; L.orc_similarity.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_parse_threshold.restype = f; L.orc_parse_threshold.argtypes = [C.c_char_p]
        L.orc_tokenize.restype = C.c_int
        L.orc_tokenize.argtypes = [C.c_char_p, C.c_int64, C.c_int, C.POINTER(i32p), i64p,
                                   C.POINTER(i32p), i64p]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_bin_count.restype = C.c_int
        L.orc_bin_count.argtypes = [i32p, C.c_int64, i64p, i64p, i32p]
        L.orc_bin_load.restype = C.c_int
        L.orc_bin_load.argtypes = [i32p, C.c_int64, i32p, i32p, i64p, i32p]
        L.orc_selfjoin.restype = C.c_int64
        L.orc_selfjoin.argtypes = [C.c_int64, i32p, i64p, i32p, C.c_int32, f, C.c_int,
                                   C.POINTER(i32p), C.POINTER(Counters)]
        L.orc_selfjoin_queries.restype = C.c_int64
        L.orc_selfjoin_queries.argtypes = [C.c_int64, i32p, i64p, i32p, C.c_int32, f, i64p, C.c_int64, C.c_int,
                                           C.POINTER(i32p), C.POINTER(Counters)]
        L.orc_selfjoin_ex.restype = C.c_int64
        L.orc_selfjoin_ex.argtypes = [C.c_int64, i32p, i64p, i32p, C.c_int32, f, C.c_int, i64p, C.c_int64, C.c_int,
                                      C.POINTER(i32p), C.POINTER(Counters), C.POINTER(C.c_double)]
        L.orc_selfjoin_literal.restype = C.c_int64
        L.orc_selfjoin_literal.argtypes = [C.c_int64, i32p, i64p, i32p, C.c_int32, f, C.c_int64,
                                           C.POINTER(i32p), C.POINTER(Counters)]
        L.orc_selfjoin_brute.restype = C.c_int64
        L.orc_selfjoin_brute.argtypes = [C.c_int64, i32p, i64p, i32p, f, C.c_int, C.POINTER(i32p)]
        L.orc_rsjoin.restype = C.c_int64
        L.orc_rsjoin.argtypes = [C.c_int64, i32p, i64p, i32p, C.c_int64, i32p, i64p, i32p, C.c_int32, f, f,
                                 C.c_int, C.POINTER(i32p), C.POINTER(Counters)]
        L.orc_rsjoin_brute.restype = C.c_int64
        L.orc_rsjoin_brute.argtypes = [C.c_int64, i32p, i64p, i32p, C.c_int64, i32p, i64p, i32p, f,
                                       C.POINTER(i32p)]
        _lib = L
    return _lib


def _p32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _take(ptr, n):
    arr = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0, np.int32)
    lib().orc_free(C.cast(ptr, C.c_void_p))
    return arr


def parse_threshold(s: str) -> float:
    """Reference threshold parse (atof -> float32), returned as a Python float."""
    return float(lib().orc_parse_threshold(s.encode()))


def bounds(fn: str, *args):
    return getattr(lib(), "orc_" + fn)(*args)


def tokenize(text: bytes, q: int):
    """-> (bin int32 array, loc int32 array)"""
    L = lib()
    b, l = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    nb, nl = C.c_int64(), C.c_int64()
    rc = L.orc_tokenize(text, len(text), q, C.byref(b), C.byref(nb), C.byref(l), C.byref(nl))
    if rc != 0:
        raise ValueError("orc_tokenize failed")
    return _take(b, nb.value).astype(np.int32), _take(l, nl.value).astype(np.int32)


class Sets:
    """A collection in the reference's internal order (file order reversed)."""

    @classmethod
    def from_arrays(cls, ids, lens, pos, tok, n_terms):
        """Wrap internal-order arrays (e.g. the product's) without re-reading a .bin."""
        s = cls.__new__(cls)
        s.id = np.ascontiguousarray(ids, np.int32)
        s.len = np.ascontiguousarray(lens, np.int32)
        s.pos = np.ascontiguousarray(pos, np.int64)
        s.tok = np.ascontiguousarray(tok if len(tok) else np.zeros(1, np.int32), np.int32)
        s.n, s.n_tokens, s.n_terms = int(s.len.size), int(np.sum(s.len)), int(n_terms)
        return s

    def __init__(self, bin_ints):
        bin_ints = np.ascontiguousarray(bin_ints, dtype=np.int32)
        L = lib()
        ns, nt, terms = C.c_int64(), C.c_int64(), C.c_int32()
        if L.orc_bin_count(_p32(bin_ints), bin_ints.size, C.byref(ns), C.byref(nt), C.byref(terms)) != 0:
            raise ValueError("Reading error.")
        self.n, self.n_tokens, self.n_terms = ns.value, nt.value, terms.value
        self.id = np.zeros(self.n, np.int32)
        self.len = np.zeros(self.n, np.int32)
        self.pos = np.zeros(self.n, np.int64)
        self.tok = np.zeros(max(self.n_tokens, 1), np.int32)
        L.orc_bin_load(_p32(bin_ints), bin_ints.size, _p32(self.id), _p32(self.len), _p64(self.pos), _p32(self.tok))

    def record(self, i):
        return self.tok[self.pos[i]:self.pos[i] + self.len[i]]


def selfjoin(sets: Sets, t: float, nthreads: int = 0, mode: str = "closed", block_size: int = 0, queries=None,
             checks: bool = True, times: list | None = None):
    """-> (triples int32 [R,3] of (q, s, overlap) internal indices sorted, counters dict).
    queries: restrict the closed form to these query indices.  checks=False verifies
    each candidate with the reference's merge alone (no closed-form cross-check; the
    CPU baseline's timed pass); then `times`, if a list, receives the seconds of the
    index build and of the query phase."""
    L = lib()
    if nthreads <= 0:
        nthreads = os.cpu_count() or 1
    out = C.POINTER(C.c_int32)()
    ctr = Counters()
    if not checks and mode == "closed":
        qs = np.ascontiguousarray(queries if queries is not None else np.zeros(0), np.int64)
        secs = (C.c_double * 2)()
        n = L.orc_selfjoin_ex(sets.n, _p32(sets.len), _p64(sets.pos), _p32(sets.tok), sets.n_terms, t, nthreads,
                              _p64(qs) if queries is not None else None, qs.size, 0, C.byref(out), C.byref(ctr),
                              secs)
        if times is not None:
            times[:] = [secs[0], secs[1]]
    elif queries is not None:
        qs = np.ascontiguousarray(queries, np.int64)
        n = L.orc_selfjoin_queries(sets.n, _p32(sets.len), _p64(sets.pos), _p32(sets.tok), sets.n_terms, t,
                                   _p64(qs), qs.size, nthreads, C.byref(out), C.byref(ctr))
    elif mode == "closed":
        n = L.orc_selfjoin(sets.n, _p32(sets.len), _p64(sets.pos), _p32(sets.tok), sets.n_terms,
                           t, nthreads, C.byref(out), C.byref(ctr))
    elif mode == "literal":
        n = L.orc_selfjoin_literal(sets.n, _p32(sets.len), _p64(sets.pos), _p32(sets.tok), sets.n_terms,
                                   t, block_size or sets.n, C.byref(out), C.byref(ctr))
    elif mode == "brute":
        n = L.orc_selfjoin_brute(sets.n, _p32(sets.len), _p64(sets.pos), _p32(sets.tok), t, nthreads,
                                 C.byref(out))
    else:
        raise ValueError(mode)
    if n < 0:
        raise ValueError("oracle join failed")
    tri = _take(out, 3 * n).reshape(-1, 3).astype(np.int32)
    return tri, ctr.as_dict()


def rsjoin(probes: Sets, sets: Sets, t: float, t_index: float | None = None, nthreads: int = 0,
           mode: str = "closed"):
    """R-S join -> (triples int32 [R,3] of (probe, source, overlap) sorted, counters dict)"""
    L = lib()
    if nthreads <= 0:
        nthreads = os.cpu_count() or 1
    out = C.POINTER(C.c_int32)()
    ctr = Counters()
    args = (probes.n, _p32(probes.len), _p64(probes.pos), _p32(probes.tok),
            sets.n, _p32(sets.len), _p64(sets.pos), _p32(sets.tok))
    if mode == "closed":
        n = L.orc_rsjoin(*args, sets.n_terms, t, t if t_index is None else t_index, nthreads, C.byref(out),
                         C.byref(ctr))
    elif mode == "brute":
        n = L.orc_rsjoin_brute(*args, t, C.byref(out))
    else:
        raise ValueError(mode)
    if n < 0:
        raise ValueError("oracle R-S join failed")
    return _take(out, 3 * n).reshape(-1, 3).astype(np.int32), ctr.as_dict()


def similarity(a: int, b: int, inter: int) -> float:
    return float(lib().orc_similarity(a, b, inter))
