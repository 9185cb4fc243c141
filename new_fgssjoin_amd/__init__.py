"""new_fgssjoin_amd -- MI355X-native fgssjoin (set-similarity self-join).

Python mirror of the reference's host interface over the C ABI in
``include/fgssjoin.h`` (library ``libfgssjoin.so``, built in-tree by
``new_fgssjoin_amd/build.py``).  The reference's host functions map as:

  ppjoin_format(path)            -> Collection.from_bin(path)      (src/fgssjoin/data.cu:31-126)
  text front end (-f, -q)        -> Collection.from_text(path, q)  (README.md:12-15)
  prepare_data(sets, t)          -> Collection.upload()            (src/fgssjoin/data.cu:133-199)
  inverted_index + process_blocks-> Collection.selfjoin(t)         (src/fgssjoin/index.cu:169-211,
                                                                    src/fgssjoin/block.cu:61-213)

Errors raise FgssError carrying the C status code; nothing here falls back to
a CPU implementation: every join runs the HIP kernels of libfgssjoin.so, and
the first call into the package raises ImportError when the library is missing.
Served R-S joins (``Server``) and batches (``Batch``) have no reference
counterpart.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfgssjoin.so")
CLI_PATH = os.path.join(_HERE, "fgssjoin")

FGSS_OK, FGSS_ERR_ARG, FGSS_ERR_IO, FGSS_ERR_FORMAT, FGSS_ERR_DEVICE, FGSS_ERR_NOMEM, FGSS_ERR_UNSUPPORTED = range(7)
GEN_TITLES, GEN_LONG = 0, 1

# every symbol declared in include/fgssjoin.h
EXPORTS = (
    "fgss_last_error", "fgss_version",
    "fgss_collection_from_bin", "fgss_collection_from_bin_buffer",
    "fgss_collection_from_text", "fgss_collection_from_text_buffer", "fgss_collection_from_text_dict",
    "fgss_collection_from_text_device",
    "fgss_collection_write_bin", "fgss_collection_get_info", "fgss_collection_arrays",
    "fgss_collection_upload", "fgss_collection_free", "fgss_device_synchronize",
    "fgss_selfjoin", "fgss_selfjoin_noalloc", "fgss_pairs_free", "fgss_pairs_print", "fgss_collection_print_index",
    "fgss_generate_text", "fgss_free",
    "fgss_batch_create", "fgss_batch_set_thresholds", "fgss_batch_run", "fgss_batch_pairs",
    "fgss_batch_pairs_all", "fgss_batch_autotune", "fgss_batch_autotune_cached", "fgss_batch_free",
    "fgss_server_create", "fgss_server_submit", "fgss_server_poll", "fgss_server_done", "fgss_server_wait",
    "fgss_server_result", "fgss_server_hold", "fgss_server_free",
)


class FgssError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


class CollectionInfo(C.Structure):
    _fields_ = [("n_sets", C.c_int64), ("n_tokens", C.c_int64), ("n_terms", C.c_int32),
                ("max_len", C.c_int32), ("average_size", C.c_double), ("sorted_desc", C.c_int32),
                ("on_device", C.c_int32)]


class JoinStats(C.Structure):
    _fields_ = [("index_entries", C.c_int64), ("probe_tokens", C.c_int64), ("postings_scanned", C.c_int64),
                ("candidates", C.c_int64), ("intersections", C.c_int64), ("results", C.c_int64),
                ("verify_tokens", C.c_int64), ("ms_index", C.c_double), ("ms_join", C.c_double),
                ("ms_copy", C.c_double), ("ms_total", C.c_double),
                ("ms_kernel", C.c_double), ("kernel_launches", C.c_int64),
                ("signature_checks", C.c_int64), ("signature_checks_128", C.c_int64), ("verified", C.c_int64),
                ("phase_cycles", C.c_int64 * 6), ("window_retries", C.c_int64), ("dense_parts", C.c_int64),
                ("dense_postings", C.c_int64)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["phase_cycles"] = list(self.phase_cycles)
        return d


class JoinOpts(C.Structure):
    _fields_ = [("load_pct", C.c_int32), ("table_slots", C.c_int32), ("stream_parts", C.c_int32),
                ("bounds_cap", C.c_int32), ("dense_pct", C.c_int32), ("query_begin", C.c_int64),
                ("query_end", C.c_int64), ("stream", C.c_void_p), ("index_groups", C.c_int32),
                ("block_records", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_pipeline", C.c_double), ("n_pairs", C.c_int64),
                ("candidates", C.c_int64), ("graph", C.c_int32), ("kernel_launches", C.c_int32)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class RequestStats(C.Structure):
    _fields_ = [("ms_queue", C.c_double), ("ms_latency", C.c_double), ("batch_requests", C.c_int32),
                ("reserved", C.c_int32), ("charged_bytes", C.c_int64), ("inflight_bytes", C.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


class _Pairs(C.Structure):
    _fields_ = [("n_pairs", C.c_int64), ("q", C.POINTER(C.c_int32)), ("s", C.POINTER(C.c_int32)),
                ("rid", C.POINTER(C.c_int32)), ("sid", C.POINTER(C.c_int32)),
                ("overlap", C.POINTER(C.c_int32)), ("similarity", C.POINTER(C.c_float)),
                ("stats", JoinStats)]


_lib = None


def lib():
    """Load libfgssjoin.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m new_fgssjoin_amd.build` "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32p, i64p = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    L.fgss_last_error.restype = C.c_char_p
    L.fgss_version.restype = C.c_char_p
    for name, args in {
        "fgss_collection_from_bin": [C.c_char_p, C.POINTER(vp)],
        "fgss_collection_from_bin_buffer": [i32p, C.c_int64, C.POINTER(vp)],
        "fgss_collection_from_text": [C.c_char_p, C.c_int, C.POINTER(vp)],
        "fgss_collection_from_text_buffer": [C.c_char_p, C.c_int64, C.c_int, C.POINTER(vp)],
        "fgss_collection_from_text_dict": [C.c_char_p, C.c_int64, vp, C.POINTER(vp)],
        "fgss_collection_from_text_device": [C.c_char_p, C.c_int64, C.c_int, C.POINTER(vp)],
        "fgss_collection_write_bin": [vp, C.c_char_p, C.c_char_p],
        "fgss_collection_get_info": [vp, C.POINTER(CollectionInfo)],
        "fgss_collection_arrays": [vp, C.POINTER(i32p), C.POINTER(i32p), C.POINTER(i64p), C.POINTER(i32p)],
        "fgss_collection_upload": [vp],
        "fgss_device_synchronize": [],
        "fgss_selfjoin": [vp, C.c_float, C.POINTER(JoinOpts), C.POINTER(C.POINTER(_Pairs))],
        "fgss_selfjoin_noalloc": [vp, vp, vp, C.c_int64, C.c_int64, C.c_int32, C.c_float, C.POINTER(JoinOpts), vp,
                                  C.c_int64, i64p, C.POINTER(JoinStats), vp],
        "fgss_pairs_print": [C.POINTER(_Pairs), vp],
        "fgss_collection_print_index": [vp, C.c_float, vp],
        "fgss_generate_text": [C.c_int, C.c_int64, C.c_uint64, C.POINTER(C.c_void_p), i64p],
        "fgss_batch_create": [C.POINTER(vp), C.POINTER(C.c_float), C.c_int32, C.POINTER(vp)],
        "fgss_batch_set_thresholds": [vp, C.POINTER(C.c_float)],
        "fgss_batch_run": [vp, C.POINTER(JoinOpts), C.c_int32, C.POINTER(BatchStats)],
        "fgss_batch_pairs": [vp, C.c_int32, C.POINTER(C.POINTER(_Pairs))],
        "fgss_batch_pairs_all": [vp, C.POINTER(C.POINTER(_Pairs)), C.POINTER(C.c_int64), C.POINTER(JoinStats)],
        "fgss_batch_autotune": [vp, C.c_int32, C.POINTER(JoinOpts), C.POINTER(C.c_double)],
        "fgss_batch_autotune_cached": [vp, C.c_int32, C.c_char_p, C.POINTER(JoinOpts), C.POINTER(C.c_double),
                                       i32p],
        "fgss_server_create": [vp, C.c_float, C.c_int32, C.c_int64, C.POINTER(JoinOpts), C.POINTER(vp)],
        "fgss_server_submit": [vp, vp, C.c_float, i64p],
        "fgss_server_poll": [vp, C.c_int32, i32p, i32p],
        "fgss_server_done": [vp, C.c_int64, i32p],
        "fgss_server_wait": [vp, C.c_int64],
        "fgss_server_hold": [vp, C.c_int32],
        "fgss_server_result": [vp, C.c_int64, C.POINTER(C.POINTER(_Pairs)), C.POINTER(RequestStats)],
    }.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.fgss_collection_free.argtypes = [vp]
    L.fgss_collection_free.restype = None
    L.fgss_pairs_free.argtypes = [C.POINTER(_Pairs)]
    L.fgss_pairs_free.restype = None
    L.fgss_free.argtypes = [vp]
    L.fgss_free.restype = None
    L.fgss_batch_free.argtypes = [vp]
    L.fgss_batch_free.restype = None
    L.fgss_server_free.argtypes = [vp]
    L.fgss_server_free.restype = None
    _lib = L
    return L


def _check(rc: int):
    if rc != FGSS_OK:
        raise FgssError(rc, lib().fgss_last_error().decode(errors="replace"))


def version() -> str:
    return lib().fgss_version().decode()


def device_synchronize() -> None:
    """Wait for all device work of this process."""
    _check(lib().fgss_device_synchronize())


def parse_threshold(s: str) -> float:
    """The CLI's threshold parse: atof to double, stored in a float (src/fgssjoin_cli.cu:31)."""
    return float(np.float32(float(s)))


def generate_text(kind: int, n_records: int, seed: int) -> bytes:
    """Seeded synthetic records, one per line (GEN_TITLES or GEN_LONG)."""
    L = lib()
    p, n = C.c_void_p(), C.c_int64()
    _check(L.fgss_generate_text(kind, n_records, seed, C.byref(p), C.byref(n)))
    try:
        return C.string_at(p, n.value)
    finally:
        L.fgss_free(p)


@dataclass
class Pairs:
    """Result pairs sorted by internal (q, s); rid/sid are the printed record ids."""
    q: np.ndarray
    s: np.ndarray
    rid: np.ndarray
    sid: np.ndarray
    overlap: np.ndarray
    similarity: np.ndarray
    stats: dict

    def __len__(self):
        return int(self.q.size)

    def lines(self) -> list[str]:
        """Reference stdout lines: "%d %d %f" (src/fgssjoin/block.cu:193)."""
        return ["%d %d %f" % (a, b, float(c)) for a, b, c in zip(self.rid, self.sid, self.similarity)]


class Collection:
    """A set collection in the reference's internal order (file order reversed)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def from_bin(cls, path: str) -> "Collection":
        h = C.c_void_p()
        _check(lib().fgss_collection_from_bin(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_bin_ints(cls, ints) -> "Collection":
        a = np.ascontiguousarray(ints, dtype=np.int32)
        h = C.c_void_p()
        _check(lib().fgss_collection_from_bin_buffer(a.ctypes.data_as(C.POINTER(C.c_int32)), a.size, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_text(cls, path: str, q: int) -> "Collection":
        h = C.c_void_p()
        _check(lib().fgss_collection_from_text(os.fsencode(path), q, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_text_bytes(cls, text: bytes, q: int, device: bool = False) -> "Collection":
        """Tokenize text (one record per non-empty line); device=True runs the
        tokenizer on the GPU (q <= 8), with the same result."""
        h = C.c_void_p()
        fn = lib().fgss_collection_from_text_device if device else lib().fgss_collection_from_text_buffer
        _check(fn(text, len(text), q, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_text_dict(cls, text: bytes, dictionary: "Collection") -> "Collection":
        """Tokenize text into the token space of `dictionary` (a collection built from
        text): probe records for Server.submit."""
        h = C.c_void_p()
        _check(lib().fgss_collection_from_text_dict(text, len(text), dictionary._h, C.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, kind: int, n_records: int, seed: int, q: int = 3, device: bool = False) -> "Collection":
        return cls.from_text_bytes(generate_text(kind, n_records, seed), q, device=device)

    def close(self):
        if self._h is not None and self._h.value:
            lib().fgss_collection_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def info(self) -> CollectionInfo:
        i = CollectionInfo()
        _check(lib().fgss_collection_get_info(self._h, C.byref(i)))
        return i

    def arrays(self):
        """(id, len, pos, tokens) numpy copies in internal order."""
        i32p, i64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64)
        idp, lp, pp, tp = i32p(), i32p(), i64p(), i32p()
        _check(lib().fgss_collection_arrays(self._h, C.byref(idp), C.byref(lp), C.byref(pp), C.byref(tp)))
        inf = self.info
        n, nt = inf.n_sets, inf.n_tokens

        def cp(p, k, dt):
            return np.ctypeslib.as_array(p, shape=(k,)).copy() if k > 0 else np.zeros(0, dt)
        return cp(idp, n, np.int32), cp(lp, n, np.int32), cp(pp, n, np.int64), cp(tp, nt, np.int32)

    def bin_ints(self) -> np.ndarray:
        """The collection as a reference .bin int32 stream (file order)."""
        ids, lens, pos, tok = self.arrays()
        parts = []
        for k in range(ids.size - 1, -1, -1):
            parts.append(np.array([ids[k], lens[k]], np.int32))
            parts.append(tok[pos[k]:pos[k] + lens[k]])
        return np.concatenate(parts) if parts else np.zeros(0, np.int32)

    def write_bin(self, path: str, loc_path: str | None = None):
        _check(lib().fgss_collection_write_bin(self._h, os.fsencode(path),
                                               os.fsencode(loc_path) if loc_path else None))

    def upload(self):
        _check(lib().fgss_collection_upload(self._h))

    def selfjoin(self, threshold: float, table_slots: int = 0, load_pct: int = 0, bounds_cap: int = 0,
                 stream_parts: int = 0, queries: tuple | None = None, stream: int = 0, dense_pct: int = 0,
                 block_records: int = 0) -> Pairs:
        """queries=(begin, end): only those queries (internal indices), each against every
        later record; stream: a hipStream_t (as int) for the call's device work;
        dense_pct, block_records: see fgss_join_opts (dense_pct -1: no dense parts)."""
        L = lib()
        qb, qe = queries if queries is not None else (0, 0)
        opts = JoinOpts(load_pct, table_slots, stream_parts, bounds_cap, dense_pct, qb, qe, stream or None, 0,
                        block_records)
        out = C.POINTER(_Pairs)()
        _check(L.fgss_selfjoin(self._h, C.c_float(threshold), C.byref(opts), C.byref(out)))
        return _take_pairs(out)


NOALLOC_MAX_SETS, NOALLOC_MAX_TOKENS, NOALLOC_MAX_TERMS = 65536, 524288, 524288


def selfjoin_noalloc(tokens: int, lengths: int, offsets: int, n_sets: int, n_tokens: int, n_terms: int,
                     threshold: float, out: int, out_cap: int, stream: int = 0, table_slots: int = 0,
                     load_pct: int = 0):
    """fgss_selfjoin_noalloc on device addresses (int): the records in device memory
    (int32 tokens, lengths, offsets in internal order) and a device output of out_cap
    (q, s, overlap) int32 triples.  Returns (number of triples written, stats dict);
    the triples are in no fixed order.  Nothing is allocated by the call."""
    n = C.c_int64()
    st = JoinStats()
    opts = JoinOpts(load_pct, table_slots, 0, 0)
    _check(lib().fgss_selfjoin_noalloc(C.c_void_p(tokens), C.c_void_p(lengths), C.c_void_p(offsets), n_sets,
                                       n_tokens, n_terms, C.c_float(threshold), C.byref(opts), C.c_void_p(out),
                                       out_cap, C.byref(n), C.byref(st), C.c_void_p(stream)))
    return n.value, st.as_dict()


def _take_pairs(out) -> Pairs:
    try:
        p = out.contents
        n = p.n_pairs

        def cp(ptr, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0, dt)
        return Pairs(cp(p.q, np.int32), cp(p.s, np.int32), cp(p.rid, np.int32), cp(p.sid, np.int32),
                     cp(p.overlap, np.int32), cp(p.similarity, np.float32), p.stats.as_dict())
    finally:
        lib().fgss_pairs_free(out)


class Batch:
    """Many independent self-joins (own collection, own threshold) run as one pipeline.

    The collections must stay alive while the batch is used (record ids come from them)."""

    def __init__(self, collections, thresholds):
        self.collections = list(collections)
        n = len(self.collections)
        hs = (C.c_void_p * n)(*[c._h.value for c in self.collections])
        ts = (C.c_float * n)(*[float(t) for t in thresholds])
        h = C.c_void_p()
        _check(lib().fgss_batch_create(hs, ts, n, C.byref(h)))
        self._h = h
        self.njobs = n

    def set_thresholds(self, thresholds):
        ts = (C.c_float * self.njobs)(*[float(t) for t in thresholds])
        _check(lib().fgss_batch_set_thresholds(self._h, ts))

    def run(self, graph: bool = True, table_slots: int = 0, load_pct: int = 0, stream: int = 0,
            index_groups: int = 0, block_records: int = 0) -> dict:
        st = BatchStats()
        opts = JoinOpts(load_pct, table_slots, 0, 0, 0, 0, 0, stream or None, index_groups, block_records)
        _check(lib().fgss_batch_run(self._h, C.byref(opts), 1 if graph else 0, C.byref(st)))
        return st.as_dict()

    def pairs(self, job: int) -> Pairs:
        out = C.POINTER(_Pairs)()
        _check(lib().fgss_batch_pairs(self._h, job, C.byref(out)))
        return _take_pairs(out)

    def all_pairs(self) -> list:
        """Every job's Pairs from the last run (as pairs(j) for each j), in one call."""
        out = C.POINTER(_Pairs)()
        first = (C.c_int64 * (self.njobs + 1))()
        stats = (JoinStats * self.njobs)()
        _check(lib().fgss_batch_pairs_all(self._h, C.byref(out), first, stats))
        try:
            p = out.contents
            n = p.n_pairs

            def cp(ptr, dt):
                return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0, dt)
            q, s, rid, sid = cp(p.q, np.int32), cp(p.s, np.int32), cp(p.rid, np.int32), cp(p.sid, np.int32)
            ov, sim = cp(p.overlap, np.int32), cp(p.similarity, np.float32)
        finally:
            lib().fgss_pairs_free(out)
        f = np.ctypeslib.as_array(first)
        return [Pairs(q[a:b], s[a:b], rid[a:b], sid[a:b], ov[a:b], sim[a:b], stats[j].as_dict())
                for j, (a, b) in enumerate(zip(f[:-1], f[1:]))]

    def autotune(self, reps: int = 3, cache_path: str | None = None):
        """The fastest options (fgss_batch_autotune); with cache_path, through the tuning
        cache there (fgss_batch_autotune_cached: a stored choice is returned untimed)."""
        best = JoinOpts()
        ms = C.c_double()
        hit = C.c_int32(0)
        if cache_path is None:
            _check(lib().fgss_batch_autotune(self._h, reps, C.byref(best), C.byref(ms)))
        else:
            _check(lib().fgss_batch_autotune_cached(self._h, reps, os.fsencode(cache_path), C.byref(best),
                                                    C.byref(ms), C.byref(hit)))
        return {"table_slots": best.table_slots, "load_pct": best.load_pct, "index_groups": best.index_groups,
                "block_records": best.block_records, "ms": ms.value, "from_cache": bool(hit.value)}

    def close(self):
        if self._h is not None and self._h.value:
            lib().fgss_batch_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """R-S joins against a resident index with continuous batching.

    `indexed` and every submitted probe collection must stay alive until their
    results are taken (record ids come from them)."""

    def __init__(self, indexed: Collection, t_min: float, slots: int = 2, budget_bytes: int = 0,
                 table_slots: int = 0, load_pct: int = 0, bounds_cap: int = 0, stream_parts: int = 0):
        self.indexed = indexed
        self._probes = {}
        opts = JoinOpts(load_pct, table_slots, stream_parts, bounds_cap)
        h = C.c_void_p()
        _check(lib().fgss_server_create(indexed._h, C.c_float(t_min), slots, budget_bytes, C.byref(opts),
                                         C.byref(h)))
        self._h = h

    def submit(self, probes: Collection, t: float) -> int:
        rid = C.c_int64()
        _check(lib().fgss_server_submit(self._h, probes._h, C.c_float(t), C.byref(rid)))
        self._probes[rid.value] = probes
        return rid.value

    def poll(self, block: bool = False):
        pend, run = C.c_int32(), C.c_int32()
        _check(lib().fgss_server_poll(self._h, 1 if block else 0, C.byref(pend), C.byref(run)))
        return pend.value, run.value

    def done(self, request_id: int) -> bool:
        d = C.c_int32()
        _check(lib().fgss_server_done(self._h, request_id, C.byref(d)))
        return bool(d.value)

    def wait(self, request_id: int):
        _check(lib().fgss_server_wait(self._h, request_id))

    def hold(self, on: bool = True):
        """Pause (or resume) admission: held requests queue and are then batched together."""
        _check(lib().fgss_server_hold(self._h, 1 if on else 0))

    def result(self, request_id: int, wait: bool = True):
        """(Pairs, request stats) of a request; it is forgotten afterwards."""
        if wait:
            self.wait(request_id)
        out = C.POINTER(_Pairs)()
        st = RequestStats()
        try:
            _check(lib().fgss_server_result(self._h, request_id, C.byref(out), C.byref(st)))
        finally:
            self._probes.pop(request_id, None)
        return _take_pairs(out), st.as_dict()

    def close(self):
        if self._h is not None and self._h.value:
            lib().fgss_server_free(self._h)
        self._h = None
        self._probes.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

