"""Benchmark: plain self-join of 1M synthetic long records (q=3, t=0.8) on one MI355X.

One step = one self-join of the HBM-resident collection, from the records on the
device to the sorted result pairs in host memory (index build, planning, candidate
generation, verification, compaction, device->host copy).  Data: 1M seeded synthetic
records of 200-1200 chars (~60 Zipf words, 5% near-duplicates with 1-10 edits),
tokenized into 3-gram sets on the GPU before timing.

Prints one JSON line: candidates verified per second (the reference's
"# Candidates" / join time) with the join time, the roofline of the join kernel and
the CPU oracle baseline on this host.  Secondary keys: 100k titles at
t = 0.6 / 0.75 / 0.9 ("titles"), 512 batched joins ("batched"), R-S serving against
a 2M-title index ("served").
"""
from __future__ import annotations

import argparse
import contextlib
import gc
import itertools
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 1024 * 16 * 2.4e9 / 1e12   # 32-bit VALU lane operations per second, in T
LONG_RECORDS, LONG_SEED, LONG_T = 1_000_000, 3, 0.8
TITLES_RECORDS, TITLES_SEED, TITLES_T = 100_000, 2, (0.6, 0.75, 0.9)
PMC = {"long": os.path.join(ROOT, "profiles", "k_join_long1m_pmc.txt"),
       "titles": os.path.join(ROOT, "profiles", "k_join_titles100k_pmc.txt")}


@contextlib.contextmanager
def gc_paused():
    """Timed regions run with the Python garbage collector paused (as `timeit` does):
    its full passes over the objects that earlier legs leave alive would otherwise land
    inside them (a served stream's p99 rose from ~4 ms to 14-20 ms that way)."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def design_bytes(st):
    """Bytes one join's kernels (k_join_dense and k_join) must read and write as designed: 4 B per probe
    token, 8 B per useful posting of a hash part (record id, position | length) and 4 B
    per posting of a dense part (record << 8 | position), one 64-byte signature line per
    candidate tested and a 128-byte one per candidate tested again, 4 B per verified
    source token, 12 B per result triple."""
    dense = st.get("dense_postings", 0)
    return (4 * st["probe_tokens"] + 8 * (st["postings_scanned"] - dense) + 4 * dense
            + 64 * st["signature_checks"] + 128 * st["signature_checks_128"] + 4 * st["verify_tokens"]
            + 12 * st["results"])


def pmc_summary(workload):
    """key -> value pairs of the committed rocprofv3 PMC summary of this workload
    (its last lines: `key value`)."""
    out = {}
    path = PMC[workload]
    if os.path.exists(path):
        for ln in open(path):
            parts = ln.split()
            if len(parts) == 2:
                try:
                    out[parts[0]] = float(parts[1])
                except ValueError:
                    pass
    return out


def roofline(workload, st_list, kernel_ms, extra=None):
    """Roofline of a join's kernels (k_join_dense, then k_join, on one stream) over joins with
    stats st_list and mean duration kernel_ms (HIP events on their stream around the two
    launches).  The kernels are bound by VALU issue, not HBM: L2 serves almost all of the
    bytes they read (the dense kernel's hit rate is 98%), so the bound reported is the VALU
    issue rate -- 1,024 SIMDs x 16 lanes x 2.4 GHz of 32-bit lane operations, a wave64 VALU
    instruction taking four cycles -- with the VALU instructions per join from the same-build
    PMC summary; the HBM figures (measured bytes and the bytes read as designed) ride beside it."""
    kbytes = sum(design_bytes(st) for st in st_list) / len(st_list)
    sec = kernel_ms / 1e3 if kernel_ms > 0 else None
    prof = pmc_summary(workload)
    fetch, write = prof.get("fetch_bytes_per_launch"), prof.get("write_bytes_per_launch")
    traffic = fetch + write if fetch is not None and write is not None else None
    vi = prof.get("valu_insts_per_launch")
    achieved = vi * 64 / sec / 1e12 if vi and sec else None
    hbm_gbs = traffic / sec / 1e9 if traffic and sec else None
    out = {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS,
           "unit": "T VALU lane-ops/s (32-bit; 1024 SIMDs x 16 lanes x 2.4 GHz)",
           "frac": achieved / VALU_PEAK_TOPS if achieved else None, "traffic": traffic,
           "kernel": "k_join_dense + k_join", "kernel_ms_per_launch": kernel_ms,
           "valu_insts_per_launch": vi,
           "basis": "VALU wave instructions of a join's two kernels (SQ_INSTS_VALU, same-build PMC summary) x 64 "
                    "lanes / HIP-event time of the two launches",
           "traffic_basis": "unscaled FETCH_SIZE + WRITE_SIZE of a join's two kernels from the same-build PMC summary",
           "traffic_source": os.path.relpath(PMC[workload], ROOT) if prof else None,
           "hbm": {"achieved_gbs": hbm_gbs, "peak_gbs": HBM_PEAK_GBS,
                   "frac": hbm_gbs / HBM_PEAK_GBS if hbm_gbs else None},
           "design_bytes_per_launch": kbytes,
           "design_gbs": kbytes / sec / 1e9 if sec else None,
           "design_note": "bytes the kernels read as designed (probe tokens, useful postings, signature lines, "
                          "verified tokens, results) per kernel second: L2 serves almost all of them, so the rate "
                          "can pass the HBM peak; a consumption rate, not an HBM figure",
           "valu_busy": prof.get("valu_busy"), "lds_bank_conflict_ratio": prof.get("lds_bank_conflict_ratio")}
    if extra:
        out.update(extra)
    return out


def host_cpu():
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "affinity": avail, "model": model}


def cpu_baseline_long(coll, pairs, t, sample_all: int, sample_one: int, seed: int = 11):
    """The oracle (each candidate verified by the reference's merge, once) on a seeded
    sample of the join's queries, each against the whole collection: all the host
    threads this process may use (at most 16, the host share of one GPU) on
    `sample_all` queries and one thread on the first `sample_one` of them, each 1
    warm-up + median of 5.  The GPU's pairs of the sampled queries must equal the
    oracle's bit for bit."""
    import numpy as np

    import oracle
    oracle.build()
    ids, lens, pos, tok = coll.arrays()
    sets = oracle.Sets.from_arrays(ids, lens, pos, tok, coll.info.n_terms)
    qs = np.sort(np.random.default_rng(seed).choice(sets.n, size=min(sample_all, sets.n), replace=False))
    cpu = host_cpu()
    threads = max(1, min(16, cpu["affinity"]))
    res = {}
    for name, nth, q in (("all", threads, qs), ("one", 1, np.sort(qs[:sample_one]))):
        runs = []
        for r in range(6):
            tm = []
            tri, ctr = oracle.selfjoin(sets, t, nthreads=nth, queries=q, checks=False, times=tm)
            if r > 0:
                runs.append(tm)
        runs.sort(key=lambda x: x[1])
        med = runs[len(runs) // 2]
        res[name] = {"threads": nth, "queries": int(q.size), "index_s": med[0], "query_s": med[1],
                     "candidates": ctr["candidates"], "candidates_per_s": ctr["candidates"] / med[1], "ctr": ctr,
                     "tri": tri}
        log(f"cpu baseline: {nth} thread(s), {q.size} queries: {med[1]:.2f} s (+ index {med[0]:.2f} s), "
            f"{ctr['candidates'] / med[1]:.3g} candidates/s")
    got = np.stack([pairs.q, pairs.s, pairs.overlap], axis=1) if len(pairs) else np.zeros((0, 3), np.int32)
    keep = np.isin(got[:, 0], qs) if len(got) else np.zeros(0, bool)
    exact = bool(np.array_equal(got[keep], res["all"]["tri"]))
    a, o = res["all"], res["one"]
    n = sets.n
    whole_s = a["index_s"] + a["query_s"] * n / a["queries"]
    out = {"value": a["candidates_per_s"], "unit": "candidates/s", "cores": a["threads"], "kind": "port",
           "sample": f"{a['queries']} random queries (seed {seed}) of the {n}-record t={t} join, each against the "
                     f"whole collection; oracle/fgss_oracle.c closed form, each candidate verified by the reference's "
                     f"merge once; query phase, median of 5 after 1 warm-up",
           "one_core": {"value": o["candidates_per_s"], "queries": o["queries"], "query_s": o["query_s"]},
           "all_cores": {"threads": a["threads"], "queries": a["queries"], "query_s": a["query_s"],
                         "index_build_s": a["index_s"]},
           "host": cpu, "whole_join_estimate_s": whole_s}
    return out, exact, a["ctr"]


def long_leg(fg, args):
    """The headline: n long records (seed 3), q=3, t=0.8, tokenized on the GPU."""
    import numpy as np
    t = float(np.float32(LONG_T))
    t0 = time.perf_counter()
    text = fg.generate_text(fg.GEN_LONG, args.long_records, LONG_SEED)
    t1 = time.perf_counter()
    coll = fg.Collection.from_text_bytes(text, 3, device=True)
    tok_s = time.perf_counter() - t1
    del text
    info = coll.info
    log(f"long: {info.n_sets} records, {info.n_tokens} tokens, {info.n_terms} terms (GPU tokenizer {tok_s:.2f} s, "
        f"set-up {time.perf_counter() - t0:.1f} s)")
    tu = time.perf_counter()
    coll.upload()
    fg.device_synchronize()
    upload_s = time.perf_counter() - tu
    coll.selfjoin(t)   # sizes the workspace
    for i in range(args.warmup):
        coll.selfjoin(t)
        log(f"warmup {i + 1}/{args.warmup} done")
    fg.device_synchronize()
    runs = []
    with gc_paused():
        a = time.perf_counter()
        for i in range(args.steps):
            runs.append(coll.selfjoin(t))
        fg.device_synchronize()
        dt = time.perf_counter() - a
    log(f"timed {args.steps} steps in {dt:.3f} s")
    p = runs[-1]
    st = p.stats
    ms_per_step = dt * 1e3 / args.steps
    kernel_ms = sum(r.stats["ms_kernel"] for r in runs) / len(runs)
    out = {"records": int(info.n_sets), "tokens": int(info.n_tokens), "terms": int(info.n_terms),
           "tokenize_s": tok_s, "upload_s": upload_s,
           "outside_step": "tokenizing and the upload (host-to-device copy of the records and their "
                           "token-bucket signatures, k_signature: about 1 ms at 1M records) precede the "
                           "timed steps; a step starts from the HBM-resident records",
           "ms_per_step": ms_per_step, "join_ms": st["ms_total"], "index_ms": st["ms_index"],
           "kernel_ms": kernel_ms, "pairs": len(p), "candidates": st["candidates"],
           "postings": st["postings_scanned"], "probe_tokens": st["probe_tokens"],
           "index_entries": st["index_entries"], "signature_checks": st["signature_checks"],
           "signature_checks_128": st["signature_checks_128"], "verified": st["verified"],
           "verify_tokens": st["verify_tokens"], "kernel_launches": st["kernel_launches"],
           "phase_cycles": st["phase_cycles"], "window_retries": st["window_retries"],
           "dense_parts": st["dense_parts"], "dense_postings": st["dense_postings"],
           "repeatable": all(np.array_equal(r.q, p.q) and np.array_equal(r.s, p.s) for r in runs)}
    cb, ctr = None, None
    if not args.no_cpu:
        log("long: cpu baseline on sampled queries ...")
        cb, exact, ctr = cpu_baseline_long(coll, p, t, args.cpu_sample, args.cpu_sample_one)
        out["bit_exact_vs_oracle_on_sample"] = exact
        cb["speedup"] = cb["whole_join_estimate_s"] / (ms_per_step / 1e3)
        log(f"long: sample bit-exact {exact}; CPU whole-join estimate {cb['whole_join_estimate_s']:.0f} s "
            f"({cb['speedup']:.0f}x)")
    extra = {}
    if ctr is not None:
        # the reference algorithm's bytes: its merge reads scale with the candidates, so
        # the sample's count is scaled to the whole join; the other terms are exact
        m_est = ctr["m_reads"] * st["candidates"] / max(ctr["candidates"], 1)
        rb = 4 * st["probe_tokens"] + 8 * st["postings_scanned"] + 4 * m_est + 8 * st["results"]
        extra = {"reference_bytes_per_launch": rb,
                 "reference_equivalent_gbs": rb / (kernel_ms / 1e3) / 1e9,
                 "reference_equivalent_note": "the reference algorithm's bytes (merge reads estimated from the "
                                              "oracle's sample) per kernel second: a work-equivalent rate, not "
                                              "bandwidth; the signature bound skips most of those reads"}
    roof = roofline("long", [r.stats for r in runs], kernel_ms, extra)
    coll.close()
    return out, roof, cb


def titles_leg(fg, args):
    """100k titles (seed 2) at t = 0.6, 0.75, 0.9: one step = the three joins."""
    import numpy as np

    import oracle
    coll = fg.Collection.generate(fg.GEN_TITLES, TITLES_RECORDS, TITLES_SEED, device=True)
    coll.upload()

    def step():
        return [coll.selfjoin(t) for t in TITLES_T]

    for _ in range(max(1, args.warmup)):
        step()
    fg.device_synchronize()
    kms, last = [], None
    with gc_paused():
        t0 = time.perf_counter()
        for _ in range(args.titles_steps):
            last = step()
            kms.append([p.stats["ms_kernel"] for p in last])
        fg.device_synchronize()
        dt = time.perf_counter() - t0
    ms = dt * 1e3 / args.titles_steps
    kms_t = [sum(k[i] for k in kms) / len(kms) for i in range(len(TITLES_T))]
    cand = sum(p.stats["candidates"] for p in last)
    per_t = {str(t): {"join_ms": round(p.stats["ms_total"], 3), "kernel_ms": round(km, 3),
                      "index_ms": round(p.stats["ms_index"], 3), "pairs": len(p),
                      "candidates": p.stats["candidates"], "postings": p.stats["postings_scanned"],
                      "sig_checks": p.stats["signature_checks"], "verified": p.stats["verified"],
                      "dense_parts": p.stats["dense_parts"], "dense_postings": p.stats["dense_postings"],
                      "phase_cycles": p.stats["phase_cycles"]}
             for t, p, km in zip(TITLES_T, last, kms_t)}
    out = {"workload": f"plain self-join, {TITLES_RECORDS} synthetic titles, q=3, t={list(TITLES_T)}; a step is the "
                       "three joins", "ms_per_step": ms, "candidates_per_s": cand / (ms / 1e3),
           "per_threshold": per_t}
    # roofline over the three launches of a step (the PMC summary's figures are per join,
    # averaged over the three)
    out["roofline"] = roofline("titles", [p.stats for p in last], sum(kms_t) / len(TITLES_T))
    if not args.no_cpu:
        oracle.build()
        ids, lens, pos, tok = coll.arrays()
        sets = oracle.Sets.from_arrays(ids, lens, pos, tok, coll.info.n_terms)
        threads = max(1, min(16, host_cpu()["affinity"]))
        exact, cpu_s, ccand = True, 0.0, 0
        for t, p in zip(TITLES_T, last):
            tm = []
            tri, ctr = oracle.selfjoin(sets, t, nthreads=threads, checks=False, times=tm)
            cpu_s += tm[0] + tm[1]
            ccand += ctr["candidates"]
            got = np.stack([p.q, p.s, p.overlap], axis=1) if len(p) else np.zeros((0, 3), np.int32)
            exact &= bool(np.array_equal(got, tri)) and p.stats["candidates"] == ctr["candidates"]
        out["bit_exact_vs_oracle"] = exact
        out["cpu_oracle"] = {"threads": threads, "step_s": cpu_s, "candidates_per_s": ccand / cpu_s}
    coll.close()
    return out


def batched_leg(fg, n_jobs: int, reps: int):
    """512 independent self-joins (titles, sizes log-uniform in 500..50k, per-job seed
    4000+j, t in {0.50, 0.55, ..., 0.95}): one batched pipeline run (autotuned, replayed
    HIP graph) vs the same joins as sequential plain launches."""
    import numpy as np
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(500), np.log(50000), n_jobs)).astype(int)
    ts = [float(np.float32(x)) for x in rng.choice(np.round(np.arange(0.5, 0.951, 0.05), 2), n_jobs)]
    t0 = time.perf_counter()
    colls = [fg.Collection.generate(fg.GEN_TITLES, int(n), 4000 + j, device=True) for j, n in enumerate(sizes)]
    log(f"batched: generated {n_jobs} collections, {int(sizes.sum())} records in {time.perf_counter() - t0:.1f} s")
    for c in colls:
        c.upload()
    # plain: one launch sequence per join, results in host memory
    for c, t in zip(colls, ts):   # warm: every collection's workspace is allocated before timing
        c.selfjoin(t)
    fg.device_synchronize()
    with gc_paused():
        t0 = time.perf_counter()
        plain = [c.selfjoin(t) for c, t in zip(colls, ts)]
        plain_s = time.perf_counter() - t0
    log(f"batched: {n_jobs} sequential plain joins in {plain_s * 1e3:.1f} ms")
    b = fg.Batch(colls, ts)
    tune = b.autotune(reps=2)   # table slots x load, then index groups and record blocks
    kw = {k: tune[k] for k in ("table_slots", "load_pct", "index_groups", "block_records")}
    b.run(graph=True, **kw)
    times = []
    exact = True
    for r in range(reps):
        with gc_paused():
            t0 = time.perf_counter()
            b.run(graph=True, **kw)
            got = b.all_pairs()
            times.append(time.perf_counter() - t0)
        if r == 0:
            for g, p in zip(got, plain):
                exact &= (np.array_equal(g.q, p.q) and np.array_equal(g.s, p.s)
                          and np.array_equal(g.overlap, p.overlap))
    times.sort()
    batch_s = times[len(times) // 2]
    log(f"batched: one graph-replayed batch run in {batch_s * 1e3:.1f} ms (median of {reps}), tune {tune}")
    b.close()
    return {"jobs": n_jobs, "records": int(sizes.sum()),
            "plain_sequential_ms": plain_s * 1e3, "batched_ms": batch_s * 1e3,
            "plain_joins_per_s": n_jobs / plain_s, "batched_joins_per_s": n_jobs / batch_s,
            "speedup": plain_s / batch_s, "autotune": tune,
            "partitioning": {"index_groups": max(1, tune["index_groups"]), "block_records": tune["block_records"]},
            "bit_exact_vs_plain": bool(exact)}


def _edit(line: bytes, rng, n_edits: int) -> bytes:
    s = bytearray(line)
    for _ in range(n_edits):
        if not s:
            break
        i = int(rng.integers(len(s)))
        op = int(rng.integers(3))
        if op == 0:
            s[i] = int(rng.integers(97, 123))
        elif op == 1:
            del s[i]
        else:
            s.insert(i, int(rng.integers(97, 123)))
    return bytes(s)


def _percentile(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs, np.float64), p)) if len(xs) else 0.0


def served_leg(fg, n_index: int, n_requests: int, slots: int, check_tokenizer: bool = True, server_kw=None):
    """R-S serving against a resident index of n_index titles (seed 5, q=3, t_min=0.7):
    requests of 1..4096 probes (log-uniform; 20% edited copies of indexed titles, the
    rest fresh titles), t ~ U[0.7, 0.95].  Plain = one request at a time (submit, wait,
    take the pairs); served = continuous batching of a Poisson stream, at a rate far
    above capacity (throughput) and at half the measured capacity (latency)."""
    idx, reqs, ts, tok = served_workload(fg, n_index, n_requests, check_tokenizer)
    out = served_measure(fg, idx, reqs, ts, slots, server_kw or {})
    out["tokenize_index"] = tok
    idx.close()
    return out


def served_workload(fg, n_index: int, n_requests: int, check_tokenizer: bool = True):
    """The resident index and the tokenized requests of the served leg."""
    import numpy as np
    t0 = time.perf_counter()
    text = fg.generate_text(fg.GEN_TITLES, n_index, 5)
    t1 = time.perf_counter()
    idx = fg.Collection.from_text_bytes(text, 3, device=True)
    dev_s = time.perf_counter() - t1
    log(f"served: index of {idx.info.n_sets} titles ({idx.info.n_tokens} tokens) tokenized on the GPU in "
        f"{dev_s:.2f} s")
    tok = {"records": int(idx.info.n_sets), "bytes": len(text), "qgrams": int(idx.info.n_tokens),
           "terms": int(idx.info.n_terms), "device_s": dev_s}
    if check_tokenizer:
        t1 = time.perf_counter()
        host = fg.Collection.from_text_bytes(text, 3)
        host_s = time.perf_counter() - t1
        exact = all(np.array_equal(a, b) for a, b in zip(host.arrays(), idx.arrays()))
        exact &= host.info.n_terms == idx.info.n_terms
        host.close()
        tok.update({"host_s": host_s, "host_threads": min(16, os.cpu_count() or 1), "speedup": host_s / dev_s,
                    "bit_exact": bool(exact)})
        log(f"served: host tokenizer {host_s:.2f} s, bit-exact {exact}")
    lines = [ln for ln in text.split(b"\n") if ln]
    del text
    rng = np.random.default_rng(55)
    sizes = np.minimum(np.exp(rng.uniform(0.0, np.log(4096.0), n_requests)).astype(int), 4096)
    sizes = np.maximum(sizes, 1)
    ts = [float(np.float32(x)) for x in rng.uniform(0.7, 0.95, n_requests)]
    fresh = [ln for ln in fg.generate_text(fg.GEN_TITLES, int(sizes.sum()), 50).split(b"\n") if ln]
    reqs, f = [], 0
    for n in sizes:
        out = []
        for _ in range(int(n)):
            if rng.random() < 0.2:
                out.append(_edit(lines[int(rng.integers(len(lines)))], rng, int(rng.integers(1, 4))))
            else:
                out.append(fresh[f % len(fresh)])
                f += 1
        reqs.append(fg.Collection.from_text_dict(b"\n".join(out) + b"\n", idx))
    del lines, fresh
    log(f"served: {n_requests} requests, {int(sizes.sum())} probes tokenized in {time.perf_counter() - t0:.1f} s")
    return idx, reqs, ts, tok


def served_measure(fg, idx, reqs, ts, slots: int, server_kw: dict):
    """Plain (one request at a time) against served (continuous batching) on the same
    requests; every served request's pairs must equal its plain run's."""
    import numpy as np
    t_min = float(np.float32(0.7))
    n_requests = len(reqs)
    sizes = np.array([r.info.n_sets for r in reqs])
    srv = fg.Server(idx, t_min, slots=1, **server_kw)
    for k in range(min(8, n_requests)):   # warm the slot's workspace
        srv.result(srv.submit(reqs[k], ts[k]))
    plain, plain_lat = [], []
    with gc_paused():
        t1 = time.perf_counter()
        for r, t in zip(reqs, ts):
            a = time.perf_counter()
            plain.append(srv.result(srv.submit(r, t))[0])
            plain_lat.append(time.perf_counter() - a)
        plain_s = time.perf_counter() - t1
    srv.close()
    log(f"served: plain sequential {n_requests} requests in {plain_s * 1e3:.1f} ms")

    def drive(rate, n_slots=slots, budget=0):
        """Open-loop Poisson arrivals at `rate` req/s; returns (makespan s, latencies s, pairs)."""
        srv = fg.Server(idx, t_min, slots=n_slots, budget_bytes=budget, **server_kw)
        for k in range(min(8, n_requests)):
            srv.result(srv.submit(reqs[k], ts[k]))
        gaps = np.random.default_rng(7).exponential(1.0 / rate, n_requests)
        arrive = np.cumsum(gaps) - gaps[0]
        out, lat = [None] * n_requests, [0.0] * n_requests
        live = {}
        i = 0
        shared = []
        with gc_paused():
            start = time.perf_counter()
            while i < n_requests or live:
                now = time.perf_counter() - start
                while i < n_requests and arrive[i] <= now:
                    live[srv.submit(reqs[i], ts[i])] = i
                    i += 1
                # with every request submitted, wait for the oldest run instead of spinning
                srv.poll(block=i >= n_requests)
                # runs are admitted first come first served, so requests finish about in
                # arrival order: look at the oldest live ones only
                finished = [rid for rid in itertools.islice(live, 64) if srv.done(rid)]
                for rid in finished:
                    k = live.pop(rid)
                    out[k], st = srv.result(rid, wait=False)
                    lat[k] = time.perf_counter() - start - arrive[k]
                    shared.append(st["batch_requests"])
                if not finished and i < n_requests and not live:
                    time.sleep(max(0.0, min(arrive[i] - (time.perf_counter() - start), 0.0005)))
            makespan = time.perf_counter() - start
        srv.close()
        return makespan, lat, out, shared

    # the saturated stream's makespan depends on how the first arrivals split into
    # runs: the median of three streams
    sat = sorted((drive(rate=1e7) for _ in range(3)), key=lambda r: r[0])
    sat_s, sat_lat, sat_out, sat_shared = sat[1]
    cap = n_requests / sat_s
    log(f"served: saturated stream {n_requests} requests in {sat_s * 1e3:.1f} ms ({cap:.0f} req/s, median of 3)")
    lat_s, lat, lat_out, lat_shared = drive(rate=0.5 * cap)
    log(f"served: Poisson at {0.5 * cap:.0f} req/s, p50 {_percentile(lat, 50) * 1e3:.2f} ms, "
        f"p99 {_percentile(lat, 99) * 1e3:.2f} ms")
    # equal load: plain (one request per run, first come first served: a work budget
    # of one byte admits a request only when nothing is in flight) and served at the
    # same Poisson rate, half the plain capacity
    eq_rate = 0.5 * n_requests / plain_s
    _, eq_plain_lat, eq_plain_out, _ = drive(rate=eq_rate, n_slots=1, budget=1)
    _, eq_lat, eq_out, _ = drive(rate=eq_rate)
    log(f"served: equal load {eq_rate:.0f} req/s: plain p50/p99 {_percentile(eq_plain_lat, 50) * 1e3:.2f}/"
        f"{_percentile(eq_plain_lat, 99) * 1e3:.2f} ms, served {_percentile(eq_lat, 50) * 1e3:.2f}/"
        f"{_percentile(eq_lat, 99) * 1e3:.2f} ms")
    exact = True
    for outs in (sat_out, lat_out, eq_plain_out, eq_out):
        for g, p in zip(outs, plain):
            exact &= (np.array_equal(g.q, p.q) and np.array_equal(g.s, p.s) and np.array_equal(g.overlap, p.overlap)
                      and g.stats["candidates"] == p.stats["candidates"])
    ms = lambda xs, q: round(_percentile(xs, q) * 1e3, 3)
    return {"index_records": int(idx.info.n_sets), "requests": n_requests, "probes": int(sizes.sum()), "slots": slots,
            "t_min": t_min, "pairs": int(sum(len(p) for p in plain)),
            "candidates": int(sum(p.stats["candidates"] for p in plain)),
            "plain_req_per_s": n_requests / plain_s, "plain_p50_ms": ms(plain_lat, 50),
            "plain_p99_ms": ms(plain_lat, 99),
            "served_req_per_s": cap, "speedup": cap / (n_requests / plain_s),
            "saturated_p50_ms": ms(sat_lat, 50), "saturated_p99_ms": ms(sat_lat, 99),
            "saturated_mean_batch": float(np.mean(sat_shared)) if sat_shared else 0.0,
            "poisson_rate": 0.5 * cap, "poisson_req_per_s": n_requests / lat_s,
            "poisson_p50_ms": ms(lat, 50), "poisson_p99_ms": ms(lat, 99),
            "poisson_mean_batch": float(np.mean(lat_shared)) if lat_shared else 0.0,
            "equal_load_rate": eq_rate, "equal_load_plain_p50_ms": ms(eq_plain_lat, 50),
            "equal_load_plain_p99_ms": ms(eq_plain_lat, 99), "equal_load_served_p50_ms": ms(eq_lat, 50),
            "equal_load_served_p99_ms": ms(eq_lat, 99),
            "bit_exact_vs_plain": bool(exact)}


def calibrate_copy(fg, nbytes: int = 1 << 30, reps: int = 3):
    """`reps` device-to-device copies of `nbytes` (the HIP runtime's copy kernel, after
    one fill of the source): the reference point for the PMC byte counters, whose
    per-dispatch FETCH_SIZE and WRITE_SIZE should read nbytes / 1024 KiB each."""
    import ctypes as C
    fg.lib()
    path = next((ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64.so" in ln), None)
    if path is None:
        raise RuntimeError("the HIP runtime is not mapped")
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(src), nbytes) or h.hipMalloc(C.byref(dst), nbytes):
        raise RuntimeError("hipMalloc failed")
    h.hipMemset(src, 7, nbytes)
    t0 = time.perf_counter()
    for _ in range(reps):
        if h.hipMemcpy(dst, src, nbytes, 3):   # device to device
            raise RuntimeError("hipMemcpy failed")
    fg.device_synchronize()
    dt = (time.perf_counter() - t0) / reps
    h.hipFree(src)
    h.hipFree(dst)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "ms_per_copy": dt * 1e3,
                      "gbs": 2 * nbytes / dt / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed joins of the headline workload")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle legs")
    ap.add_argument("--long-records", type=int, default=LONG_RECORDS,
                    help="records of the headline workload (0 skips it: profiling runs of the other legs)")
    ap.add_argument("--cpu-sample", type=int, default=200, help="queries of the all-threads CPU baseline")
    ap.add_argument("--cpu-sample-one", type=int, default=30, help="queries of the one-thread CPU baseline")
    ap.add_argument("--titles-steps", type=int, default=5, help="steps of the 100k-title leg (0 skips it)")
    ap.add_argument("--batch-jobs", type=int, default=512, help="jobs of the batched leg (0 skips it)")
    ap.add_argument("--serve-index", type=int, default=2_000_000, help="indexed titles of the served leg")
    ap.add_argument("--serve-requests", type=int, default=1000, help="requests of the served leg (0 skips it)")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 1 GiB device-to-device copies (the byte counters' reference point)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the join does not shard across GPUs "
              "in this build", file=sys.stderr)
        sys.exit(2)

    import new_fgssjoin_amd as fg
    if args.calibrate_copy:
        calibrate_copy(fg)
        return
    # fgss_selfjoin returns only after its stream has drained and the pairs are in
    # host memory, so each timed step is already bracketed by device synchronization.
    # torch is deliberately not imported: it would load a second HIP runtime.
    if args.long_records > 0:
        long, roof, cb = long_leg(fg, args)
    else:   # profiling runs of the other legs alone
        long, roof, cb = {"records": 0, "candidates": 0, "ms_per_step": None}, None, None
    out = {
        "metric": "join ms + candidates verified/s (bit-exact); batched & served p99 vs plain",
        "value": long["candidates"] / (long["ms_per_step"] / 1e3) if long["ms_per_step"] else None,
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": long["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": dict({"workload": f"plain self-join, {long['records']} synthetic long records (200-1200 chars, "
                                    f"seed {LONG_SEED}), q=3, t={LONG_T}; a step is one join"}, **long),
        "roofline": roof,
        "cpu_baseline": cb,
    }
    if args.titles_steps > 0:
        out["titles"] = titles_leg(fg, args)
    if args.batch_jobs > 0:
        out["batched"] = batched_leg(fg, args.batch_jobs, reps=3)
    if args.serve_requests > 0:
        out["served"] = served_leg(fg, args.serve_index, args.serve_requests, args.serve_slots)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
