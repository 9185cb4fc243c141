"""Benchmark: plain self-join of 100k synthetic titles (q=3) at t = 0.6, 0.75, 0.9.

One step = the three self-joins in sequence, each from the HBM-resident
collection to the sorted result pairs in host memory (index build, candidate
generation, verification, compaction, device->host copy).  Data: seeded
synthetic titles (4-15 Zipf words over a 50k-word vocabulary, 20-120 chars,
10% near-duplicates), tokenized into 3-gram sets before timing.

Prints one JSON line: candidates verified per second (the reference's
"# Candidates", summed over the step) with the join time, the roofline of the
join kernel and the CPU oracle baseline on this host.
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

THRESHOLDS = (0.6, 0.75, 0.9)
N_RECORDS = 100_000
SEED = 2
HBM_PEAK_GBS = 8000.0


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


PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_join_titles100k_pmc.txt")


def alg_bytes_join_kernel(c):
    """Algorithmic bytes of one join's candidate generation + verification (the
    k_join launch), counted by the oracle from the reference algorithm: 4 B per probe
    token, 8 B per useful posting (sid, pos), 4 B per token read of the reference's
    merge loop (checking.cu:126-160), 8 B per result pair."""
    return 4 * c["p_q"] + 8 * c["e_scan"] + 4 * c["m_reads"] + 8 * c["results"]


def alg_bytes_join(c):
    """... plus the index build: 12 B per index entry (read a token, write sid, pos)."""
    return 12 * c["e_idx"] + alg_bytes_join_kernel(c)


def design_bytes(st):
    """Bytes this build's k_join reads for one join (no oracle needed): 4 B per probe
    token, 8 B per useful posting, 64 B per 64-bucket signature test, 128 B per
    128-bucket test, 4 B per verified source token, 12 B per result."""
    return (4 * st["probe_tokens"] + 8 * st["postings_scanned"] + 64 * st["signature_checks"]
            + 128 * st["signature_checks_128"] + 4 * st["verify_tokens"] + 12 * st["results"])


def pmc_summary():
    """key -> value from the committed rocprofv3 PMC summary of this workload"""
    out = {}
    if os.path.exists(PMC_SUMMARY):
        for ln in open(PMC_SUMMARY):
            parts = ln.split()
            if len(parts) == 2:
                try:
                    out[parts[0]] = float(parts[1])
                except ValueError:
                    pass
    return out


def cpu_baseline(coll, gpu_pairs):
    """The CPU oracle on this host over the same step (all three thresholds)."""
    import numpy as np

    import oracle
    oracle.build()
    ids, lens, pos, tok = coll.arrays()
    sets = oracle.Sets.from_arrays(ids, lens, pos, tok, coll.info.n_terms)
    threads = max(1, min(16, os.cpu_count() or 1))
    cand = 0
    exact = True
    counters = []
    t0 = time.perf_counter()
    for t, p in zip(THRESHOLDS, gpu_pairs):
        tri, ctr = oracle.selfjoin(sets, t, nthreads=threads)
        counters.append(ctr)
        cand += ctr["candidates"]
        got = np.stack([p.q, p.s, p.overlap], axis=1) if len(p) else np.zeros((0, 3), np.int32)
        exact &= got.shape == tri.shape and bool(np.array_equal(got, tri)) and p.stats["candidates"] == ctr["candidates"]
    dt = time.perf_counter() - t0
    return {"value": cand / dt, "unit": "candidates/s", "cores": threads, "kind": "port",
            "sample": f"full step: {N_RECORDS} titles x t={list(THRESHOLDS)}, {dt:.1f} s, "
                      f"oracle/fgss_oracle.c closed form, {threads} threads"}, exact, counters


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
    tune = b.autotune(reps=2)
    b.run(graph=True, table_slots=tune["table_slots"], load_pct=tune["load_pct"])
    times = []
    exact = True
    for r in range(reps):
        with gc_paused():
            t0 = time.perf_counter()
            b.run(graph=True, table_slots=tune["table_slots"], load_pct=tune["load_pct"])
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
            "speedup": plain_s / batch_s, "autotune": tune, "bit_exact_vs_plain": bool(exact)}


def large_cpu_baseline(coll, pairs, t, sample: int, seed: int = 11):
    """The oracle on a uniform sample of queries of the large join (each against the
    whole collection): the CPU rate, the GPU's pairs for those queries bit-exact, and
    the sample's merge-read count for the algorithmic bytes."""
    import numpy as np

    import oracle
    oracle.build()
    ids, lens, pos, tok = coll.arrays()
    sets = oracle.Sets.from_arrays(ids, lens, pos, tok, coll.info.n_terms)
    qs = np.sort(np.random.default_rng(seed).choice(sets.n, size=min(sample, sets.n), replace=False))
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    tri, ctr = oracle.selfjoin(sets, t, nthreads=threads, queries=qs)
    dt = time.perf_counter() - t0
    got = np.stack([pairs.q, pairs.s, pairs.overlap], axis=1) if len(pairs) else np.zeros((0, 3), np.int32)
    got = got[np.isin(got[:, 0], qs)] if len(got) else got
    exact = got.shape == tri.shape and bool(np.array_equal(got, tri))
    return ({"value": ctr["candidates"] / dt, "unit": "candidates/s", "cores": threads, "kind": "port",
             "sample": f"{qs.size} random queries (seed {seed}) of the {sets.n}-record join against the whole "
                       f"collection, {dt:.1f} s, oracle/fgss_oracle.c closed form, {threads} threads",
             "sample_s": dt, "sample_queries": int(qs.size)}, exact, ctr)


def large_leg(fg, n: int, sample: int, steps: int = 2, t: float = 0.8):
    """Plain large self-join: n long records (200-1200 chars of Zipf words, 5%
    near-duplicates with 1-10 edits, seed 3), q=3, t=0.8, tokenized on the GPU."""
    import numpy as np
    t = float(np.float32(t))
    t0 = time.perf_counter()
    text = fg.generate_text(fg.GEN_LONG, n, 3)
    t1 = time.perf_counter()
    coll = fg.Collection.from_text_bytes(text, 3, device=True)
    tok_s = time.perf_counter() - t1
    del text
    info = coll.info
    log(f"large: {info.n_sets} long records, {info.n_tokens} tokens (GPU tokenizer {tok_s:.2f} s, "
        f"set-up {time.perf_counter() - t0:.1f} s)")
    coll.upload()
    coll.selfjoin(t)   # sizes the workspace
    runs = []
    for _ in range(steps):
        with gc_paused():
            a = time.perf_counter()
            p = coll.selfjoin(t)
            runs.append((time.perf_counter() - a, p))
    runs.sort(key=lambda r: r[0])
    wall_s, p = runs[len(runs) // 2]
    st = p.stats
    log(f"large: join {st['ms_total']:.0f} ms (kernel {st['ms_kernel']:.0f} ms), {len(p)} pairs, "
        f"{st['candidates']} candidates")
    out = {"workload": f"plain self-join, {info.n_sets} synthetic long records (200-1200 chars), q=3, t=0.8",
           "records": int(info.n_sets), "tokens": int(info.n_tokens), "terms": int(info.n_terms),
           "tokenize_s": tok_s, "join_ms": st["ms_total"], "kernel_ms": st["ms_kernel"], "wall_ms": wall_s * 1e3,
           "pairs": len(p), "candidates": st["candidates"], "candidates_per_s": st["candidates"] / (st["ms_total"] / 1e3),
           "postings": st["postings_scanned"], "probe_tokens": st["probe_tokens"],
           "index_entries": st["index_entries"], "signature_checks": st["signature_checks"],
           "verified": st["verified"], "kernel_launches": st["kernel_launches"]}
    if sample > 0:
        log(f"large: cpu baseline on {sample} sampled queries ...")
        cb, exact, c = large_cpu_baseline(coll, p, t, sample)
        out["cpu_baseline"] = cb
        out["bit_exact_vs_oracle_on_sample"] = exact
        # merge reads of the reference scale with the candidates: estimate the whole
        # join's from the sample, the other terms are the GPU's exact counts
        m_est = c["m_reads"] * st["candidates"] / max(c["candidates"], 1)
        kbytes = 4 * st["probe_tokens"] + 8 * st["postings_scanned"] + 4 * m_est + 8 * st["results"]
        achieved = kbytes / (st["ms_kernel"] / 1e3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "kernel": "k_join", "bytes_per_launch": kbytes,
                           "basis": "algorithmic bytes of the reference algorithm; merge reads estimated from the "
                                    "oracle's sample (scaled by candidates), the rest counted exactly"}
        cpu_full_s = cb["sample_s"] * info.n_sets / max(cb["sample_queries"], 1)
        out["cpu_full_estimate_s"] = cpu_full_s
        out["speedup_vs_cpu_estimate"] = cpu_full_s / (st["ms_total"] / 1e3)
        log(f"large: sample bit-exact {exact}; CPU estimate {cpu_full_s:.0f} s for the whole join "
            f"({out['speedup_vs_cpu_estimate']:.0f}x); roofline {achieved:.0f} GB/s")
    coll.close()
    return out


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--table-slots", type=int, default=0)
    ap.add_argument("--load-pct", type=int, default=0)
    ap.add_argument("--batch-jobs", type=int, default=512, help="jobs of the batched leg (0 skips it)")
    ap.add_argument("--serve-index", type=int, default=2_000_000, help="indexed titles of the served leg")
    ap.add_argument("--serve-requests", type=int, default=1000, help="requests of the served leg (0 skips it)")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--large-records", type=int, default=1_000_000, help="records of the large leg (0 skips it)")
    ap.add_argument("--large-sample", type=int, default=1000, help="oracle-checked queries of the large leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the join does not shard across GPUs "
              "in this build", file=sys.stderr)
        sys.exit(2)

    import new_fgssjoin_amd as fg
    # fgss_selfjoin returns only after its stream has drained and the pairs are in
    # host memory, so each timed step is already bracketed by device synchronization.
    # torch is deliberately not imported: it would load a second HIP runtime.
    sync = fg.device_synchronize

    t0 = time.perf_counter()
    coll = fg.Collection.generate(fg.GEN_TITLES, N_RECORDS, SEED, device=True)
    info = coll.info
    log(f"generated+tokenized {info.n_sets} records, {info.n_tokens} tokens, {info.n_terms} terms "
        f"in {time.perf_counter() - t0:.1f} s")
    coll.upload()

    def step():
        return [coll.selfjoin(t, table_slots=args.table_slots, load_pct=args.load_pct) for t in THRESHOLDS]

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    sync()
    last = None
    kms = []
    with gc_paused():
        t0 = time.perf_counter()
        for i in range(args.steps):
            last = step()
            kms.append([p.stats["ms_kernel"] for p in last])
        sync()
        dt = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {dt:.3f} s")
    cand = sum(p.stats["candidates"] for p in last)
    ms_per_step = dt * 1e3 / args.steps
    # k_join launch duration per threshold: mean over the timed steps (HIP events on
    # the join's own stream around the launch)
    kms_t = [sum(k[i] for k in kms) / len(kms) for i in range(len(THRESHOLDS))]
    per_t = {str(t): {"join_ms": round(p.stats["ms_total"], 3), "kernel_ms": round(km, 3),
                      "index_ms": round(p.stats["ms_index"], 3), "pairs": len(p),
                      "candidates": p.stats["candidates"], "postings": p.stats["postings_scanned"],
                      "verify_tokens": p.stats["verify_tokens"], "sig_checks": p.stats["signature_checks"],
                      "verified": p.stats["verified"], "phase_cycles": p.stats["phase_cycles"]}
             for t, p, km in zip(THRESHOLDS, last, kms_t)}
    out = {
        "metric": "join ms + candidates verified/s (bit-exact); batched & served p99 vs plain",
        "value": cand / (ms_per_step / 1e3),
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"plain self-join, {N_RECORDS} synthetic titles, q=3, t={list(THRESHOLDS)}",
                   "records": info.n_sets, "tokens": info.n_tokens, "terms": info.n_terms,
                   "per_threshold": per_t},
    }
    counters = None
    if not args.no_cpu:
        log("cpu baseline: running the oracle on the same step ...")
        cb, exact, counters = cpu_baseline(coll, last)
        out["cpu_baseline"] = cb
        out["bit_exact_vs_oracle"] = exact
    else:
        out["cpu_baseline"] = None
    kernel_s = sum(kms_t) / 1e3
    n_launch = len(THRESHOLDS)
    if counters is not None:
        kbytes = sum(alg_bytes_join_kernel(c) for c in counters)
        basis = "algorithmic bytes of the reference algorithm, counted by the oracle"
        join_s = sum(p.stats["ms_total"] for p in last) / 1e3
        join_frac = sum(alg_bytes_join(c) for c in counters) / join_s / (HBM_PEAK_GBS * 1e9)
    else:
        kbytes = sum(design_bytes(p.stats) for p in last)
        basis = "bytes this build's kernel reads (oracle skipped)"
        join_frac = None
    achieved = kbytes / kernel_s / 1e9 if kernel_s > 0 else 0.0
    prof = pmc_summary()
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("traffic_bytes_per_launch"),
                       "kernel": "k_join", "launches_per_step": n_launch,
                       "bytes_per_launch": kbytes / n_launch, "kernel_ms_per_launch": kernel_s * 1e3 / n_launch,
                       "basis": basis, "join_frac": join_frac,
                       "lds_bank_conflict_ratio": prof.get("lds_bank_conflict_ratio"),
                       "valu_busy": prof.get("valu_busy_max"),
                       "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT) if prof else None}
    if args.batch_jobs > 0:
        out["batched"] = batched_leg(fg, args.batch_jobs, reps=3)
    if args.serve_requests > 0:
        out["served"] = served_leg(fg, args.serve_index, args.serve_requests, args.serve_slots)
    if args.large_records > 0:
        out["large"] = large_leg(fg, args.large_records, 0 if args.no_cpu else args.large_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
