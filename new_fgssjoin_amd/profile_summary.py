"""Summaries of rocprofv3 CSV output (kernel statistics and PMC counters) as the
short text files kept beside the benchmark.

    python3 new_fgssjoin_amd/profile_summary.py stats <kernel_stats.csv>
    python3 new_fgssjoin_amd/profile_summary.py trace <kernel_trace.csv> --kernel k_join --last 2
    python3 new_fgssjoin_amd/profile_summary.py pmc --kernel k_join --last 3 \\
        --copy-fetch <copy FETCH_SIZE csv> --copy-write <copy WRITE_SIZE csv> \\
        --fetch <FETCH_SIZE csv> --write <WRITE_SIZE csv> [--lds <SQ_LDS_* csv>] [--sq <SQ_INSTS_* csv>]

`pmc` scales FETCH_SIZE and WRITE_SIZE (KiB per dispatch) by what the same counter
reads for the 1 GiB copy of copy_calibration.py, takes the last `--last` dispatches
of the kernel (the timed step of a run whose warm-up absorbed buffer sizing), and
prints per-dispatch HBM bytes plus the LDS bank-conflict ratio
SQ_LDS_BANK_CONFLICT / (SQ_LDS_IDX_ACTIVE - SQ_LDS_BANK_CONFLICT) when given, and
the VALU issue utilisation from SQ_INSTS_VALU when given.
The final lines are `key value` pairs that bench.py reads.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re

COPY_BYTES = 1 << 30


def short(name: str) -> str:
    name = re.sub(r"^void ", "", name).replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*$", "", name).replace("fgss::", "")


def stats(path: str, top: int = 16):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("| kernel | calls | avg us | total ms | % |")
    print("|---|---:|---:|---:|---:|")
    for r in rows[:top]:
        print(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
              f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} |")


def trace(path: str, kernel: str, last: int):
    """durations of the last `last` dispatches of one kernel, from a kernel_trace.csv"""
    rows = [r for r in csv.DictReader(open(path)) if kernel in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    rows = rows[-last:]
    print(f"| {kernel} dispatch | duration ms | grid threads |")
    print("|---:|---:|---:|")
    tot = 0.0
    for r in rows:
        ms = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        tot += ms
        print(f"| {r['Dispatch_Id']} | {ms:.3f} | {r['Grid_Size_X']} |")
    print()
    print(f"kernel_ms_per_launch {tot / max(len(rows), 1):.6g}")


def _per_dispatch(path: str, kernel: str):
    """{counter: [values in dispatch order]} of one kernel"""
    out = collections.defaultdict(list)
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    for r in rows:
        if kernel in r["Kernel_Name"]:
            out[r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def _copy_factor(path: str, counter: str) -> float:
    vals = _per_dispatch(path, "copyBuffer")[counter]
    kib = sum(vals) / len(vals)
    return COPY_BYTES / (kib * 1024.0)


def pmc(a):
    ff = _copy_factor(a.copy_fetch, "FETCH_SIZE")
    fw = _copy_factor(a.copy_write, "WRITE_SIZE")
    fetch = _per_dispatch(a.fetch, a.kernel)["FETCH_SIZE"][-a.last:]
    write = _per_dispatch(a.write, a.kernel)["WRITE_SIZE"][-a.last:]
    print(f"copy calibration: 1 GiB copy reads FETCH_SIZE x{ff:.3f}, WRITE_SIZE x{fw:.3f} of its true bytes")
    print()
    print(f"| {a.kernel} dispatch | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes (scaled) |")
    print("|---:|---:|---:|---:|")
    tot = 0.0
    for k, (f, w) in enumerate(zip(fetch, write)):
        b = f * 1024 * ff + w * 1024 * fw
        tot += b
        print(f"| {k} | {f:.0f} | {w:.0f} | {b:.4g} |")
    n = max(len(fetch), 1)
    lines = [f"traffic_bytes_per_launch {tot / n:.6g}"]
    if a.lds:
        c = _per_dispatch(a.lds, a.kernel)
        bank = sum(c["SQ_LDS_BANK_CONFLICT"][-a.last:])
        active = sum(c["SQ_LDS_IDX_ACTIVE"][-a.last:])
        insts = sum(c["SQ_INSTS_LDS"][-a.last:])
        ratio = bank / (active - bank) if active > bank else float("nan")
        print()
        print(f"LDS: {insts:.4g} instructions, {active:.4g} indexed-active cycles, {bank:.4g} bank-conflict "
              f"cycles over the last {a.last} dispatches")
        lines.append(f"lds_bank_conflict_ratio {ratio:.6g}")
        lines.append(f"lds_bank_conflict_cycles_per_inst {bank / insts if insts else float('nan'):.6g}")
    if a.sq:
        # VALU issue: a wave64 VALU instruction occupies one of the chip's SIMD16 units
        # for 4 cycles; busy = instructions x 4 / (SIMDs x dispatch cycles)
        c = _per_dispatch(a.sq, a.kernel)
        dur = _durations(a.sq, a.kernel)[-a.last:]
        valu = c["SQ_INSTS_VALU"][-a.last:]
        print()
        print(f"| {a.kernel} dispatch | us | VALU insts | SALU insts | LDS insts | VALU busy |")
        print("|---:|---:|---:|---:|---:|---:|")
        busy = []
        for k, (d, v, sa, ld) in enumerate(zip(dur, valu, c["SQ_INSTS_SALU"][-a.last:], c["SQ_INSTS_LDS"][-a.last:])):
            b = v * 4 / (SIMDS * d * 1e-9 * CLOCK_HZ) if d else float("nan")
            busy.append(b)
            print(f"| {k} | {d / 1e3:.1f} | {v:.4g} | {sa:.4g} | {ld:.4g} | {b:.2f} |")
        lines.append(f"valu_busy_max {max(busy):.4g}")
    print()
    for ln in lines:
        print(ln)


SIMDS = 256 * 4           # MI355X: 256 CUs of 4 SIMD16 units
CLOCK_HZ = 2.4e9          # peak engine clock


def _durations(path: str, kernel: str):
    """dispatch durations (ns) of one kernel, in dispatch order"""
    seen = {}
    for r in csv.DictReader(open(path)):
        if kernel in r["Kernel_Name"]:
            seen[int(r["Dispatch_Id"])] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return [seen[k] for k in sorted(seen)]


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=16)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--kernel", default="k_join")
    t.add_argument("--last", type=int, default=3)
    p = sub.add_parser("pmc")
    p.add_argument("--kernel", default="k_join")
    p.add_argument("--last", type=int, default=3)
    p.add_argument("--copy-fetch", required=True)
    p.add_argument("--copy-write", required=True)
    p.add_argument("--fetch", required=True)
    p.add_argument("--write", required=True)
    p.add_argument("--lds")
    p.add_argument("--sq", help="SQ_INSTS_VALU / SQ_INSTS_SALU / SQ_INSTS_LDS csv")
    a = ap.parse_args()
    if a.cmd == "stats":
        stats(a.csv, a.top)
    elif a.cmd == "trace":
        trace(a.csv, a.kernel, a.last)
    else:
        pmc(a)


if __name__ == "__main__":
    main()
