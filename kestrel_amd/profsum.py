"""Summaries of rocprofv3 output (its SQLite results database) for the step's kernels.

    python -m kestrel_amd.profsum trace RESULTS_DB [--bench BENCH_LOG]
    python -m kestrel_amd.profsum pmc FETCH_DB WRITE_DB [--bench BENCH_LOG]
    python -m kestrel_amd.profsum counters DB [--bench BENCH_LOG] [--kernels a,b]

`trace` tabulates a --kernel-trace run: dispatches, total and mean duration
per kernel.  `pmc` tabulates the FETCH_SIZE and WRITE_SIZE counters (KiB per
dispatch) of two --pmc runs.  `counters` tabulates every counter of one --pmc
run (mean per productive dispatch) and, where SQ_INSTS_VALU and SQ_WAVES are
present, the VALU instructions per wave.  With --bench, the JSON line that bench.py
printed in the profiled run names how many launches of each kernel did work
in one pass; bench.py runs the same steps twice (timed, then profiled with HIP
events), so the 2 x N longest dispatches (largest FETCH_SIZE for `pmc`) are
the productive ones, and their mean is set beside bench.py's event mean.  The
rest are launches that only checked the control block and returned (after a
rejected check or at the end of a run).
"""
from __future__ import annotations

import argparse
import json
import re
import sqlite3
import statistics


def short_name(name: str) -> str:
    """'void kst::stage_kernel<2, true, 1>(kst::Dev, kst::Params)' -> 'stage_kernel<2>'."""
    n = name.split("(")[0].replace("void ", "").strip()
    n = n.split("::")[-1]
    m = re.match(r"(\w+)<(\d+)", n)
    return f"{m.group(1)}<{m.group(2)}>" if m else n


def bench_line(path: str) -> dict:
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("{") and '"kernels"' in line:
                return json.loads(line)
    raise ValueError(f"no bench.py JSON line in {path}")


def productive_counts(bench: dict, passes: int = 2) -> dict:
    """Launches that did work per kernel (short name), over `passes` passes."""
    out: dict = {}
    for r in bench["kernels"]:
        base = r["kernel"].replace(" (H2)", "")
        out[base] = out.get(base, 0) + passes * r["launches"]
    return out


def bench_means(bench: dict) -> dict:
    """Launch-weighted event mean (us) per kernel short name."""
    acc: dict = {}
    for r in bench["kernels"]:
        base = r["kernel"].replace(" (H2)", "")
        t, n = acc.get(base, (0.0, 0))
        acc[base] = (t + r["mean_us"] * r["launches"], n + r["launches"])
    return {k: t / n for k, (t, n) in acc.items() if n}


def trace_table(db: str, bench: dict | None = None) -> str:
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels").fetchall()
    dur: dict = {}
    for n, s, e in rows:
        dur.setdefault(short_name(n), []).append((e - s) / 1e3)
    total = sum(sum(v) for v in dur.values())
    prod = productive_counts(bench) if bench else {}
    ev = bench_means(bench) if bench else {}
    out = ["| kernel | dispatches | total ms | share | mean us (all) | productive | mean us (productive) | event mean us |",
           "|---|---|---|---|---|---|---|---|"]
    for k, v in sorted(dur.items(), key=lambda kv: -sum(kv[1])):
        if sum(v) < 0.002 * total:
            continue
        n = prod.get(k)
        pm = statistics.mean(sorted(v)[-n:]) if n and n <= len(v) else None
        out.append(f"| `{k}` | {len(v)} | {sum(v) / 1e3:.2f} | {100 * sum(v) / total:.1f}% | "
                   f"{statistics.mean(v):.2f} | {n if n else '-'} | "
                   f"{f'{pm:.2f}' if pm is not None else '-'} | "
                   f"{f'{ev[k]:.2f}' if k in ev else '-'} |")
    rows.sort(key=lambda r: r[1])
    span = (rows[-1][2] - rows[0][1]) / 1e6
    out.append(f"\nkernel time {total / 1e3:.1f} ms over a span of {span:.1f} ms")
    return "\n".join(out)


def _counter(db: str, counter: str) -> dict:
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, dispatch_id, value from counters_collection "
                       "where counter_name = ?", (counter,)).fetchall()
    per: dict = {}
    for n, did, v in rows:
        d = per.setdefault(short_name(n), {})
        d[did] = d.get(did, 0.0) + v
    return per


def pmc_table(fetch_db: str, write_db: str, bench: dict | None = None,
              fetch_scale: float = 2.0, write_scale: float = 1.0) -> str:
    F, W = _counter(fetch_db, "FETCH_SIZE"), _counter(write_db, "WRITE_SIZE")
    prod = productive_counts(bench) if bench else {}
    out = ["| kernel | dispatches | productive | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes per launch |",
           "|---|---|---|---|---|---|"]
    for k in sorted(F, key=lambda k: -sum(F[k].values())):
        fv = sorted(F[k].values())
        wv = sorted(W.get(k, {}).values())
        n = prod.get(k)
        if n and n <= len(fv):
            fm = statistics.mean(fv[-n:])
            wm = statistics.mean(wv[-n:]) if len(wv) >= n else 0.0
        else:
            fm = statistics.mean(fv)
            wm = statistics.mean(wv) if wv else 0.0
        traffic = (fm * fetch_scale + wm * write_scale) * 1024
        out.append(f"| `{k}` | {len(fv)} | {n if n else '-'} | {fm:.1f} | {wm:.1f} | {traffic:.4g} |")
    out.append(f"\nHBM bytes = (FETCH_SIZE x {fetch_scale} + WRITE_SIZE x {write_scale}) x 1024, "
               "means over the productive dispatches")
    return "\n".join(out)


def counter_table(db: str, bench: dict | None = None, kernels=None) -> str:
    con = sqlite3.connect(db)
    names = [r[0] for r in con.execute("select distinct counter_name from counters_collection")]
    data = {c: _counter(db, c) for c in names}
    prod = productive_counts(bench) if bench else {}
    key = "SQ_WAVES" if "SQ_WAVES" in data else names[0]
    ks = sorted(data[key], key=lambda k: -sum(data[key][k].values()))
    if kernels:
        ks = [k for k in ks if k in kernels]
    out = ["| kernel | dispatches | " + " | ".join(names) + " | VALU insts / wave |",
           "|---|---|" + "---|" * len(names) + "---|"]
    for k in ks:
        # productive dispatches: the n with the most waves (or the first counter)
        ids = sorted(data[key][k], key=lambda i: data[key][k][i])
        n = prod.get(k)
        sel = ids[-n:] if n and n <= len(ids) else ids
        means = {c: statistics.mean(data[c][k].get(i, 0.0) for i in sel) for c in names}
        vpw = (means["SQ_INSTS_VALU"] / means["SQ_WAVES"]
               if means.get("SQ_WAVES") and "SQ_INSTS_VALU" in means else None)
        out.append(f"| `{k}` | {len(sel)} | " + " | ".join(f"{means[c]:.4g}" for c in names)
                   + f" | {f'{vpw:.1f}' if vpw is not None else '-'} |")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--bench")
    m = sub.add_parser("pmc")
    m.add_argument("fetch_db")
    m.add_argument("write_db")
    m.add_argument("--bench")
    c = sub.add_parser("counters")
    c.add_argument("db")
    c.add_argument("--bench")
    c.add_argument("--kernels")
    a = ap.parse_args(argv)
    bench = bench_line(a.bench) if a.bench else None
    if a.cmd == "trace":
        print(trace_table(a.db, bench))
    elif a.cmd == "counters":
        print(counter_table(a.db, bench, a.kernels.split(",") if a.kernels else None))
    else:
        print(pmc_table(a.fetch_db, a.write_db, bench))


if __name__ == "__main__":
    main()
