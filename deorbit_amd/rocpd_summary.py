"""Summarise rocprofv3 SQLite output (rocpd databases) as text.

    python -m deorbit_amd.rocpd_summary kernels <results.db> [--match NAME]
    python -m deorbit_amd.rocpd_summary counters <results.db> [--match NAME]
    python -m deorbit_amd.rocpd_summary traffic --fetch F.db --write W.db \
        --copy-fetch CF.db --copy-write CW.db --copy-bytes B --match NAME \
        --workload TAG --bench-json profiled_run.json

``kernels`` prints per-kernel dispatch count, total and mean duration (the
``--stats`` table); ``counters`` prints, per kernel, the mean per-dispatch value of
each collected counter (``--pmc`` runs).  ``traffic`` turns one FETCH_SIZE run and
one WRITE_SIZE run (KiB per dispatch) into HBM bytes per launch of one kernel,
each counter scaled by bytes-moved / counter-reading of a plain device copy of
known size, and prints a ``traffic:`` line keyed by kernel, workload and the
device-source digest that the profiled bench run reported.
"""
from __future__ import annotations

import argparse
import re
import sqlite3
from collections import defaultdict


def short(name: str, width: int = 70) -> str:
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= width else name[: width - 3] + "..."


def kernels(db: str, match: str | None = None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name "
                       "order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = []
    for name, cnt, tot, avg in rows:
        if match and match not in name:
            continue
        out.append((short(name), cnt, tot / 1e6, avg / 1e3, 100.0 * tot / total))
    return out


def counters(db: str, match: str | None = None):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, dispatch_id, sum(value) from counters_collection "
                       "group by kernel_name, counter_name, dispatch_id").fetchall()
    acc = defaultdict(list)
    for kname, cname, _, v in rows:
        if match and match not in kname:
            continue
        acc[(short(kname), cname)].append(v)
    return [(k, c, len(v), sum(v) / len(v)) for (k, c), v in sorted(acc.items())]


def counter_mean(db: str, counter: str, match: str) -> tuple[float, int]:
    """Mean per-dispatch value (summed over instances) of a counter for kernels matching."""
    rows = [(k, n, v) for k, c, n, v in counters(db, None) if c == counter and match in k]
    n = sum(r[1] for r in rows)
    if not n:
        raise SystemExit(f"no {counter} dispatches of a kernel matching {match!r} in {db}")
    return sum(r[1] * r[2] for r in rows) / n, n


def copy_kernel_kib(db: str, counter: str) -> float:
    """Largest mean per-dispatch counter value in a copy-probe run (the copy kernel)."""
    rows = [(v, k) for k, c, n, v in counters(db, None) if c == counter]
    return max(rows)[0]


def traffic(a):
    import json
    fetch, nf = counter_mean(a.fetch, "FETCH_SIZE", a.match)
    write, nw = counter_mean(a.write, "WRITE_SIZE", a.match)
    ff = a.copy_bytes / (1024.0 * copy_kernel_kib(a.copy_fetch, "FETCH_SIZE"))
    wf = a.copy_bytes / (1024.0 * copy_kernel_kib(a.copy_write, "WRITE_SIZE"))
    total = fetch * 1024 * ff + write * 1024 * wf
    digest = "unknown"
    if a.bench_json:
        with open(a.bench_json) as f:
            for line in f:
                line = line.strip()
                if line.startswith("{"):
                    digest = json.loads(line).get("roofline", {}).get("source_digest", digest)
    print(f"kernel {a.match}: FETCH_SIZE {fetch:.1f} KiB/dispatch over {nf} dispatches, "
          f"WRITE_SIZE {write:.1f} KiB/dispatch over {nw} dispatches")
    print(f"copy calibration ({a.copy_bytes} B read and written): FETCH factor {ff:.4f}, WRITE factor {wf:.4f}")
    print(f"HBM bytes per launch = {fetch:.1f}*1024*{ff:.4f} + {write:.1f}*1024*{wf:.4f} = {total:.4e}")
    print(f"traffic: kernel={a.match} workload={a.workload} digest={digest} bytes={total:.6e} "
          f"fetch_factor={ff:.4f} write_factor={wf:.4f}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["kernels", "counters", "traffic"])
    ap.add_argument("db", nargs="?")
    ap.add_argument("--match")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--workload", default="graze_1m")
    ap.add_argument("--bench-json")
    a = ap.parse_args()
    if a.what == "traffic":
        traffic(a)
        return
    if a.what == "kernels":
        print(f"{'kernel':72s} {'calls':>6s} {'total ms':>10s} {'mean us':>10s} {'%':>6s}")
        for name, cnt, tot, avg, pct in kernels(a.db, a.match):
            print(f"{name:72s} {cnt:6d} {tot:10.3f} {avg:10.2f} {pct:6.2f}")
    else:
        print(f"{'kernel':72s} {'counter':14s} {'dispatches':>10s} {'mean value':>16s}")
        for k, c, n, v in counters(a.db, a.match):
            print(f"{k:72s} {c:14s} {n:10d} {v:16.1f}")


if __name__ == "__main__":
    main()
