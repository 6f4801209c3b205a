"""Summarise rocprofv3 SQLite output (rocpd databases) as text.

    python -m deorbit_amd.rocpd_summary kernels <results.db> [--match NAME] [--tail N]
    python -m deorbit_amd.rocpd_summary counters <results.db> [--match NAME] [--tail N]
    python -m deorbit_amd.rocpd_summary traffic --fetch F.db --write W.db \
        --copy-fetch CF.db --copy-write CW.db --copy-bytes B --match NAME \
        --workload TAG --bench-json profiled_run.json [--tail N]
    python -m deorbit_amd.rocpd_summary pipes --valu V.db --mem M.db --match NAME \
        --workload TAG --bench-json profiled_run.json [--tail N]

``kernels`` prints per-kernel dispatch count, total and mean duration (the
``--stats`` table); ``counters`` prints, per kernel, the mean per-dispatch value of
each collected counter (``--pmc`` runs).  ``traffic`` turns one FETCH_SIZE run and
one WRITE_SIZE run (KiB per dispatch) into HBM bytes per launch of one kernel,
each counter scaled by bytes-moved / counter-reading of a plain device copy of
known size, and prints a ``traffic:`` line keyed by kernel, workload and the
device-source digest that the profiled bench run reported.  ``pipes`` derives
the VALU issue rate, lane utilisation, TA/TD busy fractions and the L1 hit rate
of one kernel from a VALU counter run and a memory-pipe counter run and prints
a ``pipes:`` line keyed the same way.  ``--tail N`` keeps only the last N
dispatches of each kernel (the bench's profiled steps come last in its run).
"""
from __future__ import annotations

import argparse
import re
import sqlite3
from collections import defaultdict


def short(name: str, width: int = 70) -> str:
    name = re.sub(r"\s+", " ", name)
    return name if len(name) <= width else name[: width - 3] + "..."


def kernels(db: str, match: str | None = None, tail: int = 0):
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, duration from kernels order by start").fetchall()
    per = defaultdict(list)
    for name, _, dur in rows:
        per[name].append(dur)
    if tail:
        per = {k: v[-tail:] for k, v in per.items()}
    total = sum(sum(v) for v in per.values()) or 1
    out = []
    for name, durs in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        if match and match not in name:
            continue
        tot = sum(durs)
        out.append((short(name), len(durs), tot / 1e6, tot / len(durs) / 1e3, 100.0 * tot / total))
    return out


def counters(db: str, match: str | None = None, tail: int = 0):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, dispatch_id, sum(value) from counters_collection "
                       "group by kernel_name, counter_name, dispatch_id order by dispatch_id").fetchall()
    acc = defaultdict(list)
    for kname, cname, _, v in rows:
        if match and match not in kname:
            continue
        acc[(short(kname), cname)].append(v)
    if tail:
        acc = {k: v[-tail:] for k, v in acc.items()}
    return [(k, c, len(v), sum(v) / len(v)) for (k, c), v in sorted(acc.items())]


def counter_mean(db: str, counter: str, match: str, tail: int = 0) -> tuple[float, int]:
    """Mean per-dispatch value (summed over instances) of a counter for kernels matching."""
    rows = [(k, n, v) for k, c, n, v in counters(db, None, tail) if c == counter and match in k]
    n = sum(r[1] for r in rows)
    if not n:
        raise SystemExit(f"no {counter} dispatches of a kernel matching {match!r} in {db}")
    return sum(r[1] * r[2] for r in rows) / n, n


def copy_kernel_kib(db: str, counter: str) -> float:
    """Largest mean per-dispatch counter value in a copy-probe run (the copy kernel)."""
    rows = [(v, k) for k, c, n, v in counters(db, None) if c == counter]
    return max(rows)[0]


def bench_digest(path: str | None) -> str:
    """The device-source digest a bench run reported in its JSON line."""
    import json
    digest = "unknown"
    if path:
        with open(path) as f:
            for line in f:
                line = line.strip()
                if line.startswith("{"):
                    digest = json.loads(line).get("roofline", {}).get("source_digest", digest)
    return digest


def traffic(a):
    fetch, nf = counter_mean(a.fetch, "FETCH_SIZE", a.match, a.tail)
    write, nw = counter_mean(a.write, "WRITE_SIZE", a.match, a.tail)
    ff = a.copy_bytes / (1024.0 * copy_kernel_kib(a.copy_fetch, "FETCH_SIZE"))
    wf = a.copy_bytes / (1024.0 * copy_kernel_kib(a.copy_write, "WRITE_SIZE"))
    total = fetch * 1024 * ff + write * 1024 * wf
    digest = bench_digest(a.bench_json)
    print(f"kernel {a.match}: FETCH_SIZE {fetch:.1f} KiB/dispatch over {nf} dispatches, "
          f"WRITE_SIZE {write:.1f} KiB/dispatch over {nw} dispatches")
    print(f"copy calibration ({a.copy_bytes} B read and written): FETCH factor {ff:.4f}, WRITE factor {wf:.4f}")
    print(f"HBM bytes per launch = {fetch:.1f}*1024*{ff:.4f} + {write:.1f}*1024*{wf:.4f} = {total:.4e}")
    print(f"traffic: kernel={a.match} workload={a.workload} digest={digest} bytes={total:.6e} "
          f"fetch_factor={ff:.4f} write_factor={wf:.4f}")


# MI355X: 256 CUs x 4 SIMDs, 8 XCDs; one wave64 VALU instruction occupies a SIMD
# for 4 cycles (16 lanes per cycle) unless the instruction issues at a faster rate
N_SIMD, N_XCD, N_TA = 1024, 8, 256


def pipes(a):
    """VALU issue, lane utilisation, TA/TD busy and L1 hit of one kernel (means over
    its dispatches).  GRBM_GUI_ACTIVE counts cycles of every XCD, so the kernel's
    own cycles are GRBM_GUI_ACTIVE / 8."""
    def m(db, c):
        return counter_mean(db, c, a.match, a.tail)[0]
    cyc = m(a.valu, "GRBM_GUI_ACTIVE") / N_XCD
    valu = 4.0 * m(a.valu, "SQ_INSTS_VALU") / (N_SIMD * cyc)
    lanes = m(a.valu, "SQ_THREAD_CYCLES_VALU") / (64.0 * m(a.valu, "SQ_ACTIVE_INST_VALU"))
    cyc_m = m(a.mem, "GRBM_GUI_ACTIVE") / N_XCD
    ta = m(a.mem, "TA_TA_BUSY") / N_TA / cyc_m
    td = m(a.mem, "TD_TD_BUSY") / N_TA / cyc_m
    l1 = 1.0 - m(a.mem, "TCP_TCC_READ_REQ") / m(a.mem, "TCP_TOTAL_CACHE_ACCESSES")
    digest = bench_digest(a.bench_json)
    print(f"kernel {a.match}: VALU issue {valu:.2f}, lane utilisation {lanes:.2f}, TA busy {ta:.2f}, "
          f"TD busy {td:.2f}, L1 hit {l1:.2f}")
    print(f"pipes: kernel={a.match} workload={a.workload} digest={digest} valu_issue={valu:.3f} "
          f"lane_util={lanes:.3f} ta_busy={ta:.3f} td_busy={td:.3f} l1_hit={l1:.3f}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["kernels", "counters", "traffic", "pipes"])
    ap.add_argument("db", nargs="?")
    ap.add_argument("--match")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--workload", default="graze_1m")
    ap.add_argument("--bench-json")
    ap.add_argument("--valu")
    ap.add_argument("--mem")
    ap.add_argument("--tail", type=int, default=0)
    a = ap.parse_args()
    if a.what == "traffic":
        traffic(a)
        return
    if a.what == "pipes":
        pipes(a)
        return
    if a.what == "kernels":
        print(f"{'kernel':72s} {'calls':>6s} {'total ms':>10s} {'mean us':>10s} {'%':>6s}")
        for name, cnt, tot, avg, pct in kernels(a.db, a.match, a.tail):
            print(f"{name:72s} {cnt:6d} {tot:10.3f} {avg:10.2f} {pct:6.2f}")
    else:
        print(f"{'kernel':72s} {'counter':14s} {'dispatches':>10s} {'mean value':>16s}")
        for k, c, n, v in counters(a.db, a.match, a.tail):
            print(f"{k:72s} {c:14s} {n:10d} {v:16.1f}")


if __name__ == "__main__":
    main()
