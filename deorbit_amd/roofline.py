"""Algorithmic HBM bytes per particle for each kernel of the step.

Counting rules: every per-particle array a kernel touches is counted once per
launch (neighbour records are credited once, as if every re-read were served
from cache), cell-table reads are ignored, and the sort counts only its key and
value traffic.  ``achieved GB/s = bytes_per_particle * n / kernel time``.
The figures follow the data layout of deorbit_amd/csrc/common.h:

mixed (R = float): positions, velocities, u and mass in FP64; everything else
FP32; neighbour records are float4.  fp64 (R = double): every R array doubles
and records are double4.
"""
from __future__ import annotations

MI355X_HBM_GBPS = 8000.0          # published peak HBM3E bandwidth
MI355X_FP32_TFLOPS = 157.3        # published dense FP32 (packed) vector peak
MI355X_FP64_TFLOPS = 78.6


def kernel_bytes(precision: str = "mixed") -> dict:
    r = 4 if precision == "mixed" else 8         # bytes of the pair/state real R
    rec = 4 * r                                  # one neighbour record (4 x R)
    d = 8
    out = {}
    # kick + drift + bbox: R info, x y z, v xyz, acc xyz; W v xyz, x y z
    out["kick_drift"] = 4 + 6 * d + 3 * r + 6 * d
    # keys: R info, x y z; W key, value
    out["keys"] = 4 + 3 * d + 8
    # radix sort, per pass: hist reads keys; scatter reads and writes key+value
    out["sort_pass"] = 4 + 8 + 8
    # permute: carried state R+W (8 doubles, h D eps_act, S6, info, orig) + keys/vals/dev_of R
    #          + perm W + old derivatives R+W (dudt, dSdt6) + density record W + cellc, h0 W
    carried = 8 * d + 3 * r + 6 * r + 4 + 4
    out["permute"] = 2 * carried + 12 + 4 + 2 * 7 * r + rec + 4 + r
    # density iteration: neighbour record once + own (h, cellc, info, mass) R; W rho, h
    out["density"] = rec + r + 4 + 4 + d + 2 * r
    # last iteration adds the EOS epilogue: R u, S6, D, vel; W P cs csig Pf dscale Rart sigmax + 5 records
    out["density_last"] = out["density"] + (d + 7 * r + 3 * d) + (7 * r + 5 * rec)
    if precision == "mixed":
        # DSPH_MIXED's FP64-term density: the FP32 screening record and the FP64 (x, y, z, m)
        # record of each neighbour once, the own position in FP64 besides the FP32 iteration's reads
        out["density64"] = out["density"] + 32 + 3 * d
        out["density64_last"] = out["density_last"] + 32 + 3 * d
        # the permutation also writes the FP64 records
        out["permute64"] = out["permute"] + 32
    # force: 5 neighbour records once + own (cellc, info, S6, dscale) R; W acc3 dudt dSdt6 eps_work
    out["force"] = 5 * rec + 4 + 4 + 6 * r + r + (3 * r + r + 6 * r + r)
    if precision != "mixed":
        out["force"] += 3 * d                    # own absolute position for the cell offset
    # finish: R info, sigmax, D, eps_act, rho, v3, acc3, u, dudt_old, dudt, S6, dSdt_old6, dSdt6
    #         W D, v3, u, S6
    out["finish"] = (4 + 5 * r + 3 * d + 3 * r + d + 2 * r + 18 * r) + (r + 3 * d + d + 6 * r)
    return out


def density_round_bytes(precision: str = "mixed", density64: bool = True, iterations: int = 2) -> int:
    """Algorithmic bytes of one density round (`iterations` iterations, EOS epilogue in the last)."""
    k = kernel_bytes(precision)
    tag = "density64" if precision == "mixed" and density64 else "density"
    return (iterations - 1) * k[tag] + k[tag + "_last"]


def step_bytes(precision: str = "mixed", passes: int = 3, iterations: int = 2, density64: bool = True) -> int:
    """Algorithmic bytes of one full step (one density round of `iterations`).  density64:
    DSPH_MIXED's FP64 density records (False: DSPH_MIXED_FAST)."""
    k = kernel_bytes(precision)
    d64 = precision == "mixed" and density64
    return (k["kick_drift"] + k["keys"] + passes * k["sort_pass"] + (k["permute64"] if d64 else k["permute"])
            + density_round_bytes(precision, density64, iterations) + k["force"] + k["finish"])


# Algorithmic floating-point work of one particle-step: the ordered neighbour
# pairs inside 2h of a uniform lattice particle, (4/3) pi (2 eta)^3 - 1 ~ 73 at
# eta = 1.3, times the flops of one pair in each pass: ~25 per density term (two
# iterations per round) and ~172 for the fused strain + momentum + energy pair
# (kernel slope, velocity gradient, viscosity, artificial stress, deviatoric
# stress).  Rejected candidates are overhead, not counted.
PAIRS_PER_PARTICLE = 73
FLOPS_DENSITY_PAIR = 25
FLOPS_FORCE_PAIR = 172


def step_flops(iterations: int = 2) -> int:
    return PAIRS_PER_PARTICLE * (iterations * FLOPS_DENSITY_PAIR + FLOPS_FORCE_PAIR)


def force_flops() -> int:
    """Algorithmic flops of the force pass per particle."""
    return PAIRS_PER_PARTICLE * FLOPS_FORCE_PAIR


def density_flops(iterations: int = 2) -> int:
    """Algorithmic flops of one density round per particle (FP64 in DSPH_MIXED)."""
    return PAIRS_PER_PARTICLE * iterations * FLOPS_DENSITY_PAIR


def _code_only(text: str) -> str:
    """C++ source without comments and with whitespace runs collapsed (string
    literals are kept as written)."""
    import re
    tok = re.compile(r'"(?:\\.|[^"\\])*"|//[^\n]*|/\*.*?\*/', re.S)
    text = tok.sub(lambda m: m.group(0) if m.group(0).startswith('"') else " ", text)
    return " ".join(text.split())


def source_digest() -> str:
    """Short digest of the device sources' code (comments and layout ignored);
    ties a recorded counter profile to a build."""
    import hashlib
    import os
    csrc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(csrc)):
        if name.endswith((".h", ".hip")):
            with open(os.path.join(csrc, name), encoding="utf-8") as f:
                h.update(name.encode() + b"\0" + _code_only(f.read()).encode())
    return h.hexdigest()[:16]


def _recorded(folder: str, tag: str, kernel: str, workload: str) -> dict | None:
    """Key/value pairs of a ``<tag>:`` line in any summary under `folder` recorded
    for this kernel, workload and the current source digest."""
    import glob
    import os
    import re
    digest = source_digest()
    paths = sorted(glob.glob(os.path.join(folder, "*.md")) + glob.glob(os.path.join(folder, "*.txt")))
    for path in paths:
        with open(path) as f:
            lines = [ln for ln in f if ln.startswith(tag + ":")]
        for line in lines:
            # values run to the next " key=" (kernel names hold spaces)
            kv = dict(re.findall(r"(\w+)=(.*?)(?=\s+\w+=|\s*$)", line[len(tag) + 1:].strip()))
            if kv.get("kernel") == kernel and kv.get("workload") == workload and kv.get("digest") == digest:
                return kv
    return None


def recorded_traffic(folder: str, kernel: str, workload: str) -> dict | None:
    """HBM bytes per launch recorded by ``rocpd_summary traffic`` (a ``traffic:``
    line in any summary under `folder`) for this kernel, workload and source
    digest; None when absent or recorded for other sources."""
    kv = _recorded(folder, "traffic", kernel, workload)
    if kv is None:
        return None
    return {"bytes": float(kv["bytes"]), "fetch_factor": float(kv["fetch_factor"]),
            "write_factor": float(kv["write_factor"])}


def recorded_pipes(folder: str, kernel: str, workload: str) -> dict | None:
    """Pipe utilisation recorded by ``rocpd_summary pipes`` (VALU issue, lane
    utilisation, TA and TD busy, L1 hit) for this kernel, workload and source
    digest; None when absent or recorded for other sources."""
    kv = _recorded(folder, "pipes", kernel, workload)
    if kv is None:
        return None
    return {k: float(kv[k]) for k in ("valu_issue", "lane_util", "ta_busy", "td_busy", "l1_hit")}


def limiter(pipes: dict | None) -> str:
    """The busiest pipe of a kernel by its counters: "valu" (vector ALU issue) or
    "vmem" (the vector memory path, texture address / data units), or "unmeasured"."""
    if pipes is None:
        return "unmeasured"
    return "valu" if pipes["valu_issue"] >= max(pipes["ta_busy"], pipes["td_busy"]) else "vmem"
