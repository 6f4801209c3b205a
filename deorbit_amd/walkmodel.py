"""CPU model of the pair kernels' column walk: candidates per pair in range and
the turns a wave's loops take, on a state taken from the device.

    python -m deorbit_amd.walkmodel dump --workload layered_4m --steps 500 --out state.npz   # GPU
    python -m deorbit_amd.walkmodel model state.npz [--cell-factor 1.5] [--pass force] [--waves 150]

``dump`` runs the workload's first ``--steps`` steps in DSPH_MIXED and saves the
positions (relative to their minimum, FP32) and smoothing lengths.  ``model``
rebuilds the sort grid of the device (columns of cell_factor h_init across x and
y, slices of a quarter of that along z, 48-slice blocks with their bound of h),
replays the walk of k_force (reach h_i + the block bound of h, range test
r < h_i + h_j) or of the density (reach and range 2 h_i) for every lane of
randomly chosen waves of 64 consecutive particles, and reports per particle the
candidates visited and the pairs in range, and per wave the turns of

  * ``slot``: the per-slot walk, 4 candidates per turn, a turn per 4 slots of the
    wave's longest run in each column;
  * ``mask``: the run's pairs in range only (the bit-mask walk of the kernels):
    turns = the wave's largest count of pairs in range per column;
  * ``flat``: each lane's pairs in range queued across all its columns and run
    one per turn: turns = the wave's largest count of pairs in range per lane;
  * ``ideal``: every lane busy: the wave's pairs in range / 64.

``--mirror`` replays the mirrored column order of the kernels (``y``: lanes in
the upper half of their column along y take the columns in mirrored order along
y; ``xy``: along x as well; ``none``: the common order): a turn of the walk is a
step of the stencil loop, whatever column each lane visits at that step, and
``columns_per_step`` counts the distinct columns a wave's lanes visit at a step
(the spread of its loads across memory).
"""
from __future__ import annotations

import argparse
import math

import numpy as np

SLICE_BLOCK = 48          # slices per block (kSliceBlock in csrc/common.h)


def dump(workload: str, steps: int, out: str):
    from . import dsph, ics
    ic = getattr(ics, workload)()
    s = dsph.make_system(ic, dsph.MIXED)
    s.init()
    done = 0
    while done < steps:
        k = min(50, steps - done)
        s.run(k, 0.0, True)
        done += k
    pos, h = s.get("pos"), s.get("hh")
    lo = pos.min(0)
    np.savez_compressed(out, pos=(pos - lo).astype(np.float32), h=h.astype(np.float32), lo=lo,
                        h_init=ic["params"]["h_init"], steps=steps, workload=workload)
    s.close()


def model(path: str, cell_factor: float, which: str, nwaves: int, seed: int = 1, mirror: str = "none") -> dict:
    d = np.load(path)
    P = d["pos"].astype(np.float64)
    H = d["h"].astype(np.float64)
    h0 = float(d["h_init"])
    w = cell_factor * h0
    wz = w / 4
    dim = np.floor(P.max(0) / np.array([w, w, wz])).astype(int) + 1
    C = np.minimum(np.floor(P / np.array([w, w, wz])).astype(int), dim - 1)
    key = (C[:, 0] * dim[1] + C[:, 1]) * dim[2] + C[:, 2]
    o = np.argsort(key, kind="stable")
    P, H, C, K = P[o], H[o], C[o], key[o]
    cstart = np.searchsorted(K, np.arange(int(dim.prod()) + 1))
    bkey = (C[:, 0] * dim[1] + C[:, 1]) * dim[2] + (C[:, 2] // SLICE_BLOCK) * SLICE_BLOCK
    ub, inv = np.unique(bkey, return_inverse=True)
    bmax = np.zeros(len(ub))
    np.maximum.at(bmax, inv, H)
    bound = dict(zip(ub.tolist(), bmax.tolist()))
    hmax = H.max()
    rng = np.random.default_rng(seed)
    waves = rng.choice(len(P) // 64, min(nwaves, len(P) // 64), replace=False)
    tot = dict(cand=0, inr=0, slot=0, mask=0, flat=0, ideal=0, steps=0, colsteps=0)
    for wv in waves:
        cols = {}
        visited = {}
        inr_wave = 0
        lane_inr = np.zeros(64, int)
        for lane, i in enumerate(range(wv * 64, wv * 64 + 64)):
            hi = H[i]
            reach = 2 * hi if which == "density" else hi + hmax
            cx, cy, cz = C[i]
            ox, oy, ozs = P[i, 0] - cx * w, P[i, 1] - cy * w, P[i, 2] - cz * wz
            rx = min(int(math.ceil(reach / w)), 16)
            mx = -1 if mirror == "xy" and ox >= 0.5 * w else 1
            my = -1 if mirror in ("y", "xy") and oy >= 0.5 * w else 1
            for ex in range(-rx, rx + 1):
                dx = mx * ex
                nx = cx + dx
                if not 0 <= nx < dim[0]:
                    continue
                gx = max(dx * w - ox, ox - (dx + 1) * w, 0.0)
                if gx * gx >= reach * reach:
                    continue
                for ey in range(-rx, rx + 1):
                    dy = my * ey
                    ny = cy + dy
                    if not 0 <= ny < dim[1]:
                        continue
                    gy = max(dy * w - oy, oy - (dy + 1) * w, 0.0)
                    gxy = gx * gx + gy * gy
                    if gxy >= reach * reach:
                        continue
                    rz = math.sqrt(reach * reach - gxy)
                    z0 = max(cz + math.floor((ozs - rz) / wz - 1e-3), 0)
                    z1 = min(cz + math.floor((ozs + rz) / wz + 1e-3), dim[2] - 1)
                    if z0 > z1:
                        continue
                    col = (nx * dim[1] + ny) * dim[2]
                    if which != "density":
                        f0, f1 = z0 - z0 % SLICE_BLOCK, z1 - z1 % SLICE_BLOCK
                        rb = hi + max(bound.get(col + f0, 0.0), bound.get(col + f1, 0.0))
                        if gxy >= rb * rb:
                            continue
                        rz = math.sqrt(rb * rb - gxy)
                        z0 = max(z0, cz + math.floor((ozs - rz) / wz - 1e-3))
                        z1 = min(z1, cz + math.floor((ozs + rz) / wz + 1e-3))
                        if z0 > z1:
                            continue
                    k0, k1 = cstart[col + z0], cstart[col + z1 + 1]
                    if k0 >= k1:
                        continue
                    r = np.linalg.norm(P[k0:k1] - P[i], axis=1)
                    if which == "density":
                        nin = int((r < 2 * hi).sum())
                    else:
                        nin = int(((r < hi + H[k0:k1]) & (r > 0)).sum())
                    cols.setdefault((ex, ey), np.zeros((64, 2), int))[lane] = (k1 - k0, nin)
                    visited.setdefault((ex, ey), set()).add((nx, ny))
                    inr_wave += nin
                    lane_inr[lane] += nin
                    tot["cand"] += k1 - k0
                    tot["inr"] += nin
        A = np.array(list(cols.values()))
        tot["slot"] += float((np.ceil(A[:, :, 0].max(1) / 4) * 4).sum()) / 4
        tot["mask"] += float(A[:, :, 1].max(1).sum())
        tot["flat"] += int(lane_inr.max())
        tot["ideal"] += math.ceil(inr_wave / 64)
        tot["steps"] += len(visited)
        tot["colsteps"] += sum(len(v) for v in visited.values())
    n = len(waves) * 64
    return {"particles": len(P), "waves": len(waves), "cell_factor": cell_factor, "pass": which, "mirror": mirror,
            "h_min": float(H.min()), "h_median": float(np.median(H)), "h_max": float(hmax),
            "candidates_per_particle": tot["cand"] / n, "pairs_per_particle": tot["inr"] / n,
            "candidates_per_pair": tot["cand"] / max(1, tot["inr"]),
            "slot_turns_per_wave": tot["slot"] / len(waves), "mask_turns_per_wave": tot["mask"] / len(waves),
            "flat_turns_per_wave": tot["flat"] / len(waves),
            "ideal_turns_per_wave": tot["ideal"] / len(waves),
            "mask_lane_use": tot["inr"] / 64 / max(1.0, tot["mask"]),
            "columns_per_step": tot["colsteps"] / max(1, tot["steps"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["dump", "model"])
    ap.add_argument("path", nargs="?")
    ap.add_argument("--workload", default="layered_4m")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--out", default="state.npz")
    ap.add_argument("--cell-factor", type=float, default=1.5)
    ap.add_argument("--pass", dest="which", default="force", choices=["force", "density"])
    ap.add_argument("--waves", type=int, default=150)
    ap.add_argument("--mirror", default="none", choices=["none", "y", "xy"])
    a = ap.parse_args()
    if a.what == "dump":
        dump(a.workload, a.steps, a.out)
        return
    r = model(a.path, a.cell_factor, a.which, a.waves, mirror=a.mirror)
    for k, v in r.items():
        print(f"{k:26s} {v:.4g}" if isinstance(v, float) else f"{k:26s} {v}")


if __name__ == "__main__":
    main()
