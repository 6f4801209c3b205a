"""CPU model of the pair kernels' column walk (deorbit_amd/walkmodel.py) on a small
jittered lattice: the mirrored column orders visit the same candidates and pairs as
the common order (they only reorder the walk), they take fewer pair turns, and a
wave's lanes spread over at most twice (mirror along y) or four times (x and y) as
many columns per step as in the common order."""
import numpy as np

from deorbit_amd import walkmodel


def lattice_state(tmp_path, n=24, h_init=1.3):
    rng = np.random.default_rng(7)
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    pos = (g + 0.05 * rng.standard_normal(g.shape)).astype(np.float32)
    pos -= pos.min(0)
    h = (h_init * (1 + 0.1 * rng.random(len(pos)))).astype(np.float32)
    path = tmp_path / "state.npz"
    np.savez(path, pos=pos, h=h, lo=np.zeros(3), h_init=h_init, steps=0, workload="lattice")
    return str(path)


def test_mirrored_orders_visit_the_same_pairs(tmp_path):
    path = lattice_state(tmp_path)
    for which in ("force", "density"):
        r = {m: walkmodel.model(path, 1.45, which, 40, mirror=m) for m in ("none", "y", "xy")}
        for m in ("y", "xy"):
            assert r[m]["candidates_per_particle"] == r["none"]["candidates_per_particle"]
            assert r[m]["pairs_per_particle"] == r["none"]["pairs_per_particle"]
        # waves here often straddle two of the small lattice's columns
        c = {m: r[m]["columns_per_step"] for m in r}
        assert 1.0 <= c["none"] < c["y"] < c["xy"]
        assert c["y"] <= 2.0 * c["none"] + 1e-12 and c["xy"] <= 4.0 * c["none"] + 1e-12
        # mirroring evens out the lanes' pairs per step: fewer pair turns, never
        # fewer than with every lane busy
        t = {m: r[m]["mask_turns_per_wave"] for m in r}
        assert t["xy"] < t["y"] < t["none"]
        for m in r:
            assert t[m] >= r[m]["ideal_turns_per_wave"]
