"""IC generators and the reference file formats (CPU only)."""
import os

import numpy as np
import pytest

from deorbit_amd import formats, ics


def test_impact_37k_particle_counts_and_geometry():
    ic = ics.impact_37k()
    assert ic["meta"]["n_target"] == 64 * 32 * 16
    assert ic["meta"]["n_impactor"] == 4169            # lattice points with i^2+j^2+k^2 <= 100
    assert len(ic["mass"]) == 36937
    imp = ic["mat"] == 1
    v = ic["vel"][imp][0]
    assert v == pytest.approx([8000 * np.cos(np.pi / 4), 0.0, -8000 * np.sin(np.pi / 4)])
    assert np.all(ic["vel"][~imp] == 0)
    assert ic["pos"][imp][:, 2].min() > ic["pos"][~imp][:, 2].max()   # impactor above the slab
    # jitter is bounded by 0.01 dx
    base = ics.impact(10, 45.0, 8000.0, 64, 32, [(0, 16)], jitter=0.0, seed=1000)
    assert np.max(np.abs(ic["pos"] - base["pos"])) <= 0.01 * ic["dx"] + 1e-12
    assert np.all(ic["mass"][imp] == 7800.0 * 1000) and np.all(ic["mass"][~imp] == 2680.0 * 1000)


def test_generators_are_deterministic():
    a, b = ics.impact_37k(), ics.impact_37k()
    for k in ("pos", "vel", "mass", "mat"):
        assert np.array_equal(a[k], b[k])
    c = ics.impact(10, 45.0, 8000.0, 64, 32, [(0, 16)], seed=1001)
    assert not np.array_equal(a["pos"], c["pos"])


def test_layered_target_and_relief():
    ic = ics.impact(3, 10.0, 5000.0, 12, 8, [(2, 3), (0, 5)], seed=2)
    tgt = ic["mat"] != 1
    z = ic["pos"][tgt][:, 2] / ic["dx"]
    assert np.all(ic["mat"][tgt][z > -3] == 2) and np.all(ic["mat"][tgt][z < -3] == 0)
    rel = np.zeros((12, 8))
    rel[:6] = 2
    rel[6:] = -1
    ic2 = ics.impact(3, 10.0, 5000.0, 12, 8, [(0, 8)], relief=rel, porous_top=(2, 0.98), seed=2, jitter=0)
    t2 = ic2["mat"] != 1
    assert t2.sum() == 6 * 8 * 10 + 6 * 8 * 7
    light = np.isclose(ic2["mass"][t2], 0.98 * 2680 * 1000)
    assert light.sum() == 12 * 8 * 2


def test_ensemble_members_follow_the_stated_distributions():
    ms = ics.ensemble_members()
    assert len(ms) == 256
    for m in ms:
        assert 1.0 <= m["theta"] <= 45.0 and 4000 <= m["speed"] <= 12000 and 0.05 <= m["ratio"] <= 0.3
        assert m["nx"] == 2 * m["ny"] and m["nz"] == m["ny"]
    ntot = [m["nx"] * m["ny"] * m["nz"] * (1 + m["ratio"]) for m in ms]
    assert 7000 < min(ntot) and max(ntot) < 230000


def test_gaussian_relief_statistics():
    r = ics.gaussian_relief(384, 192, 4.0, 30.0, 4000)
    assert r.shape == (384, 192)
    assert abs(r.mean()) < 1e-12 and r.std() == pytest.approx(4.0, rel=1e-12)


def test_ic_roundtrip(tmp_path):
    ic = ics.tiny_impact()
    p = os.path.join(tmp_path, "ic.bin")
    formats.write_ic(p, ic)
    assert os.path.getsize(p) == 8 * (7 + 10 * len(ic["mass"]))
    back = formats.read_ic(p)
    for k in ("pos", "vel", "mass", "u", "mat", "fixed"):
        assert np.array_equal(back[k], ic[k])
    assert back["lo"] == tuple(ic["lo"]) and back["hi"] == tuple(ic["hi"])


def test_ic_short_file_is_rejected(tmp_path):
    ic = ics.tiny_impact()
    p = os.path.join(tmp_path, "ic.bin")
    formats.write_ic(p, ic)
    with open(p, "r+b") as f:
        f.truncate(8 * (7 + 10 * 5))
    with pytest.raises(ValueError):
        formats.read_ic(p)


def test_snapshot_layout_matches_the_daemon_reader(tmp_path):
    # boundary_daemon.py:182-183 reads int64 n then n*10 doubles with mat in column 9
    rng = np.random.Generator(np.random.PCG64(0))
    n = 50
    pos, vel = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    u, rho, h = rng.random(n), rng.random(n), rng.random(n)
    mat = rng.integers(0, 3, n)
    p = os.path.join(tmp_path, "snap.bin")
    formats.write_snapshot(p, pos, vel, u, rho, h, mat)
    raw = open(p, "rb").read()
    assert int(np.frombuffer(raw[:8], "<i8")[0]) == n
    rows = np.frombuffer(raw[8:], "<f8").reshape(n, 10)
    assert np.array_equal(rows[:, 9], mat) and np.array_equal(rows[:, 6], u)
    back = formats.read_snapshot(p)
    assert np.array_equal(back["pos"], pos) and np.array_equal(back["mat"], mat)


def test_checkpoint_roundtrip_plain_and_extended(tmp_path):
    rng = np.random.Generator(np.random.PCG64(1))
    n = 40
    st = dict(pos=rng.normal(size=(n, 3)), vel=rng.normal(size=(n, 3)), mass=rng.random(n), u=rng.random(n),
              mat=rng.integers(0, 2, n), fixed=rng.integers(0, 2, n), D=rng.random(n), eps_act=rng.random(n),
              S=rng.normal(size=(n, 6)), hh=rng.random(n), acc=rng.normal(size=(n, 3)), dudt=rng.normal(size=n),
              dSdt=rng.normal(size=(n, 6)))
    p = os.path.join(tmp_path, "chk.bin")
    formats.write_checkpoint(p, st, (0, 0, 0), (1, 2, 3), 0.25, 10.0)
    assert os.path.getsize(p) == 8 * (9 + 18 * n)
    raw = np.fromfile(p, "<f8")
    assert np.array_equal(raw[9:].reshape(n, 18)[:, 12:18], st["S"][:, formats.MEM_TO_FILE])
    back = formats.read_checkpoint(p)
    assert back["t"] == 0.25 and back["dx"] == 10.0 and not back["extended"]
    for k in ("pos", "vel", "mass", "u", "D", "eps_act", "S"):
        assert np.array_equal(back[k], st[k])
    formats.write_checkpoint(p, st, (0, 0, 0), (1, 2, 3), 0.5, 10.0, extended=True, steps=17, dt_next=3.5e-4)
    assert os.path.getsize(p) == 8 * (9 + 18 * n + 11 * n + 3)
    back = formats.read_checkpoint(p)
    assert back["extended"] and back["steps"] == 17 and back["dt_next"] == 3.5e-4
    for k in ("hh", "acc", "dudt", "dSdt"):
        assert np.array_equal(back[k], st[k])
    assert not os.path.exists(p + ".tmp")
    # files with the earlier two-value trailer still read (no dt_next)
    raw = np.fromfile(p, "<f8")
    np.concatenate([raw[:-3], [formats.EXT_MAGIC_V0, 17.0]]).astype("<f8").tofile(p)
    back = formats.read_checkpoint(p)
    assert back["extended"] and back["steps"] == 17 and back["dt_next"] == 0.0


def test_grazing_ic_reproduces_the_campaign_geometry(tmp_path):
    # results/bounce_off_phase_diagram.json records N = 5,900,349 for every campaign run
    ic = ics.grazing_ic(2.0, 12000.0)
    assert len(ic["mass"]) == 5900349 and ic["dx"] == 62.5
    assert ic["meta"]["n_impactor"] == 2109
    imp = ic["mat"] == 1
    assert np.all(ic["fixed"][imp] == 0) and ic["fixed"][~imp].sum() > 0
    assert np.allclose(ic["vel"][imp][0], [12000 * np.cos(np.radians(2)), 0, -12000 * np.sin(np.radians(2))])
    small = ics.grazing_ic(10.0, 8000.0, cppr=4, D=100.0, L=(1500.0, 400.0, 500.0))
    p = os.path.join(tmp_path, "g.bin")
    formats.write_ic(p, small)
    back = formats.read_ic(p)
    assert np.array_equal(back["fixed"], small["fixed"]) and np.array_equal(back["mat"], small["mat"])


def test_ensemble_batch_and_impact_region_sample():
    E = ics.ensemble(8, 500)
    assert len(E) == 8 and all(e["meta"]["steps"] == 500 for e in E)
    assert all(8000 * 0.8 < len(e["mass"]) < 200000 * 1.2 for e in E)
    ic = ics.graze_1m()
    r = ics.impact_region(ic, 200000)
    assert len(r["mass"]) == 200000
    imp = int((r["mat"] == 1).sum())
    assert 0.2 * 200000 < imp < 0.8 * 200000          # both impactor and target in the sample
    assert ics.impact_region(ic, 10 ** 7) is ic



def test_recorded_traffic_lines_are_matched_by_kernel_workload_and_digest(tmp_path):
    from deorbit_amd import roofline
    d = roofline.source_digest()
    (tmp_path / "a.md").write_text(
        "text\n"
        f"traffic: kernel=k_force<float, false> workload=graze_1m digest={d} bytes=3.8e+08 "
        "fetch_factor=1.9999 write_factor=1.0000\n"
        "traffic: kernel=k_force<float, false> workload=graze_1m digest=0000 bytes=1 fetch_factor=1 write_factor=1\n")
    rec = roofline.recorded_traffic(str(tmp_path), "k_force<float, false>", "graze_1m")
    assert rec == {"bytes": 3.8e8, "fetch_factor": 1.9999, "write_factor": 1.0}
    assert roofline.recorded_traffic(str(tmp_path), "k_force<float, true>", "graze_1m") is None
    assert roofline.recorded_traffic(str(tmp_path), "k_force<float, false>", "layered_4m") is None


def test_source_digest_ignores_comments_and_layout():
    from deorbit_amd import roofline
    a = "int f(int x) {  // doubles x\n    return 2 * x; /* done */\n}\n"
    b = "/* header */\nint f(int x) {\n  return 2 * x;\n}"
    c = "int f(int x) { return 3 * x; }"
    assert roofline._code_only(a) == roofline._code_only(b) != roofline._code_only(c)
    assert roofline._code_only('const char* s = "a // b";') == 'const char* s = "a // b";'
