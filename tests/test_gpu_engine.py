"""Engine properties of the device step: determinism, graph replay, batching,
size-independent invariants at the benchmark sizes, and error handling."""
import numpy as np
import pytest

from conftest import scale_err
from deorbit_amd import dsph, formats, ics
from deorbit_amd import materials as M
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "u", "rho", "hh", "S", "D", "acc", "dudt")


def run_states(ic, precision, steps, graph, tuning=None):
    s = dsph.make_system(ic, precision)
    if tuning:
        s.set_tuning(**tuning)
    s.init()
    s.run(steps, 0.0, graph)
    st = s.state(list(FIELDS))
    d = s.diagnostics()
    s.close()
    return st, d


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_runs_are_bitwise_reproducible(precision, gpu):
    ic = ics.tiny_impact()
    a, _ = run_states(ic, precision, 6, False)
    b, _ = run_states(ic, precision, 6, False)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_graph_replay_equals_stream_launches(precision, gpu):
    # graphs are replayed several times, with eager work and synchronisation in between
    ic = ics.tiny_impact()
    a, da = run_states(ic, precision, 7, False)
    s = dsph.make_system(ic, precision)
    s.init()
    for _ in range(3):
        s.run(2, 0.0, True)            # one replay of the two-step graph
        s.diagnostics()                # eager kernels + host synchronisation between replays
    s.run(1, 0.0, True)                # odd remainder: one stream-launched step
    b = s.state(list(FIELDS))
    db = s.diagnostics()
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    assert da["t"] == db["t"] and da["steps"] == db["steps"] == 7


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED_FAST])
def test_neighbour_lists_give_the_cell_walk_results(precision, gpu):
    # The lists hold the density walk's pairs in its order, so every way of
    # falling back (lists off, lists overflowing, h outgrowing a tiny skin)
    # must reproduce the same numbers up to the compiler contracting
    # multiply-adds differently in the list and walk kernels and the force walk
    # taking the columns in mirrored order for half of the lanes: round-off of
    # the working precision (scale-normalised 1e-13 in FP64, 1e-5 in FP32 after
    # 12 steps).
    ic = ics.impact_37k()
    steps = 12
    ref, dref = run_states(ic, precision, steps, True, dict(neighbor_capacity=-1))
    assert dref["list_walk_steps"] == 0 and dref["list_overflow"] == 0
    n = len(ic["mass"])
    runs = {}
    on, don = run_states(ic, precision, steps, True, dict(neighbor_capacity=128))
    assert don["list_overflow"] < 1e-3 * n * (steps + 1), don     # a few interface particles at most
    runs["lists"] = on
    small, dsmall = run_states(ic, precision, steps, True, dict(neighbor_capacity=40))
    tight, dtight = run_states(ic, precision, steps, True, dict(neighbor_capacity=128, list_skin=1e-4))
    assert dsmall["list_overflow"] > 0, dsmall
    assert dtight["list_walk_steps"] > 0, dtight
    runs["overflowing lists"] = small
    runs["tiny skin"] = tight
    for name, st in runs.items():
        for f in FIELDS:
            bar = 1e-13 if precision == dsph.FP64 else 1e-5
            assert scale_err(st[f], ref[f]) < bar, (name, f, scale_err(st[f], ref[f]))


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_checkpoint_restart_is_bit_exact(precision, gpu, tmp_path):
    """Six steps straight through equal three steps, an extended checkpoint written
    to disk, a fresh context restored from it, and three more steps: bit for bit
    (the per-step sort orders particles of a cell by IC index, so the restored
    context sums in the same order)."""
    ic = ics.tiny_impact()
    a = dsph.make_system(ic, precision)
    a.init()
    a.run(6, 0.0, True)
    ref, dref = a.state(list(FIELDS)), a.diagnostics()
    a.close()
    b = dsph.make_system(ic, precision)
    b.init()
    b.run(3, 0.0, True)
    st = b.checkpoint_state()
    b.close()
    path = str(tmp_path / "ckpt.bin")
    formats.write_checkpoint(path, st, ic["lo"], ic["hi"], st["t"], ic["dx"], extended=True, steps=st["steps"],
                             dt_next=st["dt_next"])
    ck = formats.read_checkpoint(path)
    c = dsph.System(ic["params"], ic["materials"], precision)
    c.add(ck["pos"], ck["vel"], ck["mass"], ck["u"], ck["mat"], ck["fixed"])
    c.restore_state(ck)
    c.run(3, 0.0, True)
    got, dgot = c.state(list(FIELDS)), c.diagnostics()
    c.close()
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), (f, scale_err(got[f], ref[f]))
    assert dgot["t"] == dref["t"] and dgot["steps"] == dref["steps"] == 6


@pytest.mark.parametrize("cell_factor", [1.0, 1.25, 3.0])
def test_tuning_changes_speed_not_results_beyond_roundoff(cell_factor, gpu):
    ic = ics.tiny_impact()
    a, _ = run_states(ic, dsph.FP64, 4, True)
    b, _ = run_states(ic, dsph.FP64, 4, True, dict(cell_factor=cell_factor, block_pair=64))
    for f in FIELDS:
        assert scale_err(b[f], a[f]) < 1e-12, f


def test_tuning_limits_are_enforced(gpu):
    """Columns narrower than h_init or a list skin above 0.25 could let a z range
    reach past two slice blocks, and neighbour lists serve only the FP32 density
    sums; the library refuses them."""
    s = dsph.make_system(ics.tiny_impact(), dsph.MIXED)
    for bad in (dict(cell_factor=0.5), dict(list_skin=0.5), dict(neighbor_capacity=128)):
        with pytest.raises(dsph.DsphError) as e:
            s.set_tuning(**bad)
        assert e.value.code == -1
    s.set_tuning(cell_factor=1.0, list_skin=0.25)
    s.init()
    s.close()


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_batch_of_segments_equals_separate_runs(precision, gpu):
    """Three impacts of uneven size in one batched context step exactly like three
    separate contexts (each segment keeps its own grid, dt and density rounds)."""
    members = [ics.tiny_impact(seed=1), ics.tiny_impact(R=3, nx=12, ny=8, nz=5, theta=60, seed=2),
               ics.tiny_impact(R=5, nx=18, ny=12, nz=7, theta=15, speed=9000, seed=3)]
    sep = []
    for ic in members:
        s = dsph.make_system(ic, precision)
        s.init()
        s.run(5, 0.0, False)
        sep.append((s.state(list(FIELDS)), s.diagnostics()))
        s.close()
    b = dsph.System([ic["params"] for ic in members], members[0]["materials"], precision)
    for k, ic in enumerate(members):
        b.add_ic(ic, seg=np.full(len(ic["mass"]), k, np.int32))
    for k, ic in enumerate(members):
        b.seed_damage(ic["dx"], 12345, seg=k)
    b.init()
    b.run(5, 0.0, True)
    st = b.state(list(FIELDS))
    off = 0
    for k, ic in enumerate(members):
        n = len(ic["mass"])
        for f in FIELDS:
            assert np.array_equal(st[f][off:off + n], sep[k][0][f]), (k, f)
        d = b.diagnostics(k)
        assert d["t"] == sep[k][1]["t"] and d["e_tot"] == sep[k][1]["e_tot"]
        off += n
    dts = b.timestep()
    assert len(dts) == 3 and len(set(dts.tolist())) == 3       # each segment its own CFL step


def test_full_size_invariants_graze_1m(gpu):
    """~1.05M particles with strength and damage: after init and after 10 steps the
    pairwise momentum exchange cancels (sum m a = 0) and h stays in its clamp."""
    ic = ics.graze_1m()
    for precision in (dsph.MIXED, dsph.FP64):
        s = dsph.make_system(ic, precision)
        s.init()
        for k in range(2):
            m = s.get("mass")
            ma = m[:, None] * s.get("acc")
            tol = 1e-10 if precision == dsph.FP64 else 2e-5
            assert np.all(np.abs(ma.sum(0)) <= tol * np.abs(ma).sum(0)), (precision, k)
            h = s.get("hh")
            h0 = ic["params"]["h_init"]
            assert h.min() >= 0.5 * h0 * (1 - 1e-6) and h.max() <= 2 * h0 * (1 + 1e-6)
            d = s.diagnostics()
            assert d["nonfinite"] == 0 and d["n"] == len(m)
            s.run(10, 0.0, True)
        s.close()


def test_full_size_fp64_init_matches_oracle_graze_1m(gpu):
    ic = ics.graze_1m()
    o = po.make_system(ic, ic["params"], ic["materials"], seed_damage=12345, dx=ic["dx"])
    o.init()
    s = dsph.make_system(ic, dsph.FP64)
    s.init()
    for name, f in (("rho", po.RHO), ("hh", po.HH), ("P", po.P), ("acc", po.ACC), ("dudt", po.DUDT),
                    ("dSdt", po.DSDT), ("sigmax", po.SIGMAX)):
        e = scale_err(s.get(name), o.get(f))
        print(f"graze_1m init fp64 {name}={e:.2e}", flush=True)
        assert e < 1e-12, name
    assert abs(s.timestep() - o.timestep()) <= 1e-14 * o.timestep()


def test_init_above_two_million_particles_matches_oracle(gpu):
    """2.08M particles (R=30dx at 30 deg into a 160x96x128 slab): the sort uses its
    16-key tiles from 2M keys on, so the cell list of this size is checked here
    against the oracle's init, in FP64 to 1e-12 and in mixed precision to the FP32
    density bar."""
    ic = ics.impact(30, 30.0, 8000.0, 160, 96, [(M.MAT_TARGET, 128)], seed=2001, strength=True, frac=True,
                    name="impact_2m", steps=10)
    assert len(ic["mass"]) > 2_000_000
    o = po.make_system(ic, ic["params"], ic["materials"], seed_damage=12345, dx=ic["dx"])
    o.init()
    for precision, bar in ((dsph.FP64, 1e-12), (dsph.MIXED, 1e-5)):
        s = dsph.make_system(ic, precision)
        s.init()
        fields = (("rho", po.RHO), ("hh", po.HH), ("P", po.P), ("acc", po.ACC))
        for name, f in fields[:2] if precision == dsph.MIXED else fields:
            e = scale_err(s.get(name), o.get(f))
            print(f"impact_2m init {'fp64' if precision == dsph.FP64 else 'mixed'} {name}={e:.2e}", flush=True)
            assert e < bar, (precision, name)
        s.close()


def test_energy_rate_balance_impact_37k_mixed(gpu):
    # no strength, no artificial stress: sum m (v.a + du/dt) vanishes pairwise
    s = dsph.make_system(ics.impact_37k(), dsph.MIXED)
    s.init()
    m, v, a, du = s.get("mass"), s.get("vel"), s.get("acc"), s.get("dudt")
    kin = m * np.sum(v * a, axis=1)
    th = m * du
    assert abs(kin.sum() + th.sum()) <= 1e-5 * (np.abs(kin).sum() + np.abs(th).sum())


def test_argument_errors_are_reported(gpu):
    ic = ics.tiny_impact()
    p = dict(ic["params"])
    with pytest.raises(dsph.DsphError) as e:
        dsph.System(dict(p, eta=-1.0), ic["materials"])
    assert e.value.code == -1
    with pytest.raises(dsph.DsphError):       # a periodic axis needs a positive extent
        dsph.System(dict(p, periodic=(True, False, False), hi=(p["lo"][0], *p["hi"][1:])), ic["materials"])
    s = dsph.System(p, ic["materials"])
    bad = np.array(ic["mat"])
    bad[3] = 7
    with pytest.raises(dsph.DsphError):
        s.add(ic["pos"], ic["vel"], ic["mass"], ic["u"], bad)
    s.add_ic(ic)
    with pytest.raises(dsph.DsphError) as e:
        s.step(0.0)                    # no device dt before init
    assert e.value.code == -2
    with pytest.raises(ValueError):
        s.set("rho", np.zeros(3))
    s.init()
    s.step(0.0)
    s.synchronize()
    # a context holds at most 134,217,664 particles (records are addressed by 32-bit
    # byte offsets): a larger upload is refused before any array element is read
    import ctypes as C
    one3, one1, onei = np.zeros(3), np.zeros(1), np.zeros(1, np.int32)
    rc = s.lib.dsph_upload(s.h, 134_217_665 - len(ic["mass"]), dsph._ptr(one3, C.c_double),
                           dsph._ptr(one3, C.c_double), dsph._ptr(one1, C.c_double), dsph._ptr(one1, C.c_double),
                           dsph._ptr(onei, C.c_int32), None, None)
    assert rc == -5, rc
    # a refused upload leaves the particles on the device as they were
    assert s.n == len(ic["mass"])
    with pytest.raises(dsph.DsphError):
        s.add(ic["pos"][:5], ic["vel"][:5], ic["mass"][:5], ic["u"][:5], bad[:5])
    assert s.n == len(ic["mass"])
    x0 = s.get("pos")
    s.step(0.0)
    s.synchronize()
    assert np.all(np.isfinite(s.get("pos"))) and not np.array_equal(s.get("pos"), x0)


def test_nonfinite_particle_is_parked_and_reported(gpu):
    ic = ics.tiny_impact()
    s = dsph.make_system(ic, dsph.MIXED)
    s.init()
    pos = s.get("pos")
    pos[5] = np.nan
    s.set("pos", pos)
    # the step finishes (the particle sits in its segment's park cell) and
    # reports the device's flag
    with pytest.raises(dsph.DsphError) as e:
        s.step(1e-6)
    assert e.value.code == -6 and "non-finite particle position" in str(e.value)
    with pytest.raises(dsph.DsphError) as e:          # the flag is sticky
        s.synchronize()
    assert e.value.code == -6
    d = s.diagnostics()                               # still readable, with the status
    assert d["nonfinite"] >= 1 and d["n"] == len(ic["mass"]) and d["status"] == -6


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_reach_beyond_the_cell_layout_is_reported_by_step_and_run(precision, gpu):
    """A smoothing length set from outside far above 2 h_init reaches past two
    slice blocks of a column; the walk skips those pairs and flags it, and the
    very step (or run) that met it returns DSPH_E_RANGE.  The kernels only set a
    flag and finish normally."""
    dx = 10.0
    i, j, k = np.meshgrid(np.arange(3), np.arange(3), np.arange(150), indexing="ij")
    pos = np.stack([i.ravel(), j.ravel(), -k.ravel()], 1).astype(float) * dx      # a 1.5 km tall column
    n = len(pos)
    mats = M.impact_materials(False)
    prm = M.default_params(dx, pos.min(0) - 4 * dx, pos.max(0) + 4 * dx, False)
    for use_run in (False, True):
        s = dsph.System(prm, mats, precision)
        s.add(pos, np.zeros((n, 3)), np.full(n, mats[0]["rho0"] * dx ** 3), np.zeros(n), np.zeros(n, np.int32))
        s.init()
        assert s.diagnostics()["status"] == 0
        h = s.get("hh")
        h[n // 2] = 30.0 * prm["h_init"]               # reach 2h = 780 m: some 320 slices of 4.9 m
        s.set("hh", h)
        with pytest.raises(dsph.DsphError) as e:
            if use_run:
                s.run(2, 1e-4, True)
            else:
                s.step(1e-4)
        assert e.value.code == -5 and "two slice blocks" in str(e.value), str(e.value)
        assert s.diagnostics()["status"] == -5
        s.close()


def test_autotune_keeps_state(gpu):
    ic = ics.tiny_impact()
    s = dsph.make_system(ic, dsph.FP64)
    s.init()
    before = s.state(list(FIELDS))
    t = s.autotune(2)
    assert t["block_pair"] in (64, 128, 256) and 1.0 <= t["cell_factor"] <= 2.0
    after = s.state(list(FIELDS))
    for f in FIELDS:
        assert np.array_equal(before[f], after[f]), f
    s.run(2, 0.0, True)
    s.synchronize()


def test_profile_reports_every_phase(gpu):
    s = dsph.make_system(ics.tiny_impact(), dsph.MIXED)
    s.init()
    s.set_profiling(True)
    s.reset_profile()
    s.run(3, 0.0, False)
    p = s.profile()
    for ph in ("kick_drift", "sort", "permute", "density", "force", "finish"):
        assert p[ph][1] >= 3 and p[ph][0] > 0, ph


@pytest.mark.parametrize("precision", [dsph.FP64, dsph.MIXED])
def test_batch_with_an_empty_segment(precision, gpu):
    """A segment that receives no particles (an instance dropped from a sweep)
    leaves the others exactly as separate runs and reports n = 0."""
    members = [ics.tiny_impact(seed=1), ics.tiny_impact(R=3, nx=12, ny=8, nz=5, theta=60, seed=2)]
    sep = []
    for ic in members:
        s = dsph.make_system(ic, precision)
        s.init()
        s.run(4, 0.0, False)
        sep.append(s.state(list(FIELDS)))
        s.close()
    params = [members[0]["params"], members[0]["params"], members[1]["params"]]
    b = dsph.System(params, members[0]["materials"], precision)
    b.add_ic(members[0], seg=np.zeros(len(members[0]["mass"]), np.int32))
    b.add_ic(members[1], seg=np.full(len(members[1]["mass"]), 2, np.int32))
    b.seed_damage(members[0]["dx"], 12345, seg=0)
    b.seed_damage(members[1]["dx"], 12345, seg=2)
    b.init()
    b.run(4, 0.0, True)
    st = b.state(list(FIELDS))
    n0 = len(members[0]["mass"])
    for f in FIELDS:
        assert np.array_equal(st[f][:n0], sep[0][f]), f
        assert np.array_equal(st[f][n0:], sep[1][f]), f
    assert b.diagnostics(1)["n"] == 0 and b.diagnostics(2)["n"] == len(members[1]["mass"])
    dts = b.timestep()
    assert np.isfinite(dts[0]) and np.isfinite(dts[2])
    b.close()


@pytest.mark.parametrize("precision,tol", [(dsph.FP64, 1e-14), (dsph.MIXED, 1e-6)])
def test_single_particle_density_has_the_closed_form(precision, tol, gpu):
    """One particle sees only itself: rho = m W(0, h) = m / (pi h^3).  With
    eta pi^(1/3) ~ 1.90 > 1 every iteration raises h until the clamp at 2 h_init,
    where the rounds stop, so rho = m / (pi (2 h_init)^3) and h = 2 h_init
    (sph.hpp:189-201)."""
    ic = ics.tiny_impact()
    p = dict(ic["params"])
    h0 = p["h_init"]
    m = 7.8e6
    s = dsph.System(p, ic["materials"], precision)
    s.add(np.array([[1.0, 2.0, 3.0]]), np.zeros((1, 3)), np.array([m]), np.array([0.0]),
          np.array([1], np.int32))
    s.compute_density_h()
    assert abs(s.get("hh")[0] - 2 * h0) <= 1e-15 * h0
    rho = m / (np.pi * (2 * h0) ** 3)
    assert abs(s.get("rho")[0] - rho) <= tol * rho
    s.init()
    s.step(0.0)
    s.synchronize()
    assert np.all(s.get("acc") == 0.0) and s.diagnostics()["nonfinite"] == 0
    s.close()


@pytest.mark.parametrize("precision,tol", [(dsph.FP64, 1e-13), (dsph.MIXED, 1e-6)])
def test_density_moment_walks_and_rewalks_match_the_oracle(precision, tol, gpu, monkeypatch):
    """The second density iteration of a round comes from the first walk's moment
    sums while h stays in their band and from a second walk elsewhere; both
    branches run here (init moves h far from h_init, later steps little) and the
    densities and smoothing lengths follow the oracle's direct sums (sph.hpp:189-198)
    per particle.  The counters (DSPH_COUNT_REWALKS=1) show which branch ran."""
    monkeypatch.setenv("DSPH_COUNT_REWALKS", "1")
    ic = ics.impact_37k()
    o = po.make_system(ic, ic["params"], ic["materials"], dx=ic["dx"])
    o.init()
    s = dsph.make_system(ic, precision)
    s.init()
    d0 = s.diagnostics()
    n = len(ic["mass"])
    assert d0["density_lanes"] >= n and d0["density_rewalks"] > 0, d0
    for k in range(8):
        dt = o.timestep()
        o.step(dt)
        s.step(dt)
        e_rho = float(np.max(np.abs(s.get("rho") - o.get(po.RHO)) / o.get(po.RHO)))
        e_h = float(np.max(np.abs(s.get("hh") - o.get(po.HH)) / o.get(po.HH)))
        assert e_rho < tol and e_h < tol, (k, e_rho, e_h)
    d1 = s.diagnostics()
    lanes = d1["density_lanes"] - d0["density_lanes"]
    rewalks = d1["density_rewalks"] - d0["density_rewalks"]
    print(f"moment walks: init {d0['density_rewalks']}/{d0['density_lanes']} lanes walked again, "
          f"8 steps {rewalks}/{lanes}")
    assert lanes >= 8 * n and rewalks < 0.5 * lanes, (lanes, rewalks)
    s.close()
    monkeypatch.delenv("DSPH_COUNT_REWALKS")
    s2 = dsph.make_system(ic, precision)           # counters stay off without the variable
    s2.init()
    assert s2.diagnostics()["density_lanes"] == 0
    s2.close()
