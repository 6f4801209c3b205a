"""Seeded initial conditions for the SPH impact workloads.

Every generator returns an IC dict with float64 ``pos (n,3)``, ``vel (n,3)``,
``mass``, ``u``, int32 ``mat``, uint8 ``fixed`` plus ``dx``, the domain box
``lo``/``hi`` (the reference's ``set_domain`` bounds, sph.hpp:96-101, and the
header of the IC file, run_ic.cpp:136-159), the solver ``params``, the
``materials`` list and a ``meta`` dict.  The target is written first and the
impactor last, as the reference IC builders do (boundary_daemon.py:119-158), so
that Weibull draws (which follow particle index order, sph.hpp:117-128) are
reproducible from the file alone.

Geometry conventions: lattice spacing ``dx``; the target's top surface is the
plane z = 0 with particles at cell centres below it; the impactor is the set of
integer lattice points i^2+j^2+k^2 <= R^2 (R in lattice units), its lowest point
one spacing above the surface, moving along +x and down at ``theta`` degrees.
Contact happens about one third of the way along the target in x.  Each
coordinate is jittered by U(-jitter*dx, +jitter*dx) from a PCG64 stream.
"""
from __future__ import annotations

import math

import numpy as np

from . import materials as M


def sphere_offsets(R: int) -> np.ndarray:
    """Integer lattice points with i^2 + j^2 + k^2 <= R^2, in (i, j, k) lexicographic order."""
    r = np.arange(-R, R + 1)
    i, j, k = np.meshgrid(r, r, r, indexing="ij")
    sel = (i * i + j * j + k * k) <= R * R
    return np.stack([i[sel], j[sel], k[sel]], axis=1).astype(np.float64)


def slab_centres(nx: int, ny: int, nz: int) -> np.ndarray:
    """Cell centres of an nx x ny x nz block below z = 0, (i, j, k) lexicographic."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([i.ravel() + 0.5, j.ravel() + 0.5, -(k.ravel() + 0.5)], axis=1).astype(np.float64)


def _finish(pos, vel, mass, mat, dx, params, mats, meta, margin=4.0, fixed=None):
    n = len(mass)
    lo = pos.min(axis=0) - margin * dx
    hi = pos.max(axis=0) + margin * dx
    prm = dict(params)
    prm["lo"] = tuple(float(v) for v in lo)
    prm["hi"] = tuple(float(v) for v in hi)
    return dict(pos=np.ascontiguousarray(pos), vel=np.ascontiguousarray(vel),
                mass=np.ascontiguousarray(mass), u=np.zeros(n), mat=mat.astype(np.int32),
                fixed=np.zeros(n, np.uint8) if fixed is None else fixed.astype(np.uint8),
                dx=float(dx), lo=prm["lo"], hi=prm["hi"], params=prm, materials=mats, meta=meta)


def impact(R: int, theta_deg: float, speed: float, nx: int, ny: int, layers, dx: float = 10.0,
           jitter: float = 0.01, seed: int = 1000, strength: bool = False, frac: bool = False,
           relief=None, porous_top: tuple | None = None, name: str = "impact", steps: int = 100):
    """Sphere on a layered target.

    ``layers`` is a list of ``(material index, thickness in cells)`` from the top
    down.  ``relief`` (optional, nx x ny, in cells) raises/lowers each column's top.
    ``porous_top = (thickness_cells, density_factor)`` scales the particle mass of
    the top layer (there is no porosity model; only the mass changes).
    """
    mats = M.impact_materials(strength)
    nz = int(sum(t for _, t in layers))
    # --- target, column by column in (i, j, k) order.  Column (i, j) fills the
    # cells between the flat bottom z = -nz and its local top z = rel[i, j].
    rel = (np.zeros((nx, ny), np.int64) if relief is None
           else np.rint(np.asarray(relief)).astype(np.int64))
    rmax = int(max(0, rel.max()))
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz + rmax), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    zc = rmax - (kk + 0.5)
    col = rel[ii, jj]
    keep = zc < col
    k_t = (col - rmax + kk)[keep]                              # 0 = top cell of the column
    pos_t = np.stack([ii[keep] + 0.5, jj[keep] + 0.5, zc[keep]], axis=1).astype(np.float64)
    mat_t = np.full(len(pos_t), layers[-1][0], np.int32)
    edge = 0
    for m, t in layers:
        mat_t[(k_t >= edge) & (k_t < edge + t)] = m
        edge += t
    rho0 = np.array([m["rho0"] for m in mats])
    mass_t = rho0[mat_t] * dx ** 3
    if porous_top is not None:
        thick, fac = porous_top
        mass_t[k_t < thick] *= fac
    # --- impactor
    off = sphere_offsets(R)
    th = math.radians(theta_deg)
    lx, ly = nx, ny
    x_contact = lx / 3.0
    gap = 1.0
    xc = x_contact - gap / math.tan(th)
    zc = gap + R + (0.0 if relief is None else float(np.max(relief)))
    yc = ly / 2.0
    pos_i = off + np.array([xc, yc, zc])
    n_t, n_i = len(pos_t), len(pos_i)
    pos = np.concatenate([pos_t, pos_i]) * dx
    vel = np.zeros_like(pos)
    vel[n_t:, 0] = speed * math.cos(th)
    vel[n_t:, 2] = -speed * math.sin(th)
    mass = np.concatenate([mass_t, np.full(n_i, mats[M.MAT_IMPACTOR]["rho0"] * dx ** 3)])
    mat = np.concatenate([mat_t, np.full(n_i, M.MAT_IMPACTOR, np.int32)])
    rng = np.random.Generator(np.random.PCG64(seed))
    pos += rng.uniform(-jitter * dx, jitter * dx, size=pos.shape)
    params = M.default_params(dx, (0, 0, 0), (1, 1, 1), frac)
    meta = dict(name=name, n_target=n_t, n_impactor=n_i, R=R, theta=theta_deg, speed=speed,
                nx=nx, ny=ny, nz=nz, seed=seed, steps=steps, strength=strength, frac=frac)
    return _finish(pos, vel, mass, mat, dx, params, mats, meta)


# ------------------------------------------------------------------ workloads
def impact_37k():
    """Plain FP64 CPU-runnable case: R=10dx iron sphere at 45 deg, 8 km/s, onto a
    64x32x16 granite slab; no strength; 100 steps."""
    return impact(10, 45.0, 8000.0, 64, 32, [(M.MAT_TARGET, 16)], seed=1000, name="impact_37k",
                  steps=100)


def graze_1m():
    """~1.05M particles: R=40dx at 3 deg, 8 km/s, into a 128x64x96 granite slab;
    strength (iron Y=1 GPa, G=80 GPa; granite Y=2 GPa, G=30 GPa) and Weibull damage
    (k=1e39 m^-3, m=9) on; 500 steps."""
    return impact(40, 3.0, 8000.0, 128, 64, [(M.MAT_TARGET, 96)], seed=1001, strength=True,
                  frac=True, name="graze_1m", steps=500)


def layered_4m():
    """~4.0M particles: R=60dx at 2 deg, 8 km/s, into 192x96 columns of 32 water
    cells over 136 granite cells; strength and damage on; 1000 steps."""
    return impact(60, 2.0, 8000.0, 192, 96, [(M.MAT_WATER, 32), (M.MAT_TARGET, 136)],
                  seed=1002, strength=True, frac=True, name="layered_4m", steps=1000)


def gaussian_relief(nx, ny, amplitude, corr, seed):
    """Gaussian random surface: white noise filtered by a Gaussian of correlation
    length ``corr`` (cells), scaled to standard deviation ``amplitude`` (cells)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((nx, ny))
    kx = np.fft.fftfreq(nx)[:, None]
    ky = np.fft.fftfreq(ny)[None, :]
    filt = np.exp(-0.5 * (2 * math.pi) ** 2 * (kx ** 2 + ky ** 2) * corr ** 2)
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    f -= f.mean()
    sd = f.std()
    return f * (amplitude / sd if sd > 0 else 0.0)


def halfspace_32m():
    """~32M particles: R=120dx at 1.5 deg, 8 km/s, onto a 384x192 granite half-space
    ~339 cells deep with a 16-cell top layer at 0.98 rho0 (mass-scaled) and a
    Gaussian surface relief (4dx amplitude, 30dx correlation, seed 4000); 200 steps."""
    rel = gaussian_relief(384, 192, 4.0, 30.0, 4000)
    return impact(120, 1.5, 8000.0, 384, 192, [(M.MAT_TARGET, 339)], seed=4001, strength=True,
                  frac=True, relief=rel, porous_top=(16, 0.98), name="halfspace_32m", steps=200)


def ensemble_members(count: int = 256, seed0: int = 3000, nmin: int = 8000, nmax: int = 200000):
    """Parameters of the batched ensemble: per instance N ~ logU[nmin, nmax], angle
    U[1,45] deg, speed U[4,12] km/s, impactor/target particle ratio U[0.05,0.3].
    The target is a 2:1:1 slab; the impactor radius follows from its particle count."""
    out = []
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64(seed0 + i))
        n = float(np.exp(rng.uniform(math.log(nmin), math.log(nmax))))
        ang = float(rng.uniform(1.0, 45.0))
        spd = float(rng.uniform(4000.0, 12000.0))
        ratio = float(rng.uniform(0.05, 0.3))
        n_t = n / (1.0 + ratio)
        n_i = n - n_t
        ny = max(4, int(round((n_t / 2.0) ** (1.0 / 3.0))))
        R = max(2, int(round((3.0 * n_i / (4.0 * math.pi)) ** (1.0 / 3.0))))
        out.append(dict(index=i, seed=seed0 + i, R=R, theta=ang, speed=spd, nx=2 * ny, ny=ny,
                        nz=ny, ratio=ratio))
    return out


def ensemble_instance(member: dict, steps: int = 500):
    return impact(member["R"], member["theta"], member["speed"], member["nx"], member["ny"],
                  [(M.MAT_TARGET, member["nz"])], seed=member["seed"], strength=True, frac=True,
                  name=f"ensemble{member['index']:03d}", steps=steps)


def sod(nx_left: int = 100, ny_left: int = 16, gamma: float = 1.4):
    """The Sod shock tube of the reference's validation gate (sod.cpp): ideal gas,
    x in [-0.5, 0.5] (left state rho=1, P=1 on a lattice of spacing a = 0.5/nx_left;
    right state rho=0.125, P=0.1 on spacing 2a, equal particle masses), periodic in
    y and z over ny_left * a, particles within 2 h_init of either end frozen.
    Artificial viscosity alpha=1, beta=2 and Price conductivity alpha_u=1.
    Defaults give the reference's 28,800 particles."""
    aL = 0.5 / nx_left
    aR = 2.0 * aL
    nx_right, ny_right = nx_left // 2, ny_left // 2
    Lyz = ny_left * aL
    eta = 1.3
    h0 = eta * aL
    rhoL, rhoR, PL, PR = 1.0, 0.125, 1.0, 0.1
    m = rhoL * aL ** 3
    uL = PL / ((gamma - 1.0) * rhoL)
    uR = PR / ((gamma - 1.0) * rhoR)
    xfix = 2.0 * h0

    def block(x0, a, nx, ny, u):
        i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(ny), indexing="ij")
        pos = np.stack([x0 + (i.ravel() + 0.5) * a, (j.ravel() + 0.5) * a, (k.ravel() + 0.5) * a], axis=1)
        return pos, np.full(len(pos), u)

    pl, ul = block(-0.5, aL, nx_left, ny_left, uL)
    pr, ur = block(0.0, aR, nx_right, ny_right, uR)
    pos = np.concatenate([pl, pr])
    n = len(pos)
    fixed = ((pos[:, 0] < -0.5 + xfix) | (pos[:, 0] > 0.5 - xfix)).astype(np.uint8)
    params = dict(eta=eta, h_init=h0, alpha=1.0, beta=2.0, eps=0.01, alpha_u=1.0, cfl=0.2, eps_as=0.0,
                  damage_on=False, periodic=(False, True, True), lo=(-0.5, 0.0, 0.0), hi=(0.5, Lyz, Lyz),
                  g=0.0, damp=1.0)
    return dict(pos=pos, vel=np.zeros((n, 3)), mass=np.full(n, m), u=np.concatenate([ul, ur]),
                mat=np.zeros(n, np.int32), fixed=fixed, dx=aL, lo=params["lo"], hi=params["hi"], params=params,
                materials=[M.ideal(gamma)], meta=dict(name=f"sod{nx_left}x{ny_left}", steps=0, t_end=0.2,
                                                      states=dict(rhoL=rhoL, PL=PL, rhoR=rhoR, PR=PR, gamma=gamma)))


def elastic_wave(dx: float = 0.01, x0: float = -0.6, x1: float = 0.6, Lyz: float = 0.08, v0: float = 1.0,
                 width: float = 0.06):
    """The elastic longitudinal wave of the reference's strength gate
    (strength_test.cpp:11-24,64-85): a basalt bar on a cubic lattice, periodic in
    y and z, a Gaussian velocity pulse v_x = v0 exp(-x^2/w^2) at rest otherwise,
    the ends |x| > 0.55 frozen.  The pulse splits into two waves travelling at
    c_L = sqrt((A + 4G/3) / rho0)."""
    mat = M.basalt_preset()
    nx, ny = int(round((x1 - x0) / dx)), int(round(Lyz / dx))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(ny), indexing="ij")
    pos = np.stack([x0 + (i.ravel() + 0.5) * dx, (j.ravel() + 0.5) * dx, (k.ravel() + 0.5) * dx], axis=1)
    n = len(pos)
    vel = np.zeros((n, 3))
    vel[:, 0] = v0 * np.exp(-(pos[:, 0] ** 2) / (width * width))
    fixed = (np.abs(pos[:, 0]) > 0.55).astype(np.uint8)
    params = M.default_params(dx, (x0, 0.0, 0.0), (x1, Lyz, Lyz), False, periodic=(False, True, True))
    c_L = math.sqrt((mat["A"] + 4.0 / 3.0 * mat["G"]) / mat["rho0"])
    return dict(pos=pos, vel=vel, mass=np.full(n, mat["rho0"] * dx ** 3), u=np.zeros(n), mat=np.zeros(n, np.int32),
                fixed=fixed, dx=dx, lo=params["lo"], hi=params["hi"], params=params, materials=[mat],
                meta=dict(name="elastic_wave", steps=0, t_end=6.0e-5, c_L=c_L))


def impact_region(ic: dict, nmax: int) -> dict:
    """The `nmax` particles nearest (in max-norm) to the impactor's lowest point,
    where it meets the target: a bounded sample of a large impact that keeps its
    contact zone, part impactor and part target (used to time the CPU solver on a
    share of a workload too large for it)."""
    n = len(ic["mass"])
    if n <= nmax:
        return ic
    imp = np.flatnonzero(ic["mat"] == M.MAT_IMPACTOR)
    c = ic["pos"][imp[np.argmin(ic["pos"][imp, 2])]] if len(imp) else ic["pos"].mean(axis=0)
    d = np.abs(ic["pos"] - c).max(axis=1)
    keep = np.sort(np.argpartition(d, nmax)[:nmax])
    out = dict(ic)
    for k in ("pos", "vel", "mass", "u", "mat", "fixed"):
        out[k] = np.ascontiguousarray(np.asarray(ic[k])[keep])
    out["meta"] = dict(ic["meta"], name=ic["meta"]["name"] + f"-region{nmax}")
    return out


def ensemble(count: int = 256, steps: int = 500, seed0: int = 3000):
    """The batched ensemble: `count` independent iron-on-granite impacts (strength and
    damage on) drawn by ensemble_members; one IC dict per instance."""
    return [ensemble_instance(m, steps) for m in ensemble_members(count, seed0)]


def _arange(start, stop, step):
    # float stepping exactly as boundary_daemon.py:113-117 does it (accumulated sums)
    out, x = [], start
    while (x < stop) if step > 0 else (x > stop):
        out.append(x)
        x += step
    return np.array(out)


def grazing_ic(theta_deg: float, v: float, cppr: int = 8, D: float = 1000.0, L=None, tgt_mat: int = 0,
               imp_mat: int = 1, rho_t: float = 2700.0, rho_imp: float = 7800.0):
    """The bounce-off campaign's IC (boundary_daemon.py:119-158): a half-space target
    (x from -0.4 Lx to 0.6 Lx, y centred, z below 0) with a 6 dx frozen wall on the
    sides and bottom, then an iron sphere of radius D/2 released 3 dx above the
    surface at -0.35 Lx, moving at angle theta.  Rows are written target first
    (x, then y, then z) and sphere last (i, j, k), as the daemon writes them.
    ``L = (Lx, Ly, Lz)`` overrides the default max(20 km, 30D) x max(4 km, 6D) x
    max(6 km, 8D) box (used to make small test cases)."""
    a = D / 2.0
    dx = a / cppr
    th = math.radians(theta_deg)
    vx, vz = v * math.cos(th), -v * math.sin(th)
    Lx, Ly, Lz = L if L is not None else (max(20e3, 30 * D), max(4e3, 6 * D), max(6e3, 8 * D))
    wall = 6.0 * dx
    xs = _arange(-0.4 * Lx + dx / 2, 0.6 * Lx, dx)
    ys = _arange(-Ly / 2 + dx / 2, Ly / 2, dx)
    zs = _arange(-dx / 2, -Lz, -dx)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    fx = (np.abs(X - xs[0]) < wall) | (np.abs(X - xs[-1]) < wall)
    fxy = fx | (np.abs(Y - ys[0]) < wall) | (np.abs(Y - ys[-1]) < wall)
    fixed_t = (fxy | (Z < zs[-1] + wall)).astype(np.uint8)
    z0 = a + 3 * dx
    x0 = -0.35 * Lx
    na = int(math.ceil(a / dx)) + 1
    r = np.arange(-na, na + 1)
    I, J, K = np.meshgrid(r, r, r, indexing="ij")
    sel = (I * dx) ** 2 + (J * dx) ** 2 + (K * dx) ** 2 <= a * a
    I, J, K = I[sel], J[sel], K[sel]
    pos = np.concatenate([np.stack([X, Y, Z], 1), np.stack([I * dx + x0, J * dx, K * dx + z0], 1)])
    n_t, n_i = len(X), len(I)
    vel = np.zeros_like(pos)
    vel[n_t:, 0] = vx
    vel[n_t:, 2] = vz
    mass = np.concatenate([np.full(n_t, rho_t * dx ** 3), np.full(n_i, rho_imp * dx ** 3)])
    mat = np.concatenate([np.full(n_t, tgt_mat, np.int32), np.full(n_i, imp_mat, np.int32)])
    fixed = np.concatenate([fixed_t, np.zeros(n_i, np.uint8)])
    lo = (xs[0] - dx, ys[0] - dx, zs[-1] - dx)
    hi = (xs[-1] + dx, ys[-1] + dx, z0 + a + 0.5 * Lz)
    return dict(pos=pos, vel=vel, mass=mass, u=np.zeros(len(mass)), mat=mat, fixed=fixed, dx=dx, lo=lo, hi=hi,
                meta=dict(name="grazing", theta=theta_deg, v=v, n_target=n_t, n_impactor=n_i))


def tiny_impact(R=4, nx=16, ny=10, nz=6, theta=30.0, speed=6000.0, strength=True, frac=True,
                seed=7, jitter=0.01):
    """A few-thousand-particle impact with the full physics, for fast parity tests."""
    return impact(R, theta, speed, nx, ny, [(M.MAT_TARGET, nz)], seed=seed, strength=strength,
                  frac=frac, jitter=jitter, name="tiny", steps=20)


def sheared_cube(L: int = 22, dx: float = 30.0):
    """The reference's force-oracle state builder (cpu_dump.cpp:11-31): an L^3 basalt
    cube with a central iron sphere of radius 5dx, shear/convergence velocities and
    an x-gradient in u; fracture physics on.  Materials: 0 basalt, 1 iron."""
    lo = -L * dx / 2.0
    hi = L * dx / 2.0
    i, j, k = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    x = lo + (i.ravel() + 0.5) * dx
    y = lo + (j.ravel() + 0.5) * dx
    z = lo + (k.ravel() + 0.5) * dx
    R = 5.0 * dx
    iron = (x * x + y * y + z * z) < R * R
    pos = np.stack([x, y, z], axis=1)
    vel = np.stack([80.0 * (y / hi), np.zeros_like(x), -40.0 * (z / hi)], axis=1)
    u = 2.0e5 * (1.0 + x / hi)
    rho0 = np.where(iron, 7800.0, 2700.0)
    mass = rho0 * dx * dx * dx
    mat = iron.astype(np.int32)
    mats = [M.basalt_preset(), M.iron_preset()]
    params = M.default_params(dx, (lo - dx,) * 3, (hi + dx,) * 3, True)
    return dict(pos=pos, vel=vel, mass=mass, u=u, mat=mat, fixed=np.zeros(len(x), np.uint8),
                dx=dx, lo=params["lo"], hi=params["hi"], params=params, materials=mats,
                meta=dict(name="sheared_cube", steps=12, seed=None))
