"""On-disk formats of the reference SPH drivers (little-endian float64 unless noted).

* IC (run_ic.cpp:136-159, gpu_ic.cpp:44-51, boundary_daemon.py:119-158):
  header [xlo xhi ylo yhi zlo zhi n], then n rows [x y z vx vy vz mass u mat fixed].
* snapshot (run_ic.cpp:15-30, gpu_ic.cpp:118-123; read by boundary_daemon.py:181-212):
  int64 n, then n rows [x y z vx vy vz u rho h mat].  Written here in IC order.
* checkpoint (run_ic.cpp:36-55 / 76-102): header [xlo xhi ylo yhi zlo zhi n t dx],
  then n rows [x y z vx vy vz mass u mat fixed D eps_act Sxx Sxy Sxz Syy Syz Szz],
  written to <file>.tmp and renamed.
* extended checkpoint (this project): the checkpoint rows followed by n rows
  [h ax ay az dudt dSdt_xx dSdt_xy dSdt_xz dSdt_yy dSdt_yz dSdt_zz] and a trailer
  [magic steps dt_next] (dt_next: the CFL step of the last force evaluation), so a
  restart continues bit for bit (the reference checkpoint omits h, acc and the
  previous derivatives, so its restart re-runs init()).  Files with the older
  two-value trailer [magic0 steps] are still read.

Tensors inside the files use the reference's {xx, xy, xz, yy, yz, zz} order;
the in-memory arrays of this package use the Sym3 order {xx, yy, zz, xy, xz, yz}.
"""
from __future__ import annotations

import os

import numpy as np

# file order {xx,xy,xz,yy,yz,zz} -> memory order {xx,yy,zz,xy,xz,yz} and back
FILE_TO_MEM = [0, 3, 5, 1, 2, 4]
MEM_TO_FILE = [0, 3, 4, 1, 5, 2]
EXT_MAGIC = 20241204.0           # trailer [magic steps dt_next]
EXT_MAGIC_V0 = 20241203.0        # trailer [magic steps]


def write_ic(path: str, ic: dict) -> None:
    n = len(ic["mass"])
    hdr = np.array([ic["lo"][0], ic["hi"][0], ic["lo"][1], ic["hi"][1], ic["lo"][2], ic["hi"][2], n],
                   np.float64)
    rows = np.empty((n, 10), np.float64)
    rows[:, 0:3] = ic["pos"]
    rows[:, 3:6] = ic["vel"]
    rows[:, 6] = ic["mass"]
    rows[:, 7] = ic["u"]
    rows[:, 8] = ic["mat"]
    rows[:, 9] = ic["fixed"]
    with open(path, "wb") as f:
        hdr.astype("<f8").tofile(f)
        rows.astype("<f8").tofile(f)


def read_ic(path: str) -> dict:
    with open(path, "rb") as f:
        hdr = np.fromfile(f, "<f8", 7)
        if hdr.size != 7:
            raise ValueError(f"{path}: bad header")
        n = int(round(hdr[6]))
        rows = np.fromfile(f, "<f8", 10 * n)
        if rows.size != 10 * n:
            raise ValueError(f"{path}: short read ({rows.size // 10} of {n} rows)")
    rows = rows.reshape(n, 10)
    return dict(pos=rows[:, 0:3].copy(), vel=rows[:, 3:6].copy(), mass=rows[:, 6].copy(),
                u=rows[:, 7].copy(), mat=rows[:, 8].astype(np.int32), fixed=(rows[:, 9] != 0).astype(np.uint8),
                lo=(hdr[0], hdr[2], hdr[4]), hi=(hdr[1], hdr[3], hdr[5]))


def write_snapshot(path: str, pos, vel, u, rho, h, mat) -> None:
    n = len(u)
    rows = np.empty((n, 10), np.float64)
    rows[:, 0:3] = pos
    rows[:, 3:6] = vel
    rows[:, 6] = u
    rows[:, 7] = rho
    rows[:, 8] = h
    rows[:, 9] = mat
    with open(path, "wb") as f:
        np.array([n], "<i8").tofile(f)
        rows.astype("<f8").tofile(f)


def read_snapshot(path: str) -> dict:
    with open(path, "rb") as f:
        n = int(np.fromfile(f, "<i8", 1)[0])
        rows = np.fromfile(f, "<f8", 10 * n)
    if rows.size != 10 * n:
        raise ValueError(f"{path}: short read")
    rows = rows.reshape(n, 10)
    return dict(pos=rows[:, 0:3].copy(), vel=rows[:, 3:6].copy(), u=rows[:, 6].copy(),
                rho=rows[:, 7].copy(), h=rows[:, 8].copy(), mat=rows[:, 9].astype(np.int32))


def _atomic_write(path: str, chunks) -> None:
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        for c in chunks:
            np.asarray(c, "<f8").tofile(f)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def write_checkpoint(path: str, st: dict, lo, hi, t: float, dx: float, extended: bool = False,
                     steps: int = 0, dt_next: float = 0.0) -> None:
    """``st`` holds IC-order arrays: pos vel mass u mat fixed D eps_act S (memory order),
    and for ``extended`` also hh acc dudt dSdt."""
    n = len(st["mass"])
    hdr = [lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], n, t, dx]
    rows = np.empty((n, 18), np.float64)
    rows[:, 0:3] = st["pos"]
    rows[:, 3:6] = st["vel"]
    rows[:, 6] = st["mass"]
    rows[:, 7] = st["u"]
    rows[:, 8] = st["mat"]
    rows[:, 9] = st["fixed"]
    rows[:, 10] = st["D"]
    rows[:, 11] = st["eps_act"]
    rows[:, 12:18] = np.asarray(st["S"]).reshape(n, 6)[:, MEM_TO_FILE]
    chunks = [hdr, rows]
    if extended:
        ext = np.empty((n, 11), np.float64)
        ext[:, 0] = st["hh"]
        ext[:, 1:4] = st["acc"]
        ext[:, 4] = st["dudt"]
        ext[:, 5:11] = np.asarray(st["dSdt"]).reshape(n, 6)[:, MEM_TO_FILE]
        chunks += [ext, [EXT_MAGIC, steps, dt_next]]
    _atomic_write(path, chunks)


def read_checkpoint(path: str) -> dict:
    raw = np.fromfile(path, "<f8")
    if raw.size < 9:
        raise ValueError(f"{path}: bad header")
    hdr = raw[:9]
    n = int(round(hdr[6]))
    body = raw[9:9 + 18 * n]
    if body.size != 18 * n:
        raise ValueError(f"{path}: short read")
    rows = body.reshape(n, 18)
    out = dict(lo=(hdr[0], hdr[2], hdr[4]), hi=(hdr[1], hdr[3], hdr[5]), t=float(hdr[7]), dx=float(hdr[8]),
               pos=rows[:, 0:3].copy(), vel=rows[:, 3:6].copy(), mass=rows[:, 6].copy(), u=rows[:, 7].copy(),
               mat=rows[:, 8].astype(np.int32), fixed=(rows[:, 9] != 0).astype(np.uint8), D=rows[:, 10].copy(),
               eps_act=rows[:, 11].copy(), S=rows[:, 12:18][:, FILE_TO_MEM].copy(), extended=False)
    rest = raw[9 + 18 * n:]
    trailer = None
    if rest.size == 11 * n + 3 and rest[-3] == EXT_MAGIC:
        trailer = dict(steps=int(rest[-2]), dt_next=float(rest[-1]))
    elif rest.size == 11 * n + 2 and rest[-2] == EXT_MAGIC_V0:
        trailer = dict(steps=int(rest[-1]), dt_next=0.0)
    if trailer is not None:
        ext = rest[:11 * n].reshape(n, 11)
        out.update(hh=ext[:, 0].copy(), acc=ext[:, 1:4].copy(), dudt=ext[:, 4].copy(),
                   dSdt=ext[:, 5:11][:, FILE_TO_MEM].copy(), extended=True, **trailer)
    elif rest.size:
        raise ValueError(f"{path}: unrecognised trailing data")
    return out
