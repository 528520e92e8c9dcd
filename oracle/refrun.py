"""The reference program's own Run, built from its sources (oracle/Makefile
target refrun), and a reader of the state dumps its driver writes.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_run_vectors.py (which
writes the committed fixtures from it) and by tests that compare the oracle
with it live where it is built; never by the kestrel_amd package.

The driver (oracle/ref_run.f90) runs src/main.f90's set-up sequence and calls
Run (src/TimeStepper.f90:73-113) once per output interval; after interval k
it writes <prefix>_<k>.bin (and <prefix>_0.bin before the first), holding
grid%t, grid%dt, the active and ghost tile lists, every tile's state and, per
active tile, u(1:13,:,:), b0, bt and the maxima; per ghost tile u and bt.
The reference prints one line 't = ..., dt = ...' per accepted step
(TimestepMessage, src/Messages.f90:275-283), which gives the accepted-step
count of every interval.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(_HERE, "_ref", "run", "ref_run")
TIME_BIN = os.path.join(_HERE, "_ref", "time", "ref_run")

# u(1:13) field order (src/main.f90:76-103)
F13 = ("w", "rhoHnu", "rhoHnv", "Hnpsi", "Hn", "u", "v", "psi", "rho", "b0", "bt", "dbdx", "dbdy")
MAXIMA = ("Hnmax", "Hnmax_t", "umax", "umax_t", "emax", "emax_t", "dmax", "dmax_t",
          "psimax", "psimax_t", "tfirst")


def available(binary: str = BIN) -> bool:
    return os.path.exists(binary)


def read_dump(path: str) -> dict:
    """One state dump; tile ids 0-based; per-tile arrays indexed [q, j, i]."""
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=off)
        off += a.nbytes
        return a

    ver, nt, nx, ny, na, ng, ntx, nty = (int(v) for v in take("<i4", 8))
    if ver != 1:
        raise ValueError(f"{path}: dump version {ver}")
    t, dt = (float(v) for v in take("<f8", 2))
    act = take("<i4", na).astype(np.int64) - 1
    gho = take("<i4", ng).astype(np.int64) - 1
    st = take("<i4", nt).astype(np.uint8)
    u = np.zeros((na, 13, ny, nx))
    b0 = np.zeros((na, ny + 1, nx + 1))
    bt = np.zeros((na, ny + 1, nx + 1))
    mx = np.zeros((na, 11, ny, nx))
    for q in range(na):
        u[q] = take("<f8", 13 * nx * ny).reshape(ny, nx, 13).transpose(2, 0, 1)
        b0[q] = take("<f8", (nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
        bt[q] = take("<f8", (nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
        for k in range(5):
            m = take("<f8", 2 * nx * ny).reshape(2, ny, nx)
            mx[q, 2 * k] = m[0]
            mx[q, 2 * k + 1] = m[1]
        mx[q, 10] = take("<f8", nx * ny).reshape(ny, nx)
    gu = np.zeros((ng, 13, ny, nx))
    gbt = np.zeros((ng, ny + 1, nx + 1))
    for q in range(ng):
        gu[q] = take("<f8", 13 * nx * ny).reshape(ny, nx, 13).transpose(2, 0, 1)
        gbt[q] = take("<f8", (nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    return dict(t=t, dt=dt, active=act, ghost=gho, tile_state=st, nx=nx, ny=ny,
                ntx=ntx, nty=nty, u=u, b0=b0, bt=bt, maxima=mx, ghost_u=gu, ghost_bt=gbt)


def to_global(d: dict, tiles=None, which: str = "u") -> np.ndarray:
    """Scatter per-tile cell arrays [q, k, ny, nx] of the listed tiles into
    global [k, NY, NX] arrays (zero outside)."""
    tiles = d["active"] if tiles is None else tiles
    a = d[which]
    nx, ny, ntx, nty = d["nx"], d["ny"], d["ntx"], d["nty"]
    out = np.zeros((a.shape[1], nty * ny, ntx * nx))
    for q, k in enumerate(tiles):
        gi, gj = int(k) % ntx, int(k) // ntx
        out[:, gj * ny:(gj + 1) * ny, gi * nx:(gi + 1) * nx] = a[q]
    return out


def write_b0(path: str, b0v: np.ndarray, x0: float, y0: float, dx: float, dy: float) -> None:
    """Vertex elevations in the driver's sampled-topography format."""
    b0v = np.ascontiguousarray(b0v, dtype="<f8")
    nyv, nxv = b0v.shape
    with open(path, "wb") as f:
        f.write(np.array([nxv, nyv], dtype="<i4").tobytes())
        f.write(np.array([x0, y0, dx, dy], dtype="<f8").tobytes())
        f.write(b0v.tobytes())


def params(deck_path: str, workdir: str, binary: str = BIN) -> dict:
    """The RunSet scalars the reference reads from a deck (ref_run -params):
    name -> float, bool, str, or the four tile counts."""
    if not os.path.exists(binary):
        raise RuntimeError(f"{binary} is not built (make -C oracle refrun)")
    os.makedirs(workdir, exist_ok=True)
    p = subprocess.run([binary, os.path.abspath(deck_path), os.path.join(workdir, "out"), "-params"],
                       cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=60)
    out = {}
    for line in p.stdout.decode("latin-1").splitlines():
        line = line.strip()
        if not line.startswith("param "):
            continue
        name, val = (x.strip() for x in line[6:].split("=", 1))
        if name == "tiles":
            out[name] = [int(v) for v in val.split()]
        elif val in ("T", "F"):
            out[name] = val == "T"
        else:
            try:
                out[name] = float(val.replace("E", "e"))
            except ValueError:
                out[name] = val
    if p.returncode != 0 or not out:
        raise RuntimeError(f"reference parameter read failed ({p.returncode})")
    return out


def run(deck_path: str, workdir: str, b0_file: str | None = None, binary: str = BIN,
        timeout: float = 3600.0, keep: bool = True, name: str = "dump") -> dict:
    """Run the reference on a deck; returns the dumps (the set-up state, then
    one per interval run: 1..Nout, or on a restart from the restart file's
    number + 1), the accepted-step count and last printed (t, dt) of every
    interval, the number of revised-dt messages and the wall time.  The
    output folder is workdir/out."""
    import time
    if not os.path.exists(binary):
        raise RuntimeError(f"{binary} is not built (make -C oracle refrun)")
    os.makedirs(workdir, exist_ok=True)
    prefix = os.path.join(workdir, name)
    cmd = [binary, os.path.abspath(deck_path), os.path.join(workdir, "out"), prefix]
    if b0_file is not None:
        cmd.append(b0_file)
    t0 = time.perf_counter()
    p = subprocess.run(cmd, cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       timeout=timeout)
    wall = time.perf_counter() - t0
    text = p.stdout.decode("latin-1")
    if p.returncode != 0:
        raise RuntimeError(f"reference run failed ({p.returncode}):\n{text[-2000:]}")
    steps, revised, cur = [], [], 0
    nrev = 0
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("t = "):
            cur += 1
        elif s.startswith("revised nextT"):
            nrev += 1
        elif s.startswith("ref_run: interval"):
            steps.append(cur)
            revised.append(nrev)
            cur, nrev = 0, 0
    import re
    ks = sorted(int(m.group(1)) for f in os.listdir(workdir)
                if (m := re.fullmatch(re.escape(name) + r"_(\d+)\.bin", f)))
    dumps = []
    for k in ks:
        dumps.append(read_dump(f"{prefix}_{k}.bin"))
        if not keep:
            os.remove(f"{prefix}_{k}.bin")
    return dict(dumps=dumps, steps=steps, revised=revised, wall=wall, log=text, intervals=ks[1:])
