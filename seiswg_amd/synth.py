"""Seeded synthetic problems for benchmarks and large-size tests.

Layered-plus-random-field elastic models on a regular grid (k = i*nz + j,
z fastest, as the proc000000_*.bin files of the reference store them,
src/bin/gen_model.rs:42-46):

* vp rises linearly with depth from vp_top to vp_bottom;
* a Gaussian random field g (white N(0,1) from numpy's PCG64 with the given
  seed, smoothed with a separable Gaussian of standard deviation corr_m,
  scaled to max|g| = 1) perturbs it: vp <- vp * (1 + pert * g);
* vs = vp / 1.73 and rho = 310 * vp**0.25 (Gardner).

Writers produce the reference's on-disk inputs (model files, sources.dat,
stations.dat, config.ini) so the same problems run through the CLI.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np
from scipy.ndimage import gaussian_filter


def layered_random_model(nx: int, nz: int, h: float, seed: int, vp_top: float = 2000.0,
                         vp_bottom: float = 4500.0, pert: float = 0.05, corr_m: float = 200.0):
    rng = np.random.default_rng(seed)
    depth = np.arange(nz) / max(nz - 1, 1)
    vp = np.broadcast_to(vp_top + (vp_bottom - vp_top) * depth, (nx, nz)).astype(np.float64)
    g = gaussian_filter(rng.standard_normal((nx, nz)), sigma=corr_m / h, mode="reflect")
    g /= np.abs(g).max()
    vp = vp * (1.0 + pert * g)
    vs = vp / 1.73
    rho = 310.0 * vp ** 0.25
    f = lambda a: np.ascontiguousarray(a, np.float32).ravel()
    return {"vp": f(vp), "vs": f(vs), "rho": f(rho)}


def gaussian_anomalies(model: dict, nx: int, nz: int, h: float, seed: int, n: int = 3, amp: float = 0.08,
                       radius_m: float = 150.0, margin: int = 50):
    """True model of the gradient workload: vs (and through it mu) times
    (1 +/- amp * exp(-r^2 / radius^2)) around n seeded interior points."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    fac = np.ones((nx, nz))
    for q in range(n):
        ci, cj = rng.integers(margin, nx - margin), rng.integers(margin, nz - margin)
        sign = 1.0 if q % 2 == 0 else -1.0
        r2 = ((i - ci) ** 2 + (j - cj) ** 2) * h * h
        fac *= 1.0 + sign * amp * np.exp(-r2 / radius_m ** 2)
    out = dict(model)
    out["vs"] = np.ascontiguousarray(model["vs"].reshape(nx, nz) * fac, np.float32).ravel()
    return out


@dataclass
class Geometry:
    src_x: list = field(default_factory=list)
    src_z: list = field(default_factory=list)
    rec_x: list = field(default_factory=list)
    rec_z: list = field(default_factory=list)
    f0: float = 8.0
    t0: float = 0.15
    ang: float = 90.0
    amp: float = 1e10

    def src_rows(self):
        return [[x, z, 0, self.f0, self.t0, self.ang, self.amp] for x, z in zip(self.src_x, self.src_z)]


def line_geometry(nx: int, h: float, nshots: int, nrec: int, src_depth_m: float = 50.0, rec_depth_m: float = 10.0,
                  rec_first: int = 2, rec_step: int | None = None, **kw) -> Geometry:
    """Shots evenly spaced along x at src_depth; receivers on a line at rec_depth."""
    g = Geometry(**kw)
    g.src_x = [float(round((s + 0.5) * nx / nshots) * h) for s in range(nshots)]
    g.src_z = [src_depth_m] * nshots
    step = rec_step if rec_step is not None else max((nx - 2 * rec_first) // max(nrec, 1), 1)
    g.rec_x = [float((rec_first + step * r) * h) for r in range(nrec)]
    g.rec_z = [rec_depth_m] * nrec
    return g


def write_model_dir(d: str, nx: int, nz: int, h: float, model: dict):
    os.makedirs(d, exist_ok=True)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    arrs = dict(model, x=(i * h).astype(np.float32).ravel(), z=(j * h).astype(np.float32).ravel())
    for name, a in arrs.items():
        with open(os.path.join(d, f"proc000000_{name}.bin"), "wb") as f:
            np.array([a.size], np.int32).tofile(f)
            np.ascontiguousarray(a, np.float32).tofile(f)


def write_tables(d: str, geom: Geometry):
    with open(os.path.join(d, "sources.dat"), "w") as f:
        for r in geom.src_rows():
            f.write(" ".join(repr(float(v)) for v in r) + "\n")
    with open(os.path.join(d, "stations.dat"), "w") as f:
        for x, z in zip(geom.rec_x, geom.rec_z):
            f.write(f"{x!r} {z!r}\n")


# BASELINE workloads ---------------------------------------------------------
C2 = dict(nx=2000, nz=1000, h=10.0, nshots=64, nrec=400, nt=4000, dt=1e-3, f0=8.0, t0=0.15, ang=90.0,
          abs_width=40, seed=2)


def c2_problem():
    """Batched multi-shot PSV forward: 2000x1000, 64 shots at 50 m depth,
    400 receivers at 10 m (ix = 5r + 2), taper left/right/bottom width 40."""
    c = C2
    model = layered_random_model(c["nx"], c["nz"], c["h"], c["seed"])
    geom = line_geometry(c["nx"], c["h"], c["nshots"], c["nrec"], rec_first=2, rec_step=5, f0=c["f0"], t0=c["t0"],
                         ang=c["ang"])
    return model, geom


C3 = dict(nx=1000, nz=600, h=10.0, nshots=16, nrec=200, nt=3000, dt=1e-3, f0=8.0, t0=0.15, ang=90.0,
          abs_width=40, seed=2, anomaly_seed=3, ckpt=50)


def c3_problem():
    """Checkpointed P-SV gradient workload: 1000x600 initial model (the C2
    recipe, seed 2), true model with three +/-8% vs Gaussian anomalies of
    radius 150 m (seed 3); 16 shots at 50 m depth and 200 receivers at 10 m,
    evenly spaced."""
    c = C3
    init = layered_random_model(c["nx"], c["nz"], c["h"], c["seed"])
    true = gaussian_anomalies(init, c["nx"], c["nz"], c["h"], c["anomaly_seed"])
    step = (c["nx"] - 4) // c["nrec"]
    geom = line_geometry(c["nx"], c["h"], c["nshots"], c["nrec"], rec_first=2, rec_step=step, f0=c["f0"],
                         t0=c["t0"], ang=c["ang"])
    return init, true, geom
