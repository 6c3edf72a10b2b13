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


# benchmark workloads -------------------------------------------------------
PSV_BATCH = dict(nx=2000, nz=1000, h=10.0, nshots=64, nrec=400, nt=4000, dt=1e-3, f0=8.0, t0=0.15, ang=90.0,
          abs_width=40, seed=2)


def psv_batch_problem():
    """Batched multi-shot PSV forward: 2000x1000, 64 shots at 50 m depth,
    400 receivers at 10 m (ix = 5r + 2), taper left/right/bottom width 40."""
    c = PSV_BATCH
    model = layered_random_model(c["nx"], c["nz"], c["h"], c["seed"])
    geom = line_geometry(c["nx"], c["h"], c["nshots"], c["nrec"], rec_first=2, rec_step=5, f0=c["f0"], t0=c["t0"],
                         ang=c["ang"])
    return model, geom


PSV_GRADIENT = dict(nx=1000, nz=600, h=10.0, nshots=16, nrec=200, nt=3000, dt=1e-3, f0=8.0, t0=0.15, ang=90.0,
          abs_width=40, seed=2, anomaly_seed=3, ckpt=50, smooth=3.0)


def gradient_problem():
    """Checkpointed P-SV gradient workload: 1000x600 initial model (the PSV_BATCH
    recipe, seed 2), true model with three +/-8% vs Gaussian anomalies of
    radius 150 m (seed 3); 16 shots at 50 m depth and 200 receivers at 10 m,
    evenly spaced."""
    c = PSV_GRADIENT
    init = layered_random_model(c["nx"], c["nz"], c["h"], c["seed"])
    true = gaussian_anomalies(init, c["nx"], c["nz"], c["h"], c["anomaly_seed"])
    step = (c["nx"] - 4) // c["nrec"]
    geom = line_geometry(c["nx"], c["h"], c["nshots"], c["nrec"], rec_first=2, rec_step=step, f0=c["f0"],
                         t0=c["t0"], ang=c["ang"])
    return init, true, geom


UNEVEN = dict(nproblems=48, seed=4, nx=(384, 3072), nz=(384, 1536), shots=(4, 16), nrec=(100, 300), nt=(1500, 4000),
          f0=(5.0, 15.0), h=10.0, cfl=0.5, abs_width=40, ckpt=50)


def uneven_subset(k: int = 8, seed: int = 7):
    """Seeded subset of the uneven batch's problem indices (k of 48, sorted):
    the problems measured at their full nt (1500-4000)."""
    return sorted(np.random.default_rng(seed).choice(UNEVEN["nproblems"], k, replace=False).tolist())


# a batch of small uneven problems (same recipe, grids 128-448): the sizes at
# which merging problems into shared launches pays
UNEVEN_SMALL = dict(UNEVEN, seed=5, nx=(128, 448), nz=(128, 448), shots=(2, 6), nrec=(40, 100), abs_width=20)


def uneven_problems(nt_scale: float = 1.0, nproblems: int | None = None, spec: dict | None = None,
                    only=None):
    """Uneven batch of independent SH and P-SV gradient problems (seeded):
    grid nx in [384, 3072], nz in [384, 1536], each with its own
    layered-plus-random-field model (the PSV_BATCH recipe, seed 100+q) and a true
    model with Gaussian vs anomalies; 4-16 shots at 50 m depth, 100-300
    receivers at 10 m; nt in [1500, 4000] (times nt_scale, at least 60);
    dt from CFL 0.5 on the model's largest vp; Ricker f0 in [5, 15] Hz,
    t0 = 1.2 / f0.  Even problems are SH, odd ones P-SV.  `spec` replaces
    these ranges (UNEVEN_SMALL: the small-problem batch).  `only` (indices)
    builds just those problems, with the same draws as the whole batch.
    Returns a list of dicts: kind, nx, nz, h, dt, nt, init, true, geom, index."""
    c = spec or UNEVEN
    rng = np.random.default_rng(c["seed"])
    out = []
    for q in range(nproblems or c["nproblems"]):
        nx = int(rng.integers(c["nx"][0] // 8, c["nx"][1] // 8 + 1)) * 8
        nz = int(rng.integers(c["nz"][0] // 8, c["nz"][1] // 8 + 1)) * 8
        nshots = int(rng.integers(c["shots"][0], c["shots"][1] + 1))
        nrec = int(rng.integers(c["nrec"][0], c["nrec"][1] + 1))
        nt = max(60, int(round(int(rng.integers(c["nt"][0], c["nt"][1] + 1)) * nt_scale)))
        f0 = float(rng.uniform(*c["f0"]))
        if only is not None and q not in only:
            continue
        init = layered_random_model(nx, nz, c["h"], 100 + q)
        true = gaussian_anomalies(init, nx, nz, c["h"], 200 + q)
        vmax = float(max(init["vp"].max(), true["vp"].max()))
        dt = float(np.float32(c["cfl"] * c["h"] / vmax))
        step = max((nx - 4) // nrec, 1)
        geom = line_geometry(nx, c["h"], nshots, nrec, rec_first=2, rec_step=step, f0=f0, t0=1.2 / f0, ang=90.0)
        out.append(dict(kind="sh" if q % 2 == 0 else "psv", nx=nx, nz=nz, h=c["h"], dt=dt, nt=nt, init=init,
                        true=true, geom=geom, index=q))
    return out


def sh_forward_problem():
    """Plain SH forward: homogeneous 200x200, dx = dz = 100 m, vs = 3000,
    rho = 2600 (vp = 5200 stored), nt = 2000, dt = 0.006 s; one Ricker source
    at the centre (f0 = 2 Hz, t0 = 0.6 s, angle 0, amplitude 1e10); 25
    receivers at iz = 40, ix = 28 + 6r; absorbing on all edges, width 20."""
    nx = nz = 200
    model = {"rho": np.full(nx * nz, 2600.0, np.float32), "vp": np.full(nx * nz, 5200.0, np.float32),
             "vs": np.full(nx * nz, 3000.0, np.float32)}
    g = Geometry(src_x=[10000.0], src_z=[10000.0], rec_x=[float((28 + 6 * r) * 100) for r in range(25)],
                 rec_z=[4000.0] * 25, f0=2.0, t0=0.6, ang=0.0, amp=1e10)
    return dict(nx=nx, nz=nz, h=100.0, nt=2000, dt=0.006, abs_width=20, model=model, geom=g)


def spin_forward_problem(j_value: float = 1.0e4):
    """P-SV + micropolar spin forward: homogeneous 500x500, dx = dz = 5 m,
    lambda = 2e10, mu = 1.5e10, nu = 2e9, lambda_c = mu_c = nu_c = 1e5,
    rho = 2500, nt = 3000, dt = 4e-4; Ricker at the centre (f0 = 25 Hz,
    t0 = 0.05 s, 45 degrees, 1e10); 7 receivers on the 45-degree ray
    (ix = iz = 270 + 20n); absorbing on all edges, width 40.  The
    micro-inertia j: with the stated j = 1e-2 the rotation mode has
    omega*dt ~ 358 and the explicit scheme diverges; j_value (default 1e4,
    omega*dt ~ 0.36) keeps it stable."""
    nx = nz = 500
    n = nx * nz
    f = lambda v: np.full(n, v, np.float32)
    model = {"rho": f(2500.0), "lambda": f(2.0e10), "mu": f(1.5e10), "nu": f(2.0e9), "j": f(j_value),
             "lambda_c": f(1.0e5), "mu_c": f(1.0e5), "nu_c": f(1.0e5)}
    g = Geometry(src_x=[1250.0], src_z=[1250.0], rec_x=[float((270 + 20 * k) * 5) for k in range(7)],
                 rec_z=[float((270 + 20 * k) * 5) for k in range(7)], f0=25.0, t0=0.05, ang=45.0, amp=1e10)
    return dict(nx=nx, nz=nz, h=5.0, nt=3000, dt=4e-4, abs_width=40, model=model, geom=g)
