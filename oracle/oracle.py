"""ctypes wrapper of the CPU oracle (fd2d_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, never by the product package.

Besides the kernel-level oracle it restates the host-side set-up of
src/solver.rs that decides the inputs of the time loop: model-file parsing
(solver.rs:220-233), text tables (248-255), grid inference (640-658),
grid-index rounding (709-719, 745-751) and the Ricker table (350-356,
721-739).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_fd2d.so")

# field ids (fd2d.wgsl:53-105)
F = dict(VY=0, UY=1, SXY=2, SZY=3, DSY=4, DVYDX=5, DVYDZ=6, VX=7, VZ=8, UX=9, UZ=10,
         SXX=11, SZZ=12, SXZ=13, DSX=14, DSZ=15, DVXDX=16, DVXDZ=17, DVZDX=18, DVZDZ=19,
         VY_C=20, UY_C=21, SYX_C=22, SYY_C=23, SYZ_C=24, DSY_C=25, DVYDX_C=26, DVYDZ_C=27,
         DUZDX=28, DUXDZ=29, LAM=30, MU=31, NU=32, J=33, LAM_C=34, MU_C=35, NU_C=36, RHO=37,
         BOUND=38, K_LAM=39, K_MU=40, K_RHO=41, GSUM=42, GTMP=43)
OBS_X, OBS_Y, OBS_Z, OBS_YI, OBS_YC = range(5)


class or_cfg(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nz", C.c_int32), ("nt", C.c_int32),
                ("dx", C.c_float), ("dz", C.c_float), ("dt", C.c_float),
                ("sh", C.c_int32), ("psv", C.c_int32), ("spin", C.c_int32),
                ("abs_left", C.c_int32), ("abs_right", C.c_int32), ("abs_bottom", C.c_int32),
                ("abs_top", C.c_int32), ("abs_width", C.c_int32), ("abs_alpha", C.c_float),
                ("nsrc", C.c_int32), ("nrec", C.c_int32)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        fp = C.POINTER(C.c_float)
        L.or_new.restype = C.c_void_p
        L.or_new.argtypes = [C.c_int, C.POINTER(or_cfg)]
        L.or_free.argtypes = [C.c_void_p]
        for n in ("or_set_field", "or_get_field"):
            getattr(L, n).argtypes = [C.c_void_p, C.c_int32, dp]
        L.or_set_idata.argtypes = [C.c_void_p, ip]
        for n in ("or_set_stf", "or_get_obs", "or_set_obs"):
            getattr(L, n).argtypes = [C.c_void_p, dp]
        for n in ("or_init_model", "or_clear_dyn", "or_clear_obs"):
            getattr(L, n).argtypes = [C.c_void_p]
        L.or_step.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.or_adjoint_step_sh.argtypes = [C.c_void_p, C.c_int32]
        L.or_smooth.argtypes = [C.c_void_p, C.c_double]
        L.or_forward.restype = C.c_int32
        L.or_forward.argtypes = [C.c_void_p, C.c_int32, C.c_int32, dp, C.c_int32]
        L.or_adjoint_plain.restype = C.c_int32
        L.or_adjoint_plain.argtypes = [C.c_void_p, C.c_int32, ip, dp, C.c_int32, C.c_double, dp]
        L.or_ricker_table.argtypes = [C.c_int, C.c_int32, C.c_int32, C.c_float, fp, dp]
        L.or_infer_grid.argtypes = [C.c_int32, fp, fp, ip, ip, fp, fp]
        L.or_num_threads.restype = C.c_int32
        L.or_set_num_threads.argtypes = [C.c_int32]
        L.or_set_num_threads.restype = None
        L.or_objective.restype = C.c_double
        L.or_objective.argtypes = [C.c_void_p, C.c_int32, ip, dp]
        L.or_gradient.restype = C.c_double
        L.or_gradient.argtypes = [C.c_void_p, C.c_int32, ip, dp, dp, dp, dp]
        L.or_misfit.restype = C.c_double
        L.or_misfit.argtypes = [C.c_int, dp, dp, C.c_int32, C.c_int32, C.c_float, dp]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- host set-up
def read_model_field(path: str) -> np.ndarray:
    """i32 npt header + npt f32 (solver.rs:220-233); trailing bytes ignored."""
    b = open(path, "rb").read()
    if len(b) < 4:
        raise ValueError("Model data too small")
    n = int(np.frombuffer(b[:4], "<i4")[0])
    if len(b) < 4 + 4 * n:
        raise ValueError("Model data too short")
    return np.frombuffer(b[4:4 + 4 * n], "<f4").copy()


def read_table(path: str) -> list[list[float]]:
    """Whitespace table; '#' lines and blank lines skipped; bad tokens -> 0.0
    (solver.rs:248-255)."""
    rows = []
    for line in open(path).read().splitlines():
        t = line.strip()
        if not t or t.startswith("#"):
            continue
        row = []
        for tok in line.split():
            try:
                row.append(float(tok))
            except ValueError:
                row.append(0.0)
        rows.append(row)
    return rows


def infer_grid(x: np.ndarray, z: np.ndarray):
    L = lib()
    x = np.ascontiguousarray(x, np.float32)
    z = np.ascontiguousarray(z, np.float32)
    nx, nz = C.c_int32(), C.c_int32()
    dx, dz = C.c_float(), C.c_float()
    L.or_infer_grid(len(x), x.ctypes.data_as(C.POINTER(C.c_float)), z.ctypes.data_as(C.POINTER(C.c_float)),
                    C.byref(nx), C.byref(nz), C.byref(dx), C.byref(dz))
    return nx.value, nz.value, np.float32(dx.value), np.float32(dz.value)


def _round_half_away(v: np.float32) -> int:
    v = float(v)
    r = np.floor(abs(v) + 0.5)
    return int(np.copysign(r, v))


def grid_index(x: float, z: float, dx: np.float32, dz: np.float32, nz: int) -> int:
    """ix = round(x/dx), iz = round(z/dz) in f32, u32 cast (solver.rs:716-718)."""
    ix = _round_half_away(np.float32(np.float32(x) / dx))
    iz = _round_half_away(np.float32(np.float32(z) / dz))
    ix, iz = max(ix, 0), max(iz, 0)
    return (ix * nz + iz) & 0xFFFFFFFF


def ricker_table(prec: int, src_rows, nt: int, dt: float) -> np.ndarray:
    """[3][nsrc][nt] Ricker table (solver.rs:350-356, 721-739)."""
    L = lib()
    nsrc = len(src_rows)
    p = np.zeros((max(nsrc, 1), 4), np.float32)
    for s, row in enumerate(src_rows):
        get = lambda n, d: row[n] if len(row) > n else d
        p[s] = [get(3, 10.0), get(4, 0.1), get(5, 0.0), get(6, 1.0e10)]
    out = np.zeros((3, nsrc, nt), np.float64)
    L.or_ricker_table(prec, nsrc, nt, C.c_float(dt), p.ctypes.data_as(C.POINTER(C.c_float)),
                      out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def misfit(prec: int, syn, obs, dt: float):
    L = lib()
    syn, ps = _d(syn)
    obs, po = _d(obs)
    nrec, nt = syn.shape
    ad = np.zeros_like(syn)
    m = L.or_misfit(prec, ps, po, nrec, nt, C.c_float(dt), ad.ctypes.data_as(C.POINTER(C.c_double)))
    return m, ad


@dataclass
class Problem:
    """Inputs of one time-loop problem, as the reference's Solver::new builds them."""
    nx: int
    nz: int
    nt: int
    dx: float
    dz: float
    dt: float
    sh: bool = False
    psv: bool = True
    spin: bool = False
    abs_left: bool = False
    abs_right: bool = False
    abs_bottom: bool = False
    abs_top: bool = False
    abs_width: int = 20
    abs_alpha: float = 0.015
    # model: dict name -> f32[npt]; names rho, vp, vs (not spin) or
    # rho, lambda, mu, nu, j, lambda_c, mu_c, nu_c (spin)
    model: dict = field(default_factory=dict)
    src_rows: list = field(default_factory=list)   # rows x z type f0 t0 ang amp
    src_idx: list = field(default_factory=list)
    rec_idx: list = field(default_factory=list)

    @property
    def npt(self):
        return self.nx * self.nz


def load_problem(model_dir: str, sources: str, stations: str, ini: dict) -> Problem:
    """Mirror of Solver::new (solver.rs:578-760) for the oracle's inputs."""
    spin = ini.get("spin", False)
    names = ["rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c"] if spin else ["rho", "vp", "vs"]
    model = {n: read_model_field(os.path.join(model_dir, f"proc000000_{n}.bin")) for n in names}
    x = read_model_field(os.path.join(model_dir, "proc000000_x.bin"))
    z = read_model_field(os.path.join(model_dir, "proc000000_z.bin"))
    nx, nz, dx, dz = infer_grid(x, z)
    src_rows = read_table(sources)
    rec_rows = read_table(stations)
    g = lambda r, n: r[n] if len(r) > n else 0.0
    src_idx = [grid_index(g(r, 0), g(r, 1), dx, dz, nz) for r in src_rows]
    rec_idx = [grid_index(g(r, 0), g(r, 1), dx, dz, nz) for r in rec_rows]
    p = Problem(nx=nx, nz=nz, nt=ini["nt"], dx=float(dx), dz=float(dz), dt=ini["dt"],
                sh=ini.get("sh", False), psv=ini.get("psv", True), spin=spin,
                abs_left=ini.get("abs_left", False), abs_right=ini.get("abs_right", False),
                abs_bottom=ini.get("abs_bottom", False), abs_top=ini.get("abs_top", False),
                abs_width=ini.get("abs_width", 20), abs_alpha=ini.get("abs_alpha", 0.015),
                model=model, src_rows=src_rows, src_idx=src_idx, rec_idx=rec_idx)
    return p


class Oracle:
    """One oracle instance (prec 4 = FP32 restatement, 8 = FP64 oracle)."""

    def __init__(self, p: Problem, prec: int = 4, stf: np.ndarray | None = None):
        self.p, self.prec = p, prec
        L = lib()
        c = or_cfg(p.nx, p.nz, p.nt, p.dx, p.dz, p.dt, int(p.sh), int(p.psv), int(p.spin),
                   int(p.abs_left), int(p.abs_right), int(p.abs_bottom), int(p.abs_top),
                   p.abs_width, p.abs_alpha, len(p.src_idx), len(p.rec_idx))
        self.h = L.or_new(prec, C.byref(c))
        self.L = L
        m = p.model
        if p.spin:
            slots = dict(LAM="lambda", MU="mu", NU="nu", J="j", LAM_C="lambda_c", MU_C="mu_c",
                         NU_C="nu_c", RHO="rho")
        else:
            slots = dict(LAM="vp", MU="vs", RHO="rho")
        for fid, name in slots.items():
            self.set_field(F[fid], m[name])
        idata = np.array(list(p.src_idx) + list(p.rec_idx), dtype=np.int64).astype(np.uint32).view(np.int32)
        idata = np.ascontiguousarray(idata)
        if len(idata):
            L.or_set_idata(self.h, idata.ctypes.data_as(C.POINTER(C.c_int32)))
        if stf is None:
            stf = ricker_table(prec, p.src_rows, p.nt, p.dt)
        if len(p.src_idx):
            _, sp = _d(stf)
            L.or_set_stf(self.h, sp)
        L.or_init_model(self.h)

    def __del__(self):
        try:
            self.L.or_free(self.h)
        except Exception:
            pass

    def set_field(self, fid: int, v):
        v, pv = _d(np.asarray(v).reshape(-1))
        assert v.size == self.p.npt
        self.L.or_set_field(self.h, fid, pv)

    def field(self, fid: int) -> np.ndarray:
        out = np.zeros(self.p.npt)
        self.L.or_get_field(self.h, fid, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def obs(self) -> np.ndarray:
        out = np.zeros((5, len(self.p.rec_idx), self.p.nt))
        self.L.or_get_obs(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_obs(self, v):
        v, pv = _d(v)
        self.L.or_set_obs(self.h, pv)

    def clear(self):
        self.L.or_clear_dyn(self.h)
        self.L.or_clear_obs(self.h)

    def step(self, it: int, isrc: int = -1):
        self.L.or_step(self.h, it, isrc)

    def forward(self, isrc: int = -1, snap: int = 0):
        """Returns (obs[5][nrec][nt], snaps[n][4][npt] for vy, vx, vz, vy_c)."""
        nsnap = (self.p.nt - 1) // snap if snap > 0 else 0
        snaps = np.zeros((max(nsnap, 1), 4, self.p.npt))
        n = self.L.or_forward(self.h, isrc, snap, snaps.ctypes.data_as(C.POINTER(C.c_double)), nsnap)
        return self.obs(), snaps[:n]

    def adjoint_plain(self, isrc_list, ref, adj_interval: int, smooth: float):
        isrc = np.ascontiguousarray(isrc_list, np.int32)
        ref, pr = _d(ref)
        mis = np.zeros(len(isrc))
        rc = self.L.or_adjoint_plain(self.h, len(isrc), isrc.ctypes.data_as(C.POINTER(C.c_int32)), pr,
                                     adj_interval, smooth, mis.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise ValueError(f"adjoint failed rc={rc}")
        return self.field(F["K_MU"]), mis

    def objective(self, isrc_list, obs) -> float:
        """J = sum (taper * (traces - obs))^2 over runs, components and time;
        obs: [nruns][ncomp][nrec][nt], ncomp 1 (SH: uy) or 2 (P-SV: ux, uz)."""
        isrc = np.ascontiguousarray(isrc_list, np.int32)
        obs, po = _d(obs)
        return self.L.or_objective(self.h, len(isrc), isrc.ctypes.data_as(C.POINTER(C.c_int32)), po)

    def gradient(self, isrc_list, obs):
        """(J, K_rho, K_mu, K_lambda): exact adjoint gradient of objective()."""
        isrc = np.ascontiguousarray(isrc_list, np.int32)
        obs, po = _d(obs)
        ks = [np.zeros(self.p.npt) for _ in range(3)]
        J = self.L.or_gradient(self.h, len(isrc), isrc.ctypes.data_as(C.POINTER(C.c_int32)), po,
                               *[k.ctypes.data_as(C.POINTER(C.c_double)) for k in ks])
        return J, ks[0], ks[1], ks[2]

    def smooth(self, sigma: float):
        self.L.or_smooth(self.h, sigma)


def smooth_array(o: Oracle, v, sigma: float) -> np.ndarray:
    """Gaussian smoothing (solver.rs:1249-1274) of an arbitrary field through
    the oracle's K_mu slot; the slot's previous contents are restored."""
    saved = o.field(F["K_MU"])
    o.set_field(F["K_MU"], v)
    o.smooth(sigma)
    out = o.field(F["K_MU"])
    o.set_field(F["K_MU"], saved)
    return out


def velocity_gradient(model: dict, k_mu, k_lam, psv: bool):
    """Gradient of the misfit with respect to (vs, vp) from the one with
    respect to (mu, lambda), through vps2lm (fd2d.wgsl:197-210):
    mu = rho vs^2 and lambda = rho (vp^2 - 2 vs^2) where vp > vs, else 0, so
        dJ/dvs = 2 rho vs K_mu - 4 rho vs K_lambda [vp > vs]
        dJ/dvp = 2 rho vp K_lambda [vp > vs]
    (SH: K_lambda = 0).  The chain rule of the inversion workflow."""
    rho, vp, vs = (np.asarray(model[k], np.float64) for k in ("rho", "vp", "vs"))
    k_mu = np.asarray(k_mu, np.float64)
    k_lam = np.asarray(k_lam, np.float64) if psv else np.zeros_like(k_mu)
    on = vp > vs
    gvs = 2.0 * rho * vs * k_mu + np.where(on, -4.0 * rho * vs * k_lam, 0.0)
    gvp = np.where(on, 2.0 * rho * vp * k_lam, 0.0)
    return gvs, gvp


def steepest_descent_iteration(p: Problem, runs, obs, step: float, tries: int, sigma: float):
    """One iteration of the inversion workflow (mode = inversion; the
    reference lists it as planned, README.md:61, without a specification, so
    the rules are this build's own, restated here in FP64):
    gradient = the exact adjoint one (or_gradient) with K_mu and K_lambda
    Gaussian-smoothed (sigma cells), taken to (vs, vp) by velocity_gradient;
    trial t (t = 0 .. tries-1) uses s = step / 2^t and moves
    vs <- max(vs - s max(vs) / max|g_vs| g_vs, 0) and likewise vp (P-SV only,
    rho fixed); the first trial whose misfit is below the current one is
    accepted.  Returns (J0, vs1, vp1, J1, accepted trial) or (J0, None, None,
    None, None) when no trial lowers the misfit."""
    o = Oracle(p, 8)
    J0, _, km, kl = o.gradient(runs, obs)
    km = smooth_array(o, km, sigma)
    kl = smooth_array(o, kl, sigma) if p.psv else np.zeros_like(km)
    gvs, gvp = velocity_gradient(p.model, km, kl, p.psv)
    vs = np.asarray(p.model["vs"], np.float64)
    vp = np.asarray(p.model["vp"], np.float64)
    gs, gp = np.abs(gvs).max(), np.abs(gvp).max()
    for t in range(tries):
        s = step / 2.0 ** t
        a_s = s * vs.max() / gs if gs > 0 else 0.0
        a_p = s * vp.max() / gp if (p.psv and gp > 0) else 0.0
        vs1 = np.maximum(vs - a_s * gvs, 0.0)
        vp1 = np.maximum(vp - a_p * gvp, 0.0) if p.psv else vp
        m1 = dict(p.model, vs=vs1.astype(np.float32), vp=vp1.astype(np.float32))
        p1 = Problem(**{**p.__dict__, "model": m1})
        J1 = Oracle(p1, 8).objective(runs, obs)
        if J1 < J0:
            return J0, vs1, vp1, J1, t
    return J0, None, None, None, None


def max_rel_err(got, ref) -> float:
    """Peak-normalised max error (tests/integration.rs:153-170)."""
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    peak = np.abs(ref).max() if ref.size else 0.0
    denom = peak if peak > np.finfo(np.float32).eps else 1.0
    return float(np.abs(got - ref).max() / denom) if ref.size else 0.0


def rel_l2(got, ref) -> float:
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(got - ref) / (n if n > 0 else 1.0))
