"""Pins the CPU oracle (oracle/fd2d_oracle.cpp) before it is trusted.

* Against the reference's own golden data: the three integration-test
  fixtures (tests/integration.rs:201-282) and the reference's forward traces
  on the checkerboard model (examples/adjoint/obs_traces, every 16th sample).
* Against closed forms that a plausible mistake would break: the 0.75 slope of
  the reference's FD operators on a ramp (fd2d.wgsl:151-180), Hooke's law and
  the micropolar couple-stress update on linear velocity fields, the Gaussian
  taper, brute-force Gaussian smoothing, the Ricker wavelet and the misfit.
"""
import os

import numpy as np
import pytest

import examples as E
from oracle import oracle as O


def load_example(tmp_path, ex, ini_name="config.ini", model_key="model_init"):
    d = E.materialise(ex, str(tmp_path))
    ini = E.ini_values(os.path.join(d, ini_name))
    mdir = ini["paths"].get("model_init") or ini["paths"].get("model_true")
    p = O.load_problem(os.path.join(d, mdir), os.path.join(d, ini["paths"]["sources"]),
                       os.path.join(d, ini["paths"]["stations"]), ini)
    return p, ini


@pytest.mark.parametrize("prec", [4, 8])
def test_forward_example_matches_reference_fixtures(tmp_path, prec):
    p, ini = load_example(tmp_path, "forward")
    assert (p.nx, p.nz, p.dx, p.dz) == (200, 200, 100.0, 100.0)
    o = O.Oracle(p, prec)
    obs, snaps = o.forward(-1, ini["save_snapshot"])
    tr, ref_tr = obs[O.OBS_Y], E.fixture("forward", "uy_000000.npy")
    sn, ref_sn = snaps[0][0], E.fixture("forward", "proc000100_vy.bin")
    print(f"prec {prec}: uy rl2 {O.rel_l2(tr, ref_tr):.2e}  vy(100) rl2 {O.rel_l2(sn, ref_sn):.2e}")
    assert O.max_rel_err(tr, ref_tr) <= 1e-4 and O.max_rel_err(sn, ref_sn) <= 1e-4
    assert O.rel_l2(tr, ref_tr) <= 1e-5 and O.rel_l2(sn, ref_sn) <= 1e-5


@pytest.mark.parametrize("prec", [4, 8])
def test_spin_example_matches_reference_fixtures(tmp_path, prec):
    p, ini = load_example(tmp_path, "spin")
    o = O.Oracle(p, prec)
    obs, snaps = o.forward(-1, ini["save_snapshot"])
    ry, ref_ry = obs[O.OBS_Y], E.fixture("spin", "ry_000000.npy")
    vx, ref_vx = snaps[0][1], E.fixture("spin", "proc000200_vx.bin")
    print(f"prec {prec}: ry rl2 {O.rel_l2(ry, ref_ry):.2e}  vx(200) rl2 {O.rel_l2(vx, ref_vx):.2e}")
    assert O.max_rel_err(ry, ref_ry) <= 1e-4 and O.max_rel_err(vx, ref_vx) <= 1e-4
    assert O.rel_l2(ry, ref_ry) <= 1e-5 and O.rel_l2(vx, ref_vx) <= 1e-5


@pytest.mark.parametrize("src", [0, 5, 11])
def test_checkerboard_forward_matches_reference_traces(tmp_path, src):
    p, ini = load_example(tmp_path, "adjoint", "config_true.ini")
    assert (p.nx, p.nz) == (201, 201) and len(p.src_idx) == 12 and len(p.rec_idx) == 144
    o = O.Oracle(p, 4)
    obs, _ = o.forward(src, 0)
    got = obs[O.OBS_Y][:, ::16]
    ref = E.obs_traces_every16()[f"uy_{src:06d}"]
    assert O.max_rel_err(got, ref) <= 1e-4 and O.rel_l2(got, ref) <= 1e-5


@pytest.mark.slow
def test_adjoint_example_matches_reference_kmu(tmp_path):
    """End to end as run_example.bash does it: forward on model_true for the
    observed traces, then the adjoint workflow on model_init."""
    pt, _ = load_example(tmp_path, "adjoint", "config_true.ini")
    ot = O.Oracle(pt, 4)
    obs = np.stack([ot.forward(s, 0)[0][O.OBS_Y] for s in range(12)])
    p, ini = load_example(tmp_path, "adjoint")
    o = O.Oracle(p, 4)
    kmu, mis = o.adjoint_plain(list(range(12)), obs, ini["adj_interval"], ini["smooth"])
    ref = E.fixture("adjoint", "proc000000_kmu.bin")
    print(f"kmu rl2 {O.rel_l2(kmu, ref):.2e} max {O.max_rel_err(kmu, ref):.2e}")
    assert O.max_rel_err(kmu, ref) <= 1e-4 and O.rel_l2(kmu, ref) <= 1e-5
    assert np.all(mis > 0)


def small_problem(nx=12, nz=10, sh=False, psv=True, spin=False, dx=2.0, dz=3.0, dt=0.5):
    m = {"rho": np.ones(nx * nz, np.float32)}
    if spin:
        m.update({k: np.ones(nx * nz, np.float32) for k in ("lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c")})
    else:
        m.update(vp=np.full(nx * nz, 2.0, np.float32), vs=np.ones(nx * nz, np.float32))
    return O.Problem(nx=nx, nz=nz, nt=4, dx=dx, dz=dz, dt=dt, sh=sh, psv=psv, spin=spin, model=m)


def test_fd_operators_have_reference_weights_on_a_ramp():
    """SH velocity from a stress ramp: the reference's D- returns 0.75 of the
    true slope (weights 9/8, -1/8 of the 3-cell difference) and 0 on the
    two rows/columns next to each edge."""
    p = small_problem(sh=True, psv=False)
    o = O.Oracle(p, 8)
    i, j = np.meshgrid(np.arange(p.nx), np.arange(p.nz), indexing="ij")
    o.set_field(O.F["SXY"], (5.0 * i).ravel())       # slope 5/dx per cell index
    o.set_field(O.F["SZY"], (-7.0 * j).ravel())
    o.step(0, -1)
    vy = o.field(O.F["VY"]).reshape(p.nx, p.nz)
    # vps2lm: rho*vs^2 = 1, taper 1, rho 1 -> vy = dt * (Dx- sxy + Dz- szy)
    ex = np.where((i >= 2) & (i < p.nx - 2), 0.75 * 5.0 / p.dx, 0.0)
    ez = np.where((j >= 2) & (j < p.nz - 2), 0.75 * -7.0 / p.dz, 0.0)
    np.testing.assert_allclose(vy, p.dt * (ex + ez), rtol=1e-12, atol=1e-12)
    # stress update uses the NEW velocity with D+ (zero for i<1, i>=nx-2)
    sxy = o.field(O.F["SXY"]).reshape(p.nx, p.nz)
    dvx = np.zeros_like(vy)
    dvx[1:p.nx - 2] = (9 * (vy[2:p.nx - 1] - vy[1:p.nx - 2]) - (vy[3:] - vy[:p.nx - 3])) / (8 * p.dx)
    np.testing.assert_allclose(sxy, 5.0 * i + p.dt * dvx, rtol=1e-12, atol=1e-10)


def test_hooke_and_couple_stress_closed_forms():
    """One PSV+spin step from linear velocity fields with zero stress: the
    velocities are unchanged (zero divergence) and the stresses follow
    Hooke's law / the couple-stress law with the 0.75 operator slope."""
    nx, nz = 14, 11
    p = small_problem(nx, nz, spin=True)
    lam, mu, nu, muc, nuc = 3.0, 2.0, 0.5, 0.25, 0.125
    p.model.update({"lambda": np.full(nx * nz, lam, np.float32), "mu": np.full(nx * nz, mu, np.float32),
                    "nu": np.full(nx * nz, nu, np.float32), "mu_c": np.full(nx * nz, muc, np.float32),
                    "nu_c": np.full(nx * nz, nuc, np.float32)})
    o = O.Oracle(p, 8)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    a, b, c, d = 0.3, -0.2, 0.7, 0.11
    o.set_field(O.F["VX"], (a * i + b * j).ravel())
    o.set_field(O.F["VZ"], (c * i + d * j).ravel())
    o.step(0, -1)
    dt = p.dt
    inx = (i >= 1) & (i < nx - 2)
    inz = (j >= 1) & (j < nz - 2)
    dvxdx = np.where(inx, 0.75 * a / p.dx, 0)
    dvzdx = np.where(inx, 0.75 * c / p.dx, 0)
    dvxdz = np.where(inz, 0.75 * b / p.dz, 0)
    dvzdz = np.where(inz, 0.75 * d / p.dz, 0)
    get = lambda n: o.field(O.F[n]).reshape(nx, nz)
    np.testing.assert_allclose(get("VX"), a * i + b * j, atol=1e-12)
    np.testing.assert_allclose(get("SXX"), dt * ((lam + 2 * mu) * dvxdx + lam * dvzdz), atol=1e-12)
    np.testing.assert_allclose(get("SZZ"), dt * ((lam + 2 * mu) * dvzdz + lam * dvxdx), atol=1e-12)
    np.testing.assert_allclose(get("SXZ"), dt * mu * (dvxdz + dvzdx), atol=1e-12)
    np.testing.assert_allclose(get("SYY_C"), 2 * dt * nu * (0 - 0.5 * (dvzdx - dvxdz)), atol=1e-12)
    np.testing.assert_allclose(get("SYX_C"), 0, atol=1e-12)


def test_taper_closed_form():
    nx, nz, w, al = 50, 45, 12, 0.05
    p = small_problem(nx, nz, sh=True, psv=False)
    p.abs_left = p.abs_right = p.abs_bottom = p.abs_top = True
    p.abs_width, p.abs_alpha = w, al
    o = O.Oracle(p, 8)
    b = o.field(O.F["BOUND"]).reshape(nx, nz)
    al = float(np.float32(al))

    def f(idx, n):
        out = np.ones(n)
        lo = idx + 1 < w
        out[lo] *= np.exp(-(al * (w - idx[lo] - 1)) ** 2)
        hi = idx > n - w
        out[hi] *= np.exp(-(al * (w + idx[hi] - n)) ** 2)
        return out

    np.testing.assert_allclose(b, np.outer(f(np.arange(nx), nx), f(np.arange(nz), nz)), rtol=1e-12)


def test_smoothing_matches_brute_force_gaussian():
    nx, nz, sig = 9, 7, 1.7
    p = small_problem(nx, nz, sh=True, psv=False)
    o = O.Oracle(p, 8)
    rng = np.random.default_rng(1)
    k = rng.standard_normal((nx, nz))
    o.set_field(O.F["K_MU"], k.ravel())
    o.smooth(sig)
    got = o.field(O.F["K_MU"]).reshape(nx, nz)
    wx = np.exp(-(np.arange(nx)[:, None] - np.arange(nx)[None]) ** 2 / (2 * sig * sig))
    wz = np.exp(-(np.arange(nz)[:, None] - np.arange(nz)[None]) ** 2 / (2 * sig * sig))
    want = (wx @ k @ wz.T) / np.outer(wx.sum(1), wz.sum(1))
    np.testing.assert_allclose(got, want, rtol=1e-12)


def test_ricker_closed_form():
    rows = [[0, 0, 0, 2.0, 0.6, 0.0, 1e10], [0, 0, 0, 25.0, 0.05, 45.0, 3e9], [0, 0, 0, 8.0, 0.15, 90.0, 1e10]]
    nt, dt = 300, 1e-3
    for prec in (4, 8):
        tab = O.ricker_table(prec, rows, nt, dt)
        for s, r in enumerate(rows):
            # the scheme's inputs are f32 values (solver.rs:710-715)
            f0, t0, ang, amp = (float(np.float32(v)) for v in (r[3], r[4], r[5], r[6]))
            t = np.arange(nt) * float(np.float32(dt)) - t0
            a = (np.pi * f0) ** 2
            st = -amp * (1 - 2 * a * t * t) * np.exp(-a * t * t)
            th = np.deg2rad(ang)
            tol = 1e-5 if prec == 4 else 1e-9
            np.testing.assert_allclose(tab[1, s], st, rtol=tol, atol=tol * amp)
            np.testing.assert_allclose(tab[0, s], st * np.sin(th), rtol=tol, atol=tol * amp)
            np.testing.assert_allclose(tab[2, s], -st * np.cos(th), rtol=tol, atol=tol * amp)


def test_misfit_closed_form():
    rng = np.random.default_rng(3)
    nrec, nt, dt = 3, 200, 0.01
    syn, obs = rng.standard_normal((nrec, nt)), rng.standard_normal((nrec, nt))
    m, ad = O.misfit(8, syn, obs, dt)
    t = np.arange(nt) * dt
    te = (nt - 1) * dt
    tw = te / 10
    tap = np.ones(nt)
    lo, hi = t <= tw, t >= te - tw
    tap[lo] = 0.5 + 0.5 * np.cos(np.pi * (tw - t[lo]) / tw)
    tap[hi] = 0.5 + 0.5 * np.cos(np.pi * (te - tw - t[hi]) / tw)
    diff = (syn - obs) * tap
    assert m == pytest.approx(np.sqrt((diff ** 2).sum()), rel=1e-12)
    np.testing.assert_allclose(ad, (2 * diff * tap)[:, ::-1], rtol=1e-12)


def test_grid_inference_rule():
    def grid(nx, nz, h):
        i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
        return (i * h).astype(np.float32).ravel(), (j * h).astype(np.float32).ravel()

    for nx, nz in [(200, 200), (201, 201), (1000, 600), (1536, 1024)]:
        got = O.infer_grid(*grid(nx, nz, 10.0))
        assert got[:2] == (nx, nz)
    # the reference's rounding rule mis-factors a 2:1 grid (solver.rs:653-654)
    assert O.infer_grid(*grid(2000, 1000, 10.0))[:2] == (2001, 1000)


def test_fp32_restatement_error_vs_fp64_oracle(tmp_path):
    """States the FP32 scheme's error against the FP64 oracle per output."""
    p, ini = load_example(tmp_path, "forward")
    r32 = O.Oracle(p, 4).forward(-1, ini["save_snapshot"])
    r64 = O.Oracle(p, 8).forward(-1, ini["save_snapshot"])
    e_tr = O.rel_l2(r32[0][O.OBS_Y], r64[0][O.OBS_Y])
    e_sn = O.rel_l2(r32[1][:, 0], r64[1][:, 0])
    print(f"FP32 vs FP64: uy traces rl2 {e_tr:.2e}, vy snapshots rl2 {e_sn:.2e}")
    assert e_tr < 1e-5 and e_sn < 1e-5


def test_vps2lm_closed_form():
    """vps2lm (fd2d.wgsl:188-201): mu = rho vs^2; lambda = rho (vp^2 - 2 vs^2)
    where vp > vs, else 0 (the branch no reference fixture reaches)."""
    nx, nz = 9, 8
    rng = np.random.default_rng(5)
    p = small_problem(nx, nz)
    rho = rng.uniform(1.0, 3.0, nx * nz).astype(np.float32)
    vs = rng.uniform(0.5, 2.0, nx * nz).astype(np.float32)
    vp = (vs * rng.uniform(0.5, 2.5, nx * nz)).astype(np.float32)
    vp[:5] = vs[:5]                       # vp == vs: the else branch
    p.model.update(rho=rho, vp=vp, vs=vs)
    assert np.any(vp < vs) and np.any(vp > vs)
    for prec in (4, 8):
        o = O.Oracle(p, prec)
        r, s, v = (x.astype(np.float64) for x in (rho, vs, vp))
        want_mu = r * s * s
        want_lam = np.where(v > s, r * (v * v - 2 * s * s), 0.0)
        tol = 1e-6 if prec == 4 else 1e-14
        np.testing.assert_allclose(o.field(O.F["MU"]), want_mu, rtol=tol)
        np.testing.assert_allclose(o.field(O.F["LAM"]), want_lam, rtol=tol, atol=tol * np.abs(want_lam).max())
        assert np.all(o.field(O.F["LAM"])[v <= s] == 0)


def test_spin_traces_closed_forms():
    """Spin receivers (fd2d.wgsl:449-492): ux, uz and ry are the displacement
    and micro-rotation fields at the receiver cells, and yi = ry - curl/2 with
    curl = -(Dx+ uz - Dz+ ux), D+ the forward difference
    (9 (f[i+1] - f[i]) - (f[i+2] - f[i-1])) / 8h, zero on the first and the
    last two rows/columns.  Checked on the last sample of every receiver
    against the fields the run ends with."""
    nx, nz, nt = 24, 20, 30
    rng = np.random.default_rng(8)
    p = small_problem(nx, nz, spin=True, dt=0.05)
    for k in ("rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c"):
        p.model[k] = rng.uniform(0.5, 1.5, nx * nz).astype(np.float32)
    p.nt = nt
    p.src_idx = [10 * nz + 9, 4 * nz + 15]
    cells = [(12, 10), (0, 0), (1, 1), (nx - 2, 3), (5, nz - 2), (20, 17)]
    p.rec_idx = [i * nz + j for i, j in cells]
    stf = rng.standard_normal((3, len(p.src_idx), nt))
    o = O.Oracle(p, 8, stf=stf)
    obs, _ = o.forward(-1, 0)
    ux, uz = o.field(O.F["UX"]).reshape(nx, nz), o.field(O.F["UZ"]).reshape(nx, nz)
    uyc = o.field(O.F["UY_C"]).reshape(nx, nz)
    duzdx, duxdz = np.zeros((nx, nz)), np.zeros((nx, nz))
    duzdx[1:nx - 2] = (9 * (uz[2:nx - 1] - uz[1:nx - 2]) - (uz[3:nx] - uz[0:nx - 3])) / (8 * p.dx)
    duxdz[:, 1:nz - 2] = (9 * (ux[:, 2:nz - 1] - ux[:, 1:nz - 2]) - (ux[:, 3:nz] - ux[:, 0:nz - 3])) / (8 * p.dz)
    curl = -(duzdx - duxdz)
    assert np.abs(curl).max() > 0
    for r, (i, j) in enumerate(cells):
        assert obs[O.OBS_X][r, -1] == ux[i, j] and obs[O.OBS_Z][r, -1] == uz[i, j]
        assert obs[O.OBS_Y][r, -1] == uyc[i, j]
        np.testing.assert_allclose(obs[O.OBS_YI][r, -1], uyc[i, j] - 0.5 * curl[i, j], rtol=1e-12,
                                   atol=1e-12 * np.abs(uyc).max())


@pytest.mark.parametrize("phys", ["sh", "psv"])
def test_taper_multiplies_the_velocities(phys):
    """add_vy / add_vxz (fd2d.wgsl:263-271, 355-368): with zero stresses and
    no source one step leaves v = b v0, b the taper of all four edges; the
    P-SV case is the one no reference fixture has a taper for."""
    nx, nz = 30, 26
    p = small_problem(nx, nz, sh=phys == "sh", psv=phys == "psv")
    p.abs_left = p.abs_right = p.abs_bottom = p.abs_top = True
    p.abs_width, p.abs_alpha = 9, 0.07
    o = O.Oracle(p, 8)
    rng = np.random.default_rng(4)
    names = ["VY"] if phys == "sh" else ["VX", "VZ"]
    v0 = {n: rng.standard_normal(nx * nz) for n in names}
    for n in names:
        o.set_field(O.F[n], v0[n])
    o.step(0, -1)
    b = o.field(O.F["BOUND"])
    assert b.min() < 0.9
    for n in names:
        np.testing.assert_allclose(o.field(O.F[n]), b * v0[n], rtol=1e-15, atol=0)
