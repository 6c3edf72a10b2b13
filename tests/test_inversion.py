"""mode = inversion (the reference lists it as planned, README.md:61, with no
specification; the rules are this build's, restated in FP64 by
oracle.steepest_descent_iteration).

Pinned by:
* the chain rule from (K_mu, K_lambda) to (dJ/dvs, dJ/dvp) through vps2lm
  against FP64 central finite differences of the oracle's misfit with respect
  to vs and vp themselves, including cells on the vp <= vs branch (a wrong
  factor, sign or branch fails it); the K fields are pinned by
  tests/test_gradient.py, the smoothing by tests/test_oracle.py;
* the CLI's first iteration (GPU gradient, FP32) against the FP64
  restatement: initial misfit, accepted trial, the model update and the
  misfit after it.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O


def tiny_problem(sh, nx=18, nz=15, nt=60, seed=4):
    rng = np.random.default_rng(seed)
    npt = nx * nz
    model = {"rho": rng.uniform(1.0, 2.0, npt).astype(np.float32),
             "vp": rng.uniform(2.0, 2.6, npt).astype(np.float32),
             "vs": rng.uniform(0.8, 1.2, npt).astype(np.float32)}
    model["vp"][5 * nz + 3: 5 * nz + 9] = 0.5      # lambda = 0 there (vps2lm's vp <= vs branch)
    src = [(nx // 3) * nz + nz // 2, (2 * nx // 3) * nz + nz // 3]
    rec = [3 * nz + 3, (nx - 3) * nz + nz - 4, (nx // 2) * nz + 2, 5 * nz + 5]
    p = O.Problem(nx=nx, nz=nz, nt=nt, dx=1.0, dz=1.25, dt=float(np.float32(0.1)), sh=sh, psv=not sh,
                  abs_left=True, abs_right=True, abs_bottom=True, abs_top=True, abs_width=5,
                  abs_alpha=float(np.float32(0.2)), model=model, src_idx=src, rec_idx=rec)
    stf = rng.standard_normal((3, len(src), nt))
    return p, stf


@pytest.mark.parametrize("sh", [True, False], ids=["sh", "psv"])
def test_velocity_chain_rule_matches_finite_differences(sh):
    p, stf = tiny_problem(sh)
    nc = 1 if sh else 2
    runs = [0, 1]
    obs = np.random.default_rng(2).standard_normal((len(runs), nc, len(p.rec_idx), p.nt)) * 0.01
    J, _, km, kl = O.Oracle(p, 8, stf=stf).gradient(runs, obs)
    gvs, gvp = O.velocity_gradient(p.model, km, kl, p.psv)

    def j_with(name, k, h):
        m = {n: v.astype(np.float64) for n, v in p.model.items()}
        m[name][k] += h
        q = O.Problem(**{**p.__dict__, "model": m})
        return O.Oracle(q, 8, stf=stf).objective(runs, obs)

    nz = p.nz
    branch = [5 * nz + 4, 5 * nz + 6]                       # vp < vs: lambda = 0
    rng = np.random.default_rng(9)
    cand = np.flatnonzero(np.abs(gvs) > 1e-3 * np.abs(gvs).max())
    cells = list(rng.choice(cand, 5, replace=False)) + branch
    for name, g in [("vs", gvs)] + ([] if sh else [("vp", gvp)]):
        fd, ad = [], []
        for k in cells:
            h = 1e-4 * float(p.model[name][k])
            fd.append((j_with(name, k, h) - j_with(name, k, -h)) / (2 * h))
            ad.append(g[k])
        fd, ad = np.array(fd), np.array(ad)
        print(name, fd, ad)
        assert np.abs(fd - ad).max() <= 1e-5 * np.abs(ad).max(), (name, fd, ad)
    if not sh:
        assert np.all(gvp[branch] == 0.0)


def write_model_dir(d, nx, nz, h, model):
    os.makedirs(d, exist_ok=True)
    npt = nx * nz
    x = np.repeat(np.arange(nx, dtype=np.float32) * h, nz)
    z = np.tile(np.arange(nz, dtype=np.float32) * h, nx)
    for name, v in list(model.items()) + [("x", x), ("z", z)]:
        with open(os.path.join(d, f"proc000000_{name}.bin"), "wb") as f:
            np.array([npt], "<i4").tofile(f)
            np.asarray(v, "<f4").tofile(f)


@pytest.mark.gpu
def test_cli_inversion_iteration_matches_fp64_restatement(seiswg, tmp_path):
    """A P-SV inversion on a 64 x 48 grid (three sources, eight receivers,
    observed traces from a model with a slow anomaly): one iteration of the
    CLI (batched GPU gradient, smoothing, chain rule, line search) against
    oracle.steepest_descent_iteration in FP64."""
    from test_gpu_parity import run_cli
    nx, nz, h, nt, dt = 64, 48, 10.0, 240, 0.002
    rng = np.random.default_rng(21)
    npt = nx * nz
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    vs = (1000.0 + 8.0 * jj + rng.uniform(-5, 5, (nx, nz))).astype(np.float32).ravel()
    vp = (1.8 * vs).astype(np.float32)
    rho = np.full(npt, 2000.0, np.float32)
    blob = np.exp(-(((ii - 30) ** 2 + (jj - 26) ** 2) / 40.0)).ravel()
    vs_true = (vs * (1.0 - 0.06 * blob)).astype(np.float32)
    d = str(tmp_path)
    write_model_dir(os.path.join(d, "model_init"), nx, nz, h, {"rho": rho, "vp": vp, "vs": vs})
    write_model_dir(os.path.join(d, "model_true"), nx, nz, h, {"rho": rho, "vp": vp, "vs": vs_true})
    srcs = [(120.0, 60.0), (320.0, 60.0), (520.0, 60.0)]
    with open(os.path.join(d, "sources.dat"), "w") as f:
        f.write("# x z type f0 t0 ang amp\n")
        for x, z in srcs:
            f.write(f"{x} {z} 0 25.0 0.05 30.0 1.0e6\n")
    recs = [(60.0 + 70.0 * q, 400.0) for q in range(8)]
    with open(os.path.join(d, "stations.dat"), "w") as f:
        f.write("# x z\n")
        for x, z in recs:
            f.write(f"{x} {z}\n")
    solver = (f"[solver]\nsh = no\npsv = yes\nnt = {nt}\ndt = {dt}\nabs_top = no\nabs_bottom = yes\n"
              "abs_left = yes\nabs_right = yes\nabs_width = 12\nabs_alpha = 0.05\nsmooth = 2.0\n"
              "combine_sources = no\n")
    open(os.path.join(d, "true.ini"), "w").write(
        "[workflow]\nmode = forward\n\n" + solver +
        "\n[path]\noutput = obs\nmodel_init = model_true\nsources = sources.dat\nstations = stations.dat\n")
    run_cli(os.path.join(d, "true.ini"))
    ini = ("[workflow]\nmode = inversion\n\n" + solver +
           "\n[path]\noutput = output\nmodel_init = model_init\nsources = sources.dat\nstations = stations.dat\n"
           "traces = obs/traces\n\n[inversion]\nniter = 1\nstep = 0.02\nlinesearch = 4\ncheckpoint = 40\n")
    open(os.path.join(d, "inv.ini"), "w").write(ini)
    run_cli(os.path.join(d, "inv.ini"))

    # the oracle, on the same files
    vals = dict(nt=nt, dt=float(np.float32(dt)), sh=False, psv=True, spin=False, abs_left=True, abs_right=True,
                abs_bottom=True, abs_top=False, abs_width=12, abs_alpha=float(np.float32(0.05)))
    p = O.load_problem(os.path.join(d, "model_init"), os.path.join(d, "sources.dat"),
                       os.path.join(d, "stations.dat"), vals)
    assert (p.nx, p.nz) == (nx, nz)
    runs = [0, 1, 2]
    obs = np.stack([np.stack([E_npy(os.path.join(d, "obs", "traces", f"{c}_{r:06d}.npy")).reshape(len(recs), nt)
                              for c in ("ux", "uz")]) for r in runs]).astype(np.float64)
    J0, vs1, vp1, J1, trial = O.steepest_descent_iteration(p, runs, obs, 0.02, 4, float(np.float32(2.0)))
    assert trial is not None and J1 < 0.99 * J0, "the test problem should accept a step clearly"
    log = np.loadtxt(os.path.join(d, "output", "misfit.txt"))
    print(f"misfit J0 cli {log[0, 1]:.9e} oracle {J0:.9e}; J1 cli {log[1, 1]:.9e} oracle {J1:.9e}; trial {trial}")
    # measured: 1e-7 and 3e-7
    assert log[0, 1] == pytest.approx(J0, rel=1e-5)
    assert log[1, 1] == pytest.approx(J1, rel=1e-5)
    out = os.path.join(d, "output", "iter_001")
    got_vs = O.read_model_field(os.path.join(out, "proc000000_vs.bin")).astype(np.float64)
    got_vp = O.read_model_field(os.path.join(out, "proc000000_vp.bin")).astype(np.float64)
    for name, got, want, base in (("vs", got_vs, vs1, vs), ("vp", got_vp, vp1, vp)):
        err = O.rel_l2(got - base, want - base)
        print(f"{name} update: rel-L2 {err:.3e} (max |update| {np.abs(want - base).max():.4g})")
        assert err <= 1e-4, name      # measured: vs 7.7e-6, vp 1.2e-5
    assert np.array_equal(O.read_model_field(os.path.join(out, "proc000000_rho.bin")), rho)


def E_npy(path):
    from examples import read_npy_payload
    return read_npy_payload(path)
