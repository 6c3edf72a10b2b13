"""Differentiable gradient (K_rho, K_mu, K_lambda; no counterpart in the
reference, parity with it is unpinned).

Pinned instead by:
* the oracle's FP64 exact discrete adjoint against FP64 central finite
  differences of the same scheme (relative step 1e-4) on tiny grids, at
  sampled cells and along a random direction through every cell at once
  (<grad J, dm> = dJ/de) - this pins the adjoint derivation (a dropped term,
  a wrong sign or a transposed stencil fails it);
* the GPU's FP32 checkpointed adjoint against the oracle's FP64 adjoint, and
  its invariance to the checkpoint interval and to shot batching.
"""
import numpy as np
import pytest

from oracle import oracle as O


def problem(sh, nx, nz, nt, seed=0, nrec=5):
    rng = np.random.default_rng(seed)
    npt = nx * nz
    model = {"rho": rng.uniform(1.0, 2.0, npt).astype(np.float32),
             "vp": rng.uniform(2.0, 2.6, npt).astype(np.float32),
             "vs": rng.uniform(0.8, 1.2, npt).astype(np.float32)}
    src = [(nx // 3) * nz + nz // 2, (2 * nx // 3) * nz + nz // 3, 4 * nz + nz - 3]
    rec = [3 * nz + 3, (nx - 3) * nz + nz - 4, (nx // 2) * nz + 2, 3 * nz + 3, (nx // 2) * nz + nz // 2][:nrec]
    p = O.Problem(nx=nx, nz=nz, nt=nt, dx=float(np.float32(1.0)), dz=float(np.float32(1.25)),
                  dt=float(np.float32(0.1)), sh=sh, psv=not sh, abs_left=True, abs_right=True, abs_bottom=True,
                  abs_top=True, abs_width=5, abs_alpha=float(np.float32(0.2)), model=model, src_idx=src,
                  rec_idx=rec)
    stf = rng.standard_normal((3, len(src), nt)).astype(np.float32)
    ncomp = 1 if sh else 2
    return p, stf, ncomp


@pytest.mark.parametrize("sh", [True, False], ids=["sh", "psv"])
def test_oracle_adjoint_matches_fp64_finite_differences(sh):
    p, stf, nc = problem(sh, 18, 15, 60)
    o = O.Oracle(p, 8, stf=stf.astype(np.float64))
    runs = [0, -1]
    obs = np.random.default_rng(1).standard_normal((len(runs), nc, len(p.rec_idx), p.nt)) * 0.01
    J, kr, km, kl = o.gradient(runs, obs)
    assert J == pytest.approx(o.objective(runs, obs), rel=1e-14)
    rng = np.random.default_rng(5)
    params = [("RHO", kr), ("MU", km)] + ([] if sh else [("LAM", kl)])
    for name, K in params:
        base = o.field(O.F[name])
        # points where the kernel is not negligible (inside the domain of dependence)
        cand = np.flatnonzero(np.abs(K) > 1e-3 * np.abs(K).max())
        fd, ad = [], []
        for k in rng.choice(cand, 6, replace=False):
            h = 1e-4 * abs(base[k])
            f = base.copy()
            f[k] += h
            o.set_field(O.F[name], f)
            jp = o.objective(runs, obs)
            f[k] -= 2 * h
            o.set_field(O.F[name], f)
            jm = o.objective(runs, obs)
            o.set_field(O.F[name], base)
            fd.append((jp - jm) / (2 * h))
            ad.append(K[k])
        fd, ad = np.array(fd), np.array(ad)
        assert np.abs(fd - ad).max() <= 1e-6 * np.abs(ad).max(), (name, fd, ad)
    if sh:
        assert np.all(kl == 0)


def gpu_problem(seiswg, p, stf):
    c = seiswg.Context.from_problem(p)
    c.set_sources(p.src_idx, stf)
    c.set_receivers(p.rec_idx)
    return c


@pytest.mark.parametrize("sh", [True, False], ids=["sh", "psv"])
def test_oracle_gradient_equals_directional_derivative(sh):
    """<grad J, dm> against the central difference (J(m + e dm) - J(m - e dm))
    / 2e for a random perturbation dm of every cell of rho, mu (and lambda)
    at once: the dot-product form of the linearisation test, pinning the
    whole kernel fields rather than sampled points."""
    p, stf, nc = problem(sh, 16, 14, 50, seed=7)
    o = O.Oracle(p, 8, stf=stf.astype(np.float64))
    runs = [1, -1]
    obs = np.random.default_rng(3).standard_normal((len(runs), nc, len(p.rec_idx), p.nt)) * 0.01
    J, kr, km, kl = o.gradient(runs, obs)
    rng = np.random.default_rng(11)
    names = ["RHO", "MU"] + ([] if sh else ["LAM"])
    grads = {"RHO": kr, "MU": km, "LAM": kl}
    base = {n: o.field(O.F[n]).copy() for n in names}
    dm = {n: rng.standard_normal(base[n].size) * np.abs(base[n]) for n in names}
    ad = sum(float(np.dot(grads[n], dm[n])) for n in names)
    eps = 1e-5

    def j_at(sign):
        for n in names:
            o.set_field(O.F[n], base[n] + sign * eps * dm[n])
        return o.objective(runs, obs)

    fd = (j_at(1) - j_at(-1)) / (2 * eps)
    for n in names:
        o.set_field(O.F[n], base[n])
    print(f"directional derivative: adjoint {ad:.12e} finite difference {fd:.12e}")
    assert abs(fd - ad) <= 1e-7 * abs(ad)


@pytest.mark.gpu
@pytest.mark.parametrize("sh", [True, False], ids=["sh", "psv"])
@pytest.mark.parametrize("threads", [64, 256])
def test_gpu_gradient_matches_oracle_adjoint(seiswg, sh, threads):
    p, stf, nc = problem(sh, 46, 40, 120, seed=3)
    runs = [0, 1, -1]
    rng = np.random.default_rng(2)
    obs = (rng.standard_normal((len(runs), nc, len(p.rec_idx), p.nt)) * 1e-3).astype(np.float32)
    print(f"case sh={sh} threads={threads}", flush=True)
    c = gpu_problem(seiswg, p, stf)
    c.set_tiling(threads, 9)
    mis, kr, km, kl = c.gradient(runs, obs, ckpt_every=16)
    o = O.Oracle(p, 8, stf=stf.astype(np.float64))
    J, okr, okm, okl = o.gradient(runs, obs.astype(np.float64))
    per_run = [o.objective([r], obs[q:q + 1].astype(np.float64)) for q, r in enumerate(runs)]
    np.testing.assert_allclose(mis, per_run, rtol=1e-5)
    for name, g, w in (("K_rho", kr, okr), ("K_mu", km, okm)) + ((("K_lam", kl, okl),) if not sh else ()):
        err = O.rel_l2(g, w)
        print(f"{name}: rel-L2 vs FP64 oracle adjoint {err:.2e}")
        assert err <= 1e-4, name
    if sh:
        assert not np.any(kl)


@pytest.mark.gpu
@pytest.mark.parametrize("sh", [True, False], ids=["sh", "psv"])
def test_gpu_gradient_independent_of_checkpoint_interval_and_batching(seiswg, sh):
    """One shot per workgroup: misfits and kernels bit for bit whatever the
    checkpoint interval and the shots per launch.  Two and four shots per
    workgroup (their kernel contributions summed per step in the workgroup,
    then over time): the same kernels up to FP32 rounding, and bit for bit
    across checkpoint intervals."""
    p, stf, nc = problem(sh, 50, 44, 90, seed=4)
    runs = [0, 1, 2, -1]
    obs = (np.random.default_rng(7).standard_normal((len(runs), nc, len(p.rec_idx), p.nt)) * 1e-3).astype(np.float32)
    c = gpu_problem(seiswg, p, stf)
    c.set_tiling(128, 9, 1)
    ref = c.gradient(runs, obs, ckpt_every=90)
    for ck, spl in ((7, 0), (25, 1), (11, 3)):
        got = c.gradient(runs, obs, ckpt_every=ck, shots_per_launch=spl)
        assert np.array_equal(got[0], ref[0])
        for a, b in zip(got[1:], ref[1:]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (ck, spl)
    for group in (2, 4):
        c.set_tiling(128, 9, group)
        g90 = c.gradient(runs, obs, ckpt_every=90)
        assert c.tiling()[2] == group
        g7 = c.gradient(runs, obs, ckpt_every=7)
        assert np.array_equal(g90[0], ref[0]) and np.array_equal(g7[0], ref[0])
        for a, b, r in zip(g90[1:], g7[1:], ref[1:]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), group
            if np.any(r):
                err = O.rel_l2(a, r)
                print(f"group {group}: rel-L2 vs one shot per workgroup {err:.2e}")
                assert err <= 1e-6, group
            else:
                assert not np.any(a)


@pytest.mark.gpu
@pytest.mark.parametrize("ck", [7, 25, 90])
def test_gpu_gradient_two_step_launches_equal_one_step_launches(seiswg, ck):
    """P-SV gradient with the forward and the aux-field recompute in two-step
    launches (the default) against one step per launch: misfits and kernels
    bit for bit, for segments of 7 (one pair and three single steps), 25 and
    90 steps (captured graphs)."""
    p, stf, nc = problem(False, 50, 44, 90, seed=6)
    runs = [0, 1, 2, -1]
    obs = (np.random.default_rng(9).standard_normal((len(runs), nc, len(p.rec_idx), p.nt)) * 1e-3).astype(np.float32)
    out = []
    for spl in (1, 2):
        c = gpu_problem(seiswg, p, stf)
        c.set_steps_per_launch(spl)
        out.append(c.gradient(runs, obs, ckpt_every=ck))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1:], out[1][1:]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), ck


def multi_problems(seiswg):
    """Uneven problems: SH and P-SV, different grids, nt, dt and tile shapes."""
    specs = [(True, 40, 36, 70, 0.1, (64, 7, 1)), (False, 52, 30, 90, 0.08, (128, 16, 1)),
             (False, 34, 70, 55, 0.1, (64, 32, 1)), (True, 60, 44, 90, 0.05, (128, 9, 1))]
    out = []
    for q, (sh, nx, nz, nt, dt, tl) in enumerate(specs):
        p, stf, nc = problem(sh, nx, nz, nt, seed=10 + q)
        p.dt = float(np.float32(dt))
        c = gpu_problem(seiswg, p, stf)
        c.set_tiling(*tl)
        out.append((p, stf, nc, c))
    return out


@pytest.mark.gpu
def test_multi_problem_forward_equals_single_runs(seiswg):
    probs = multi_problems(seiswg)
    runs = [[0, -1], [1], [0, 1, 2], [-1, 2]]
    single = []
    for (p, stf, nc, c), r in zip(probs, runs):
        c.forward_batch(r)
        comp = 1 if p.sh else 0
        single.append([c.obs(s, comp).copy() for s in range(len(r))])
    m = seiswg.Multi([c for *_, c in probs])
    m.forward(runs)
    for (p, stf, nc, c), r, ref in zip(probs, runs, single):
        comp = 1 if p.sh else 0
        for s in range(len(r)):
            assert np.array_equal(c.obs(s, comp), ref[s])


@pytest.mark.gpu
@pytest.mark.parametrize("merge_limit", [None, 1 << 40, 3000])
def test_multi_problem_gradient_equals_single_gradients(seiswg, merge_limit):
    """Every problem on its own launches (the default), all four in merged
    launches (a limit above every problem), and a split (limit 3000
    shot-cells: two problems merged, two on their own) give the
    single-problem gradients (one shot per workgroup in both, as merged
    launches run)."""
    probs = multi_problems(seiswg)
    runs = [[0, -1], [1, 2], [0], [2, 0, 1]]
    refs = [(np.random.default_rng(q).standard_normal((len(r), nc, len(p.rec_idx), p.nt)) * 1e-3).astype(np.float32)
            for q, ((p, stf, nc, c), r) in enumerate(zip(probs, runs))]
    single = [c.gradient(r, o, ckpt_every=20) for (p, stf, nc, c), r, o in zip(probs, runs, refs)]
    m = seiswg.Multi([c for *_, c in probs])
    if merge_limit is not None:
        m.set_merge_limit(merge_limit)
    got = m.gradient(runs, refs, ckpt_every=20)
    for a, b in zip(got, single):
        assert np.array_equal(a[0], b[0])
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("workers", [1, 2, 16])
def test_multi_problem_gradient_independent_of_workers(seiswg, workers):
    """Problems on their own launches (merge limit 0) run `workers` at a time
    from host threads; the gradients equal the single-problem ones bit for
    bit whatever the count."""
    probs = multi_problems(seiswg)
    runs = [[0, -1], [1, 2], [0], [2, 0, 1]]
    refs = [(np.random.default_rng(q).standard_normal((len(r), nc, len(p.rec_idx), p.nt)) * 1e-3).astype(np.float32)
            for q, ((p, stf, nc, c), r) in enumerate(zip(probs, runs))]
    single = [c.gradient(r, o, ckpt_every=20) for (p, stf, nc, c), r, o in zip(probs, runs, refs)]
    m = seiswg.Multi([c for *_, c in probs])
    m.set_merge_limit(0)
    m.set_workers(workers)
    got = m.gradient(runs, refs, ckpt_every=20)
    for a, b in zip(got, single):
        assert np.array_equal(a[0], b[0])
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.gpu
def test_released_memory_cache_keeps_results(seiswg):
    """Buffers of a closed context go to the library's memory cache;
    sw_release_cached_memory hands them back, and a new context after that
    (fresh allocations) gives the same gradient bit for bit."""
    import gc
    gc.collect()        # contexts of earlier tests release their buffers now, not mid-test
    p, stf, nc = problem(False, 50, 44, 60, seed=12)
    runs = [0, 1]
    obs = (np.random.default_rng(3).standard_normal((len(runs), nc, len(p.rec_idx), p.nt)) * 1e-3).astype(np.float32)
    c = gpu_problem(seiswg, p, stf)
    ref = c.gradient(runs, obs, ckpt_every=16)
    c.close()
    assert seiswg.release_cached_memory() > 0
    assert seiswg.release_cached_memory() == 0
    c = gpu_problem(seiswg, p, stf)
    got = c.gradient(runs, obs, ckpt_every=16)
    c.close()
    for a, b in zip(got, ref):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.slow
def test_oracle_adjoint_matches_finite_differences_at_the_1000x600_workload():
    """The checkpointed P-SV gradient workload's grid, model and geometry
    (1000x600, dx = dz = 10 m, the layered-plus-random-field model, 200
    receivers at 10 m, shot 8 of the 16 at 50 m, dt = 1 ms) with nt cut to
    200 (the Ricker of f0 = 8 Hz, t0 = 0.15 s has fired) and zero observed
    data: at 20 seeded points (seed 33) per parameter, drawn among the cells
    the wavefield reaches (|K| > 1e-3 max|K|), FP64 central finite
    differences of J with the relative step 1e-4 against the FP64 adjoint
    kernels, rel-L2 <= 1e-3 over the 20 points of each of K_rho, K_mu and
    K_lambda."""
    from seiswg_amd import synth
    c = synth.PSV_GRADIENT
    init, _, geom = synth.gradient_problem()
    h, nz, nt = np.float32(c["h"]), c["nz"], 200
    rows = geom.src_rows()
    src_idx = [O.grid_index(r[0], r[1], h, h, nz) for r in rows]
    rec_idx = [O.grid_index(x, z, h, h, nz) for x, z in zip(geom.rec_x, geom.rec_z)]
    p = O.Problem(nx=c["nx"], nz=nz, nt=nt, dx=c["h"], dz=c["h"], dt=c["dt"], sh=False, psv=True, spin=False,
                  abs_left=True, abs_right=True, abs_bottom=True, abs_top=False, abs_width=c["abs_width"],
                  abs_alpha=0.015, model=init, src_rows=rows, src_idx=src_idx, rec_idx=rec_idx)
    o = O.Oracle(p, 8)
    runs = [8]
    obs = np.zeros((1, 2, len(rec_idx), nt))
    J, kr, km, kl = o.gradient(runs, obs)
    assert J > 0
    rng = np.random.default_rng(33)
    for name, K in (("RHO", kr), ("MU", km), ("LAM", kl)):
        base = o.field(O.F[name]).copy()
        cand = np.flatnonzero(np.abs(K) > 1e-3 * np.abs(K).max())
        pts = rng.choice(cand, 20, replace=False)
        fd = []
        for k in pts:
            step = 1e-4 * abs(base[k])
            f = base.copy()
            f[k] = base[k] + step
            o.set_field(O.F[name], f)
            jp = o.objective(runs, obs)
            f[k] = base[k] - step
            o.set_field(O.F[name], f)
            jm = o.objective(runs, obs)
            fd.append((jp - jm) / (2 * step))
        o.set_field(O.F[name], base)
        fd, ad = np.array(fd), K[pts]
        err = O.rel_l2(ad, fd)
        print(f"K_{name}: rel-L2 of the adjoint against FD at 20 points {err:.2e} (|K| {np.abs(ad).min():.2e}.."
              f"{np.abs(ad).max():.2e})")
        assert err <= 1e-3, name


@pytest.mark.gpu
def test_gpu_gradient_without_receivers_has_zero_misfits(seiswg):
    """No receivers: no traces, no adjoint source; every shot's misfit and
    the kernels are exactly zero."""
    p, stf, nc = problem(False, 40, 36, 40, seed=2)
    c = gpu_problem(seiswg, p, stf)
    runs = [0, 1]
    obs = np.ones((len(runs), nc, 5, p.nt), np.float32)
    c.gradient(runs, obs, ckpt_every=16)       # the misfit buffer now holds nonzero values
    c.set_receivers([])
    mis, kr, km, kl = c.gradient(runs, np.zeros((len(runs), nc, 0, p.nt), np.float32), ckpt_every=16)
    assert np.array_equal(mis, np.zeros(len(runs)))
    assert not np.any(kr) and not np.any(km) and not np.any(kl)
