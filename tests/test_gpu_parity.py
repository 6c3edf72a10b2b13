"""GPU parity: the fused gfx950 kernels against the CPU oracle (same inputs,
same FP32 operation order) and against the reference's golden fixtures.

Bar: state fields and traces equal to the FP32 oracle within 1 ulp-level
relative error (rel-L2 <= 1e-6 per field; typically bit-identical), and
reference fixtures within the reference's own tolerance (max-rel <= 1e-4,
tests/integration.rs:40) and rel-L2 <= 1e-5.
"""
import os
import subprocess

import numpy as np
import pytest

import examples as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STATE = {"sh": ["VY", "SXY", "SZY"], "psv": ["VX", "VZ", "SXX", "SZZ", "SXZ"],
         "spin": ["VY_C", "SYX_C", "SYY_C", "SYZ_C"]}


def random_problem(nx, nz, nt, sh, psv, spin, seed, nsrc=4, nrec=7):
    rng = np.random.default_rng(seed)
    npt = nx * nz
    if spin:
        names = ["rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c"]
        model = {n: rng.uniform(0.5, 1.5, npt).astype(np.float32) for n in names}
    else:
        model = {"rho": rng.uniform(1.0, 2.0, npt).astype(np.float32),
                 "vp": rng.uniform(1.5, 2.5, npt).astype(np.float32),
                 "vs": rng.uniform(0.5, 1.2, npt).astype(np.float32)}
        model["vp"][:7] = 0.3      # vp <= vs branch of vps2lm
    # sources: interior, near edges and a duplicated cell (index order kept)
    cells = [(nx // 2, nz // 2), (1, 2), (nx - 2, nz - 1), (nx // 2, nz // 2), (3, nz - 3)][:nsrc]
    src_idx = [i * nz + j for i, j in cells]
    rcells = [(nx // 3, nz // 4), (0, 0), (nx - 1, nz - 2), (2, 1), (nx // 3, nz // 4), (nx - 3, 5), (7, nz - 3)]
    rec_idx = [i * nz + j for i, j in rcells[:nrec]]
    p = O.Problem(nx=nx, nz=nz, nt=nt, dx=float(np.float32(1.7)), dz=float(np.float32(2.3)),
                  dt=float(np.float32(0.05)), sh=sh, psv=psv, spin=spin, abs_left=True, abs_right=True,
                  abs_bottom=True, abs_top=True, abs_width=6, abs_alpha=float(np.float32(0.1)), model=model,
                  src_rows=[], src_idx=src_idx, rec_idx=rec_idx)
    stf = rng.standard_normal((3, len(src_idx), nt)).astype(np.float32)
    return p, stf


def gpu_context(p, stf, seiswg):
    c = seiswg.Context.from_problem(p)
    c.set_sources(p.src_idx, stf)
    c.set_receivers(p.rec_idx)
    return c


def compare(got, want, label, tol=1e-6):
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float64).astype(np.float32)
    same = np.array_equal(got.view(np.uint32), want.view(np.uint32))
    err = O.rel_l2(got, want)
    assert err <= tol, f"{label}: rel-L2 {err:.3e}"
    return same


@pytest.mark.parametrize("phys", ["sh", "psv", "spin", "all"])
@pytest.mark.parametrize("tiling", [(64, 7), (128, 1), (256, 64), (64, 200)])
def test_steps_from_random_state_match_oracle(seiswg, phys, tiling):
    sh = phys in ("sh", "all")
    psv = phys in ("psv", "spin", "all")
    spin = phys in ("spin", "all")
    nx, nz, nsteps = 45, 131, 4
    p, stf = random_problem(nx, nz, 6, sh, psv, spin, seed=11)
    print(f"case phys={phys} tiling={tiling}", flush=True)
    c = gpu_context(p, stf, seiswg)
    c.set_tiling(*tiling)
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    o.clear()
    rng = np.random.default_rng(5)
    fields = (STATE["sh"] if sh else []) + (STATE["psv"] if psv else []) + (STATE["spin"] if spin else [])
    c.prepare([-1])
    for name in fields:
        v = rng.standard_normal(nx * nz).astype(np.float32)
        c.write_field(getattr(seiswg.abi, "F_" + name), v)
        o.set_field(O.F[name], v)
    c.advance(nsteps)
    for it in range(nsteps):
        o.step(it, -1)
    exact = []
    for name in fields:
        exact.append(compare(c.field(getattr(seiswg.abi, "F_" + name)), o.field(O.F[name]), name))
    obs = o.obs()
    comps = ([O.OBS_Y] if sh or spin else []) + ([O.OBS_X, O.OBS_Z] if psv else []) + \
            ([O.OBS_YI, O.OBS_YC] if spin else [])
    for comp in comps:
        exact.append(compare(c.obs(0, comp)[:, :nsteps], obs[comp][:, :nsteps], f"obs{comp}"))
    print(f"bit-identical: {sum(exact)}/{len(exact)}")
    assert all(exact)


@pytest.mark.parametrize("spacing", [(1.7, 2.3), (5.02, 5.6)], ids=["admitted", "rejected"])
@pytest.mark.parametrize("nsteps", [3, 6])
@pytest.mark.parametrize("phys", ["sh", "psv", "spin"])
def test_steps_through_subnormal_range_match_oracle_bitwise(seiswg, phys, nsteps, spacing):
    """State fields spanning zero (both signs), the subnormal range and
    normal magnitudes: numerators outside the fast forms' ranges take the
    IEEE division in the kernels, the others the fast forms; both must equal
    the oracle's IEEE divisions bit for bit.  Spacings (5.02, 5.6) give
    divisors 8h that the host's check rejects for the two-operation form
    (the kernels then divide by them the IEEE way); six steps include a
    two-step launch."""
    sh, psv, spin = phys == "sh", phys in ("psv", "spin"), phys == "spin"
    nx, nz = 40, 75
    p, stf = random_problem(nx, nz, 8, sh, psv, spin, seed=21)
    p.dx, p.dz = float(np.float32(spacing[0])), float(np.float32(spacing[1]))
    stf = (stf * 1e-40).astype(np.float32)
    print(f"case phys={phys} nsteps={nsteps} spacing={spacing}", flush=True)
    c = gpu_context(p, stf, seiswg)
    c.set_tiling(64, 9)
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    o.clear()
    rng = np.random.default_rng(8)
    fields = (STATE["sh"] if sh else []) + (STATE["psv"] if psv else []) + (STATE["spin"] if spin else [])
    c.prepare([-1])
    for name in fields:
        mag = 10.0 ** rng.uniform(-45, 2, nx * nz)
        v = (rng.standard_normal(nx * nz) * mag * (rng.uniform(size=nx * nz) > 0.3)).astype(np.float32)
        c.write_field(getattr(seiswg.abi, "F_" + name), v)
        o.set_field(O.F[name], v)
    c.advance(nsteps)
    for it in range(nsteps):
        o.step(it, -1)
    for name in fields:
        got = c.field(getattr(seiswg.abi, "F_" + name))
        want = np.asarray(o.field(O.F[name]), np.float64).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name


@pytest.mark.parametrize("threads,group", [(64, 1), (128, 2), (256, 4), (256, 2)])
def test_single_source_runs_and_batched_shots_agree(seiswg, threads, group):
    """Shots batched in one launch, several per workgroup, equal single runs."""
    nx, nz, nt = 60, 70, 40
    p, stf = random_problem(nx, nz, nt, False, True, False, seed=3)
    print(f"case threads={threads} group={group}", flush=True)
    c = gpu_context(p, stf, seiswg)
    c.set_tiling(threads, 16, group)
    singles = []
    for s in range(4):
        c.forward(s)
        singles.append((c.obs(0, O.OBS_X).copy(), c.field(seiswg.abi.F_VZ).copy()))
    c.forward_batch([0, 1, 2, 3])
    assert c.tiling() == (threads, 16, group)
    for s in range(4):
        assert np.array_equal(c.obs(s, O.OBS_X), singles[s][0])
        assert np.array_equal(c.field(seiswg.abi.F_VZ, shot=s), singles[s][1])
    # per-shot oracle check for one shot
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    obs, _ = o.forward(2, 0)
    compare(c.obs(2, O.OBS_X), obs[O.OBS_X], "batched shot 2 ux")


@pytest.mark.parametrize("phys", ["sh", "spin", "all"])
@pytest.mark.parametrize("threads,group", [(128, 2), (256, 2), (128, 4)])
def test_grouped_shots_equal_single_runs(seiswg, phys, threads, group):
    """Every physics with several shots per workgroup against one shot per run."""
    sh, psv, spin = phys in ("sh", "all"), phys in ("spin", "all"), phys in ("spin", "all")
    nx, nz, nt = 50, 140, 30
    p, stf = random_problem(nx, nz, nt, sh, psv, spin, seed=17)
    print(f"case phys={phys} threads={threads} group={group}", flush=True)
    c = gpu_context(p, stf, seiswg)
    c.set_tiling(threads, 11, group)
    comps = ([O.OBS_Y] if sh or spin else []) + ([O.OBS_X, O.OBS_Z] if psv else []) + \
            ([O.OBS_YI, O.OBS_YC] if spin else [])
    fields = (STATE["sh"] if sh else []) + (STATE["psv"] if psv else []) + (STATE["spin"] if spin else [])
    runs = [0, 1, 2, 3]
    single = []
    for r in runs:
        c.forward(r)
        single.append(([c.obs(0, k).copy() for k in comps],
                       [c.field(getattr(seiswg.abi, "F_" + n)).copy() for n in fields]))
    c.forward_batch(runs)
    eff = c.tiling()[2]
    assert eff == (2 if spin and threads * group > 512 else group)
    for q in range(len(runs)):
        for k, ref in zip(comps, single[q][0]):
            assert np.array_equal(c.obs(q, k).view(np.uint32), ref.view(np.uint32)), (q, k)
        for n, ref in zip(fields, single[q][1]):
            assert np.array_equal(c.field(getattr(seiswg.abi, "F_" + n), shot=q).view(np.uint32),
                                  ref.view(np.uint32)), (q, n)


@pytest.mark.parametrize("threads,seglen,group", [(64, 5, 1), (128, 16, 2), (256, 9, 4), (128, 3, 4), (64, 200, 1)])
@pytest.mark.parametrize("nt", [11, 37])
@pytest.mark.parametrize("phys", ["psv", "sh"])
def test_two_step_launches_equal_one_step_launches(seiswg, threads, seglen, group, nt, phys):
    """P-SV and SH forward with steps in pairs of two-step launches (the
    default) against one step per launch: state and traces bit for bit, for
    short segments (recomputed rows at both ends), several shots per
    workgroup, sources and receivers on edges, halos and duplicated cells, an
    odd step count (single steps after the pairs) and, with 37 steps,
    captured graphs."""
    nx, nz = 61, 150
    sh = phys == "sh"
    p, stf = random_problem(nx, nz, nt, sh, not sh, False, seed=29)
    print(f"case phys={phys} threads={threads} seglen={seglen} group={group} nt={nt}", flush=True)
    comps = [O.OBS_Y] if sh else [O.OBS_X, O.OBS_Z]
    fields = STATE[phys]
    runs = [0, 1, 2, 3]
    out = {}
    for spl in (1, 2):
        c = gpu_context(p, stf, seiswg)
        c.set_tiling(threads, seglen, group)
        c.set_steps_per_launch(spl)
        c.forward_batch(runs)
        tm = c.timing()
        out[spl] = ([[c.obs(q, k).copy() for k in comps] for q in range(len(runs))],
                    [[c.field(getattr(seiswg.abi, "F_" + n), shot=q).copy() for n in fields]
                     for q in range(len(runs))], tm.step_launches)
        c.close()
    assert out[1][2] == nt and out[2][2] == nt // 4 * 2 + nt % 4, (out[1][2], out[2][2])
    for q in range(len(runs)):
        for k, a, b in zip(comps, out[1][0][q], out[2][0][q]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (q, k)
        for n, a, b in zip(fields, out[1][1][q], out[2][1][q]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (q, n)
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    obs, _ = o.forward(1, 0)
    compare(out[2][0][1][0], obs[comps[0]], "two-step shot 1 trace")


@pytest.mark.parametrize("threads,seglen,group", [(64, 5, 1), (256, 9, 2), (128, 16, 4), (320, 7, 1), (512, 11, 2)])
@pytest.mark.parametrize("phys", ["psv", "sh"])
def test_receiver_kernel_equals_recording_in_the_march(seiswg, monkeypatch, threads, seglen, group, phys):
    """Traces of two-step launches recorded by the receiver kernel after each
    launch (what launches of many shot-cells do) and in the march itself
    (small launches), forced both ways, against one step per launch: traces
    and state bit for bit, with receivers on source cells (the receiver kernel
    repeats the launch's first-stage update there, injection included), on
    edges and duplicated, several shots per workgroup, 37 steps (captured
    graphs, a single step at the end)."""
    nx, nz, nt = 61, 150, 37
    sh = phys == "sh"
    p, stf = random_problem(nx, nz, nt, sh, not sh, False, seed=31, nrec=7)
    # receivers also on the source cells, one of them (the duplicated source
    # cell) twice
    p.rec_idx = list(p.rec_idx) + list(p.src_idx) + [p.src_idx[0]]
    print(f"case phys={phys} threads={threads} seglen={seglen} group={group}", flush=True)
    comps = [O.OBS_Y] if sh else [O.OBS_X, O.OBS_Z]
    fields = STATE[phys]
    runs = [0, 1, 2, 3]
    out = {}
    for mode, spl, env in (("one-step", 1, None), ("march", 2, "0"), ("kernel", 2, "1")):
        if env is None:
            monkeypatch.delenv("SEISWG_RECORD_KERNEL", raising=False)
        else:
            monkeypatch.setenv("SEISWG_RECORD_KERNEL", env)
        c = gpu_context(p, stf, seiswg)
        c.set_tiling(threads if threads <= 256 else 256, seglen, group)
        c.set_tiling2(threads, seglen, group)
        c.set_steps_per_launch(spl)
        c.forward_batch(runs)
        assert spl == 1 or c.tiling2()[0] == threads
        out[mode] = ([[c.obs(q, k).copy() for k in comps] for q in range(len(runs))],
                     [[c.field(getattr(seiswg.abi, "F_" + n), shot=q).copy() for n in fields]
                      for q in range(len(runs))])
        c.close()
    monkeypatch.delenv("SEISWG_RECORD_KERNEL", raising=False)
    for mode in ("march", "kernel"):
        for q in range(len(runs)):
            for k, a, b in zip(comps, out["one-step"][0][q], out[mode][0][q]):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (mode, q, k)
            for n, a, b in zip(fields, out["one-step"][1][q], out[mode][1][q]):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (mode, q, n)
    # the traces on source cells carry the injection
    assert np.any(out["kernel"][0][0][0][7:] != 0)
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    obs, _ = o.forward(2, 0)
    compare(out["kernel"][0][2][0], obs[comps[0]], "receiver-kernel shot 2 trace")


def test_random_shapes_two_step_receiver_paths_equal_one_step(seiswg, monkeypatch):
    """40 seeded random problems (SH or P-SV, grids 6-260 x 6-700, 1-6 shots,
    5-39 steps, random two-step tile widths, segment lengths and shots per
    workgroup, sources and receivers anywhere, receivers also on source cells
    and the grid corners): two-step launches with the receiver kernel and
    with recording in the march against one step per launch, traces and
    state bit for bit."""
    rng = np.random.default_rng(2024)
    shapes2 = [(64, 1), (128, 1), (128, 2), (128, 4), (256, 1), (256, 2), (256, 4), (320, 1), (320, 2), (512, 1),
               (512, 2)]
    for case in range(40):
        sh = bool(rng.integers(2))
        phys = "sh" if sh else "psv"
        nx, nz = int(rng.integers(6, 260)), int(rng.integers(6, 700))
        nshots, nt, npt = int(rng.integers(1, 7)), int(rng.integers(5, 40)), nx * nz
        model = {"rho": rng.uniform(1.0, 2.0, npt).astype(np.float32),
                 "vp": rng.uniform(1.5, 2.5, npt).astype(np.float32),
                 "vs": rng.uniform(0.5, 1.2, npt).astype(np.float32)}
        nsrc = int(rng.integers(1, 6))
        src_idx = [int(rng.integers(nx)) * nz + int(rng.integers(nz)) for _ in range(nsrc)]
        rec_idx = [int(rng.integers(nx)) * nz + int(rng.integers(nz)) for _ in range(int(rng.integers(1, 12)))]
        rec_idx += src_idx[:2] + [npt - 1, 0]
        p = O.Problem(nx=nx, nz=nz, nt=nt, dx=float(np.float32(1.7)), dz=float(np.float32(2.3)),
                      dt=float(np.float32(0.05)), sh=sh, psv=not sh, spin=False, abs_left=True, abs_right=True,
                      abs_bottom=True, abs_top=bool(rng.integers(2)), abs_width=int(rng.integers(0, 8)),
                      abs_alpha=float(np.float32(0.1)), model=model, src_rows=[], src_idx=src_idx, rec_idx=rec_idx)
        stf = rng.standard_normal((3, nsrc, nt)).astype(np.float32)
        nt2, group = shapes2[rng.integers(len(shapes2))]
        seglen = int(rng.integers(2, 120))
        runs = [int(rng.integers(nsrc)) for _ in range(nshots)]
        comps = [O.OBS_Y] if sh else [O.OBS_X, O.OBS_Z]
        print(f"case {case} {phys} {nx}x{nz} shots={nshots} nt={nt} tiling2=({nt2},{seglen},{group})", flush=True)
        out = {}
        for mode, spl, env in (("one-step", 1, None), ("march", 2, "0"), ("kernel", 2, "1")):
            if env is None:
                monkeypatch.delenv("SEISWG_RECORD_KERNEL", raising=False)
            else:
                monkeypatch.setenv("SEISWG_RECORD_KERNEL", env)
            c = gpu_context(p, stf, seiswg)
            c.set_steps_per_launch(spl)
            if spl == 2:
                c.set_tiling2(nt2, seglen, group)
            c.forward_batch(runs)
            out[mode] = [c.obs(q, k).copy() for q in range(nshots) for k in comps] + \
                        [c.field(getattr(seiswg.abi, "F_" + n), shot=q).copy() for q in range(nshots)
                         for n in STATE[phys]]
            c.close()
        monkeypatch.delenv("SEISWG_RECORD_KERNEL", raising=False)
        for mode in ("march", "kernel"):
            for q, (a, b) in enumerate(zip(out["one-step"], out[mode])):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (case, mode, q)


@pytest.mark.parametrize("tiling2", [(512, 7, 1), (512, 13, 2), (256, 20, 4), (128, 5, 2), (64, 9, 1)])
def test_two_step_shape_set_apart_from_one_step_shape(seiswg, tiling2):
    """Two-step launches with a tile shape of their own (sw_set_tiling2),
    other segment lengths and tile widths than the one-step launches of the
    same batch, against one step per launch, bit for bit."""
    nx, nz, nt = 61, 150, 23
    p, stf = random_problem(nx, nz, nt, False, True, False, seed=31)
    print(f"case tiling2={tiling2}", flush=True)
    runs = [0, 1, 2, 3]
    out = {}
    for spl in (1, 2):
        c = gpu_context(p, stf, seiswg)
        c.set_steps_per_launch(spl)
        if spl == 2:
            c.set_tiling2(*tiling2)
        c.forward_batch(runs)
        if spl == 2:
            assert c.tiling2() == tiling2
        out[spl] = [[c.field(getattr(seiswg.abi, "F_" + n), shot=q).copy() for n in STATE["psv"]] +
                    [c.obs(q, O.OBS_X).copy(), c.obs(q, O.OBS_Z).copy()] for q in range(len(runs))]
        c.close()
    for q in range(len(runs)):
        for k, (a, b) in enumerate(zip(out[1][q], out[2][q])):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (q, k)


@pytest.mark.parametrize("tiling2", [(320, 7, 1), (320, 13, 2)])
@pytest.mark.parametrize("phys", ["psv", "sh"])
def test_five_wave_two_step_tiles_equal_one_step_launches(seiswg, tiling2, phys):
    """320-thread two-step tiles (308-column strips: three strips and a
    partly filled last one on 700 columns) against one step per launch, bit
    for bit."""
    nx, nz, nt = 37, 700, 9
    sh = phys == "sh"
    p, stf = random_problem(nx, nz, nt, sh, not sh, False, seed=33)
    print(f"case phys={phys} tiling2={tiling2}", flush=True)
    comps = [O.OBS_Y] if sh else [O.OBS_X, O.OBS_Z]
    runs = [0, 1, 2, 3]
    out = {}
    for spl in (1, 2):
        c = gpu_context(p, stf, seiswg)
        c.set_steps_per_launch(spl)
        if spl == 2:
            c.set_tiling2(*tiling2)
        c.forward_batch(runs)
        if spl == 2:
            assert c.tiling2() == tiling2
        out[spl] = [[c.field(getattr(seiswg.abi, "F_" + n), shot=q).copy() for n in STATE[phys]] +
                    [c.obs(q, k).copy() for k in comps] for q in range(len(runs))]
        c.close()
    for q in range(len(runs)):
        for k, (a, b) in enumerate(zip(out[1][q], out[2][q])):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (q, k)


@pytest.mark.parametrize("tiling2", [(256, 5, 1), (384, 7, 1), (512, 13, 1), (384, 200, 1), (256, 1, 1)])
@pytest.mark.parametrize("nt", [13, 38])
def test_three_step_launches_equal_one_step_launches(seiswg, tiling2, nt):
    """P-SV forward with three steps per launch (sw_set_steps_per_launch(3):
    9-column halos, 7 + 8 recomputed rows per segment) against one step per
    launch, state and traces bit for bit: 700 columns (strips of 238, 366 and
    494 columns, the last one partly filled), segments of 1 to 200 rows,
    sources and receivers on edges and on duplicated cells, step counts that
    leave one or two single steps after the three-step launches, and with 38
    steps captured graph blocks."""
    nx, nz = 61, 700
    p, stf = random_problem(nx, nz, nt, False, True, False, seed=37)
    print(f"case tiling2={tiling2} nt={nt}", flush=True)
    runs = [0, 1, 2, 3]
    out, launches = {}, {}
    for spl in (1, 3):
        c = gpu_context(p, stf, seiswg)
        c.set_steps_per_launch(spl)
        if spl == 3:
            c.set_tiling2(*tiling2)
        c.forward_batch(runs)
        if spl == 3:
            assert c.tiling2() == tiling2
        launches[spl] = c.timing().step_launches
        out[spl] = [[c.field(getattr(seiswg.abi, "F_" + n), shot=q).copy() for n in STATE["psv"]] +
                    [c.obs(q, O.OBS_X).copy(), c.obs(q, O.OBS_Z).copy()] for q in range(len(runs))]
        c.close()
    assert launches[1] == nt and launches[3] == nt // 3 + nt % 3, launches
    for q in range(len(runs)):
        for k, (a, b) in enumerate(zip(out[1][q], out[3][q])):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (q, k)
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    obs, _ = o.forward(2, 0)
    compare(out[3][2][5], obs[O.OBS_X], "three-step shot 2 trace")


@pytest.mark.parametrize("phys", ["psv", "sh"])
def test_snapshots_equal_the_state_after_their_step(seiswg, phys):
    """Snapshots of a batched run (copied asynchronously while later steps
    run) against the velocity fields read after advancing to the same step:
    bit for bit, every shot, every snapshot."""
    nx, nz, nt, every = 48, 90, 23, 5
    sh = phys == "sh"
    p, stf = random_problem(nx, nz, nt, sh, not sh, False, seed=41)
    runs = [0, 1, 2]
    c = gpu_context(p, stf, seiswg)
    snaps = {}
    c.forward_batch(runs, snap=every, snapshots=snaps)
    fids = [seiswg.abi.F_VY] if sh else [seiswg.abi.F_VX, seiswg.abi.F_VZ]
    steps = sorted({k[1] for k in snaps})
    assert steps == list(range(every, nt, every)), steps
    c.prepare(runs)
    done = 0
    for it in steps:
        c.advance(it + 1 - done)
        done = it + 1
        for s in range(len(runs)):
            for f in fids:
                ref = c.field(f, shot=s)
                assert np.array_equal(snaps[(s, it, f)].view(np.uint32), ref.view(np.uint32)), (s, it, f)


@pytest.mark.parametrize("phys", ["psv", "sh", "spin"])
def test_autotuned_shapes_keep_results(seiswg, phys):
    """sw_autotune times one-step shapes and, where steps run two per launch,
    two-step shapes; it keeps the fastest, and runs with them give the same
    traces and states as the heuristic shapes, bit for bit."""
    nx, nz, nt = 40, 150, 22
    sh, psv, spin = phys == "sh", phys in ("psv", "spin"), phys == "spin"
    p, stf = random_problem(nx, nz, nt, sh, psv, spin, seed=51)
    runs = [0, 1, 2, 3]
    comps = [O.OBS_Y] if sh or spin else [O.OBS_X, O.OBS_Z]
    fields = STATE["sh"] if sh else STATE["psv"] + (STATE["spin"] if spin else [])
    out = []
    for tune in (False, True):
        c = gpu_context(p, stf, seiswg)
        if tune:
            threads, seglen, group, ms = c.autotune(len(runs), steps=4)
            print(f"case {phys}: tuned one-step {c.tiling()} two-step {c.tiling2()} ({ms:.4f} ms/step)")
            assert ms > 0 and threads in (64, 128, 256)
        c.forward_batch(runs)
        if tune:
            assert c.tiling() == (threads, seglen, group)
            # two-step shapes: marching or LDS tiles, LDS tiles only with spin
            assert c.tiling2()[0] > 0 and (not spin or c.tiling2()[0] == 1024)
        out.append([c.obs(q, k).copy() for q in range(len(runs)) for k in comps] +
                   [c.field(getattr(seiswg.abi, "F_" + n), shot=q).copy() for q in range(len(runs)) for n in fields])
        c.close()
    for a, b in zip(*out):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def load_example(tmp_path, ex, ini_name="config.ini"):
    d = E.materialise(ex, str(tmp_path))
    ini = E.ini_values(os.path.join(d, ini_name))
    mdir = ini["paths"].get("model_init") or ini["paths"].get("model_true")
    p = O.load_problem(os.path.join(d, mdir), os.path.join(d, ini["paths"]["sources"]),
                       os.path.join(d, ini["paths"]["stations"]), ini)
    return d, p, ini


@pytest.mark.parametrize("reciprocal", [False, True], ids=["exact", "reciprocal"])
@pytest.mark.parametrize("ex,comp,fid,snap_file,trace_file",
                         [("forward", O.OBS_Y, 0, "proc000100_vy.bin", "uy_000000.npy"),
                          ("spin", O.OBS_Y, 7, "proc000200_vx.bin", "ry_000000.npy")])
def test_examples_match_fixtures_and_oracle(seiswg, tmp_path, ex, comp, fid, snap_file, trace_file, reciprocal):
    """Exact divisions: the oracle bit for bit (compare's 1-ulp-level bar) and
    the reference fixtures within their tolerance.  Divisions by the rounded
    reciprocal (sw_set_division): the same fixture tolerance, and rel-L2 <=
    5e-6 against the oracle over the whole run (1200 / 2000 steps; measured
    1.1e-6 on the forward example's traces, where the exact path itself is
    1.2e-6 from the reference's own outputs)."""
    d, p, ini = load_example(tmp_path, ex)
    c = seiswg.Context.from_problem(p)
    if reciprocal:
        c.set_division(True)
    stf = seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in p.src_rows], p.nt, p.dt)
    c.set_sources(p.src_idx, stf)
    c.set_receivers(p.rec_idx)
    snaps = {}
    c.forward(-1, ini["save_snapshot"], snaps)
    step = int(snap_file[4:10])
    snap = snaps[(0, step, fid)]
    tr = c.obs(0, comp)
    ref_sn, ref_tr = E.fixture(ex, snap_file), E.fixture(ex, trace_file)
    print(f"{ex}: trace rl2 {O.rel_l2(tr, ref_tr):.2e} max {O.max_rel_err(tr, ref_tr):.2e}; "
          f"snap rl2 {O.rel_l2(snap, ref_sn):.2e}")
    assert O.max_rel_err(tr, ref_tr) <= 1e-4 and O.rel_l2(tr, ref_tr) <= 1e-5
    assert O.max_rel_err(snap, ref_sn) <= 1e-4 and O.rel_l2(snap, ref_sn) <= 1e-5
    o = O.Oracle(p, 4)
    oobs, osnaps = o.forward(-1, ini["save_snapshot"])
    q = {0: 0, 7: 1}[fid]
    e_tr, e_sn = O.rel_l2(tr, oobs[comp]), O.rel_l2(snap, osnaps[step // ini["save_snapshot"] - 1][q])
    print(f"{ex} ({'reciprocal' if reciprocal else 'exact'} division): vs oracle trace {e_tr:.2e} snapshot {e_sn:.2e}")
    tol = 5e-6 if reciprocal else 1e-6
    compare(tr, oobs[comp], "trace vs oracle", tol)
    compare(snap, osnaps[step // ini["save_snapshot"] - 1][q], "snapshot vs oracle", tol)


def run_cli(ini):
    import seiswg_amd
    r = subprocess.run([seiswg_amd.CLI_PATH, ini], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, RUST_LOG="error"))
    assert r.returncode == 0, r.stderr
    return r


def test_cli_forward_and_spin_examples_reproduce_fixtures(seiswg, tmp_path):
    for ex, files in [("forward", [("proc000100_vy.bin", "proc000100_vy.bin"),
                                   ("traces/uy_000000.npy", "uy_000000.npy")]),
                      ("spin", [("proc000200_vx.bin", "proc000200_vx.bin"), ("ry_000000.npy", "ry_000000.npy")])]:
        d = E.materialise(ex, str(tmp_path))
        run_cli(os.path.join(d, "config.ini"))
        for got_name, fx in files:
            gp = os.path.join(d, "output", got_name)
            fp = os.path.join(E.GOLDEN, "fixtures", ex, fx)
            if fx.endswith(".npy"):
                gb, fb = open(gp, "rb").read(), open(fp, "rb").read()
                hl = 10 + int.from_bytes(fb[8:10], "little")
                assert gb[:hl] == fb[:hl], "NPY header differs from the reference's"
                got, ref = E.read_npy_payload(gp), E.read_npy_payload(fp)
            else:
                got, ref = np.fromfile(gp, "<f4"), np.fromfile(fp, "<f4")
            assert got.size == ref.size
            assert O.max_rel_err(got, ref) <= 1e-4 and O.rel_l2(got, ref) <= 1e-5, (ex, got_name)
    # the forward example writes 11 snapshots (it = 100..1100) and raw vy traces
    d = os.path.join(str(tmp_path), "forward", "output")
    snaps = sorted(f for f in os.listdir(d) if f.startswith("proc"))
    assert snaps == [f"proc{it:06d}_vy.bin" for it in range(100, 1200, 100)]
    raw = np.fromfile(os.path.join(d, "traces", "vy_000000.npy"), "<f4")
    assert np.array_equal(raw, E.read_npy_payload(os.path.join(d, "traces", "uy_000000.npy")))


def test_cli_adjoint_example_reproduces_kmu_fixture(seiswg, tmp_path):
    """run_example.bash's adjoint flow: config_true forward writes the observed
    traces, then the adjoint workflow; K_mu against the reference fixture."""
    d = E.materialise("adjoint", str(tmp_path))
    run_cli(os.path.join(d, "config_true.ini"))
    sub = E.obs_traces_every16()
    for s in (0, 6, 11):
        got = np.load(os.path.join(d, "obs_traces", f"uy_{s:06d}.npy"))[:, ::16]
        assert O.max_rel_err(got, sub[f"uy_{s:06d}"]) <= 1e-4 and O.rel_l2(got, sub[f"uy_{s:06d}"]) <= 1e-5
    run_cli(os.path.join(d, "config.ini"))
    got = np.fromfile(os.path.join(d, "output", "proc000000_kmu.bin"), "<f4")
    ref = E.fixture("adjoint", "proc000000_kmu.bin")
    print(f"kmu rl2 {O.rel_l2(got, ref):.2e} max {O.max_rel_err(got, ref):.2e}")
    assert O.max_rel_err(got, ref) <= 1e-4 and O.rel_l2(got, ref) <= 1e-5
    hdr = np.fromfile(os.path.join(d, "output", "gradient_kmu.bin"), "<i4", count=1)[0]
    assert hdr == got.size


def test_cli_adjoint_example_with_combined_sources_matches_oracle(seiswg, tmp_path):
    """The adjoint example with combine_sources = yes in both configs (one
    run in which all twelve sources fire, traces under index 0): the CLI's
    K_mu bit for bit against the FP32 oracle's plain adjoint of the same
    single run."""
    d = E.materialise("adjoint", str(tmp_path))
    for name in ("config_true.ini", "config.ini"):
        path = os.path.join(d, name)
        text = open(path).read().replace("combine_sources = no", "combine_sources = yes")
        assert "combine_sources = yes" in text
        open(path, "w").write(text)
    run_cli(os.path.join(d, "config_true.ini"))
    files = sorted(os.listdir(os.path.join(d, "obs_traces")))
    assert "uy_000000.npy" in files and not any(f.startswith("uy_") and f != "uy_000000.npy" for f in files), files
    r = run_cli(os.path.join(d, "config.ini"))
    got = np.fromfile(os.path.join(d, "output", "proc000000_kmu.bin"), "<f4")
    _, p, ini = load_example(tmp_path / "o", "adjoint")
    ref_obs = E.read_npy_payload(os.path.join(d, "obs_traces", "uy_000000.npy")).reshape(len(p.rec_idx), p.nt)
    o = O.Oracle(p, 4)
    kmu_o, mis_o = o.adjoint_plain([-1], ref_obs[None], ini["adj_interval"], ini["smooth"])
    err = O.rel_l2(got, kmu_o)
    print(f"combined-source K_mu vs oracle: rel-L2 {err:.2e}, max-rel {O.max_rel_err(got, kmu_o):.2e}")
    want = np.asarray(kmu_o, np.float64).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))     # the same operations in the same order
    assert np.any(got != 0)


def example_files(d, dirs):
    """The files of the example's model directories, keyed by relative path."""
    return {f"{sub}/{n}": open(os.path.join(d, sub, n), "rb").read()
            for sub in dirs for n in sorted(os.listdir(os.path.join(d, sub)))}


@pytest.mark.parametrize("ex,nframes,fields", [("forward", 11, ["vy"]), ("spin", 9, ["vx", "vz", "ry"])])
def test_in_memory_run_equals_cli_files(seiswg, tmp_path, monkeypatch, ex, nframes, fields):
    """seiswg_run_simulation on the forward and spin examples: every trace
    file the CLI writes, byte for byte under the same relative path; the
    snapshot frames equal to the CLI's proc<step>_<field>.bin files in step
    order; the grid scalars; and nothing written to disk."""
    d = E.materialise(ex, str(tmp_path))
    run_cli(os.path.join(d, "config.ini"))
    empty = tmp_path / "cwd"
    empty.mkdir()
    monkeypatch.chdir(empty)
    text = lambda n: open(os.path.join(d, n)).read()
    out = seiswg.run_simulation(text("config.ini"), text("sources.dat"), text("stations.dat"),
                                example_files(d, ["model"]))
    assert os.listdir(empty) == []
    written = set()
    for root, _, names in os.walk(os.path.join(d, "output")):
        for n in names:
            if not n.startswith("proc"):
                written.add(os.path.relpath(os.path.join(root, n), d))
    traces = {k for k in out if k.endswith(".npy")}
    assert traces == written and traces
    for k in traces:
        assert out[k] == open(os.path.join(d, k), "rb").read(), k
    assert np.frombuffer(out["snapshot/nframes"], "<u4")[0] == nframes
    snap_dir = os.path.join(d, "output")
    steps = sorted(int(n[4:10]) for n in os.listdir(snap_dir) if n.startswith("proc") and n.endswith(f"_{fields[0]}.bin"))
    assert len(steps) == nframes
    for f in fields:
        cli = b"".join(open(os.path.join(snap_dir, f"proc{it:06d}_{f}.bin"), "rb").read() for it in steps)
        assert out[f"snapshot/{f}.f32"] == cli, f
    u32 = lambda k: int(np.frombuffer(out[k], "<u4")[0])
    ini = E.ini_values(os.path.join(d, "config.ini"))
    assert u32("nt") == ini["nt"] and np.frombuffer(out["dt"], "<f4")[0] == np.float32(ini["dt"])
    rows = [ln for ln in text("stations.dat").splitlines() if ln.strip() and not ln.strip().startswith("#")]
    assert u32("nx") * u32("nz") == u32("npt") and u32("nrec") == len(rows)


def test_in_memory_adjoint_with_true_config_reproduces_kmu_fixture(seiswg, tmp_path, monkeypatch):
    """The adjoint example in memory, as lib.rs runs it: the true-model
    forward (config_true.ini) feeds its traces to the adjoint run; K_mu
    against the reference fixture, the model-format copy, and nothing on disk."""
    d = E.materialise("adjoint", str(tmp_path))
    empty = tmp_path / "cwd"
    empty.mkdir()
    monkeypatch.chdir(empty)
    text = lambda n: open(os.path.join(d, n)).read()
    out = seiswg.run_simulation(text("config.ini"), text("sources.dat"), text("stations.dat"),
                                example_files(d, ["model_init", "model_true"]), text("config_true.ini"))
    assert os.listdir(empty) == []
    got = np.frombuffer(out["output/proc000000_kmu.bin"], "<f4")
    ref = E.fixture("adjoint", "proc000000_kmu.bin")
    print(f"kmu rl2 {O.rel_l2(got, ref):.2e} max {O.max_rel_err(got, ref):.2e}")
    assert O.max_rel_err(got, ref) <= 1e-4 and O.rel_l2(got, ref) <= 1e-5
    g = out["output/gradient_kmu.bin"]
    assert np.frombuffer(g[:4], "<i4")[0] == got.size and g[4:] == out["output/proc000000_kmu.bin"]
    assert not any(k.startswith("obs_traces/") for k in out)


def test_in_memory_per_source_runs_and_observed_traces_from_the_map(seiswg, tmp_path):
    """The adjoint example's true-model forward in memory with snapshots on
    (12 per-source runs, a snapshot every 1600 of 4800 steps): 24 frames,
    run after run and step after step, equal to the device's snapshots of the
    same batch; then the adjoint workflow with those traces passed in the
    input map (no true config) reproduces the reference's K_mu fixture."""
    d = E.materialise("adjoint", str(tmp_path))
    text = lambda n: open(os.path.join(d, n)).read()
    true_ini = text("config_true.ini").replace("save_snapshot   = 0", "save_snapshot   = 1600")
    assert "save_snapshot   = 1600" in true_ini
    files = example_files(d, ["model_init", "model_true"])
    fwd = seiswg.run_simulation(true_ini, text("sources.dat"), text("stations.dat"), files)
    assert np.frombuffer(fwd["snapshot/nframes"], "<u4")[0] == 24
    _, p, _ = load_example(tmp_path, "adjoint", "config_true.ini")
    c = seiswg.Context.from_problem(p)
    c.set_sources(p.src_idx, seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in p.src_rows], p.nt, p.dt))
    c.set_receivers(p.rec_idx)
    snaps = {}
    c.forward_batch(list(range(12)), 1600, snaps)
    want = b"".join(snaps[(s, it, seiswg.abi.F_VY)].tobytes() for s in range(12) for it in (1600, 3200))
    assert fwd["snapshot/vy.f32"] == want
    c.close()
    traces = {k: v for k, v in fwd.items() if k.startswith("obs_traces/")}
    assert sorted(k for k in traces if "/uy_" in k) == [f"obs_traces/uy_{s:06d}.npy" for s in range(12)]
    adj = seiswg.run_simulation(text("config.ini"), text("sources.dat"), text("stations.dat"), {**files, **traces})
    got = np.frombuffer(adj["output/proc000000_kmu.bin"], "<f4")
    ref = E.fixture("adjoint", "proc000000_kmu.bin")
    assert O.max_rel_err(got, ref) <= 1e-4 and O.rel_l2(got, ref) <= 1e-5


def test_cli_inversion_lowers_the_misfit(seiswg, tmp_path):
    """workflow = inversion on the adjoint example: observed traces from the
    true model, two steepest-descent iterations from the initial model; the
    misfit log decreases, the models are written in the proc000000_*.bin
    format (rho and the coordinates unchanged, vs changed)."""
    d = E.materialise("adjoint", str(tmp_path))
    run_cli(os.path.join(d, "config_true.ini"))
    ini = open(os.path.join(d, "config.ini")).read().replace("mode = adjoint", "mode = inversion")
    ini += "\n[inversion]\nniter = 2\nstep = 0.02\nlinesearch = 4\ncheckpoint = 100\n"
    cfg = os.path.join(d, "config_inv.ini")
    open(cfg, "w").write(ini)
    run_cli(cfg)
    log = np.loadtxt(os.path.join(d, "output", "misfit.txt"))
    print("misfit per iteration:", log[:, 1])
    assert log.shape[0] >= 2 and np.all(np.diff(log[:, 1]) < 0)
    last = os.path.join(d, "output", f"iter_{int(log[-1, 0]):03d}")
    npt = 201 * 201

    def field(dirname, name):
        raw = np.fromfile(os.path.join(dirname, f"proc000000_{name}.bin"), "<f4")
        assert raw[:1].view("<i4")[0] == npt
        return raw[1:]

    init = os.path.join(d, "model_init")
    for name in ("rho", "x", "z"):
        assert np.array_equal(field(last, name), field(init, name)), name
    assert not np.array_equal(field(last, "vs"), field(init, "vs"))


def test_in_memory_inversion_with_true_config(seiswg, tmp_path, monkeypatch):
    """workflow = inversion in memory: the true-model forward of
    config_true.ini feeds the observed traces, two iterations lower the
    misfit, and the accepted models come back in the model-file format."""
    d = E.materialise("adjoint", str(tmp_path))
    empty = tmp_path / "cwd"
    empty.mkdir()
    monkeypatch.chdir(empty)
    text = lambda n: open(os.path.join(d, n)).read()
    ini = text("config.ini").replace("mode = adjoint", "mode = inversion")
    ini += "\n[inversion]\nniter = 2\nstep = 0.02\nlinesearch = 4\ncheckpoint = 100\n"
    files = example_files(d, ["model_init", "model_true"])
    out = seiswg.run_simulation(ini, text("sources.dat"), text("stations.dat"), files, text("config_true.ini"))
    assert os.listdir(empty) == []
    log = np.array([[float(x) for x in ln.split()] for ln in out["output/misfit.txt"].decode().splitlines()])
    print("misfit per iteration:", log[:, 1])
    assert log.shape[0] >= 2 and np.all(np.diff(log[:, 1]) < 0)
    last = f"output/iter_{int(log[-1, 0]):03d}"
    npt = 201 * 201
    for name in ("rho", "x", "z", "vs"):
        raw = out[f"{last}/proc000000_{name}.bin"]
        assert np.frombuffer(raw[:4], "<i4")[0] == npt and len(raw) == 4 + 4 * npt
        same = raw == files[f"model_init/proc000000_{name}.bin"][:4 + 4 * npt]
        assert same == (name != "vs"), name


def test_adjoint_plain_matches_oracle(seiswg):
    nx, nz, nt = 50, 46, 300
    p, _ = random_problem(nx, nz, nt, True, False, False, seed=8, nsrc=3, nrec=6)
    p.src_rows = [[0, 0, 0, 3.0, 0.3, 0.0, 1.0], [0, 0, 0, 4.0, 0.25, 0.0, 1.0], [0, 0, 0, 5.0, 0.2, 0.0, 1.0]]
    stf = O.ricker_table(4, p.src_rows, nt, p.dt).astype(np.float32)
    rng = np.random.default_rng(2)
    ref = rng.standard_normal((3, len(p.rec_idx), nt)).astype(np.float32) * 1e-3
    c = gpu_context(p, stf, seiswg)
    kmu, mis = c.adjoint_plain([0, 1, 2], ref, 10, 2.5)
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    kmu_o, mis_o = o.adjoint_plain([0, 1, 2], ref, 10, 2.5)
    compare(kmu, kmu_o, "kmu", tol=1e-6)
    np.testing.assert_allclose(mis, mis_o, rtol=1e-9)


def test_smoothing_matches_oracle(seiswg):
    nx, nz = 37, 53
    p, stf = random_problem(nx, nz, 4, True, False, False, seed=1)
    c = gpu_context(p, stf, seiswg)
    rng = np.random.default_rng(4)
    k = rng.standard_normal(nx * nz).astype(np.float32)
    c.adjoint_plain([0], np.zeros((1, len(p.rec_idx), 4), np.float32), 1, 3.0)   # allocates K_mu
    c.write_field(seiswg.abi.F_K_MU, k)
    c.smooth(seiswg.abi.F_K_MU, 3.0)
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    o.set_field(O.F["K_MU"], k)
    o.smooth(3.0)
    compare(c.field(seiswg.abi.F_K_MU), o.field(O.F["K_MU"]), "smoothed", tol=1e-6)


def test_setters_discard_a_prepared_batch(seiswg):
    """sw_set_receivers / sw_set_sources after sw_prepare: the batch's traces,
    accumulators and lists have the old counts, so the library refuses to
    advance (or read its traces) until the batch is prepared again."""
    p, stf = random_problem(30, 40, 12, False, True, False, seed=1)
    c = gpu_context(p, stf, seiswg)
    c.prepare([0])
    c.advance(2)
    c.set_receivers(p.rec_idx[:3])
    with pytest.raises(seiswg.SeiswgError, match="no prepared batch"):
        c.advance(1)
    with pytest.raises(seiswg.SeiswgError, match="not in last run"):
        c.obs(0, seiswg.abi.OBS_X)
    c.prepare([0])
    c.advance(2)
    c.set_sources(p.src_idx[:2], stf[:, :2])
    with pytest.raises(seiswg.SeiswgError, match="no prepared batch"):
        c.advance(1)
    c.prepare([1])
    c.advance(3)


def _spin_problem(seiswg, j_value):
    from seiswg_amd import synth
    q = synth.spin_forward_problem(j_value)
    nx, nz, h, nt, dt, g, m = q["nx"], q["nz"], q["h"], q["nt"], q["dt"], q["geom"], q["model"]
    ctx = seiswg.Context(nx, nz, nt, h, h, dt, sh=False, psv=True, spin=True, abs_left=True, abs_right=True,
                         abs_bottom=True, abs_top=True, abs_width=q["abs_width"], abs_alpha=0.015)
    ctx.set_model(seiswg.abi.MODEL_LAME, [m[k] for k in ("rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c")])
    rows = g.src_rows()
    ctx.set_sources([seiswg.grid_index(r[0], r[1], h, h, nz) for r in rows],
                    seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in rows], nt, dt))
    ctx.set_receivers([seiswg.grid_index(x, z, h, h, nz) for x, z in zip(g.rec_x, g.rec_z)])
    return ctx


def test_diverging_run_is_reported_with_its_step(seiswg):
    """The 500x500 P-SV + spin configuration with its stated micro-inertia
    j = 1e-2 (omega * dt ~ 358: the explicit scheme diverges): the forward
    run returns an error naming the first snapshot step whose state is not
    finite; with j = 1e4 the run is finite and passes; with the check off
    the diverged run completes."""
    print("case spin 500x500 j=1e-2 snap=500", flush=True)
    bad = _spin_problem(seiswg, 1e-2)
    snaps = {}
    with pytest.raises(seiswg.SeiswgError, match=r"not finite .*after step 500\b"):
        bad.forward_batch([-1], snap=500, snapshots=snaps)
    bad.set_finite_check(False)
    bad.forward_batch([-1])
    assert not np.all(np.isfinite(bad.obs(0, seiswg.abi.OBS_X)))
    print("case spin 500x500 j=1e4", flush=True)
    good = _spin_problem(seiswg, 1e4)
    good.forward_batch([-1], snap=500, snapshots=snaps)
    assert np.all(np.isfinite(good.obs(0, seiswg.abi.OBS_X)))


def test_graph_blocks_reused_across_prepared_batches(seiswg, monkeypatch):
    """Runs of 16 steps or more replay captured 16-step blocks; a batch
    prepared again with the same buffers and shapes replays the same blocks
    (no new capture), and the results equal launches without graphs."""
    p, stf = random_problem(60, 70, 80, False, True, False, seed=3)
    c = gpu_context(p, stf, seiswg)
    c.prepare([0, 1])
    c.advance(40)
    state, n, msg = c.graph_state()
    assert state == 1 and n >= 1 and msg == ""
    got = c.field(seiswg.abi.F_VX, 1)
    c.prepare([0, 1])
    c.advance(40)
    assert c.graph_state()[1] == n
    assert np.array_equal(c.field(seiswg.abi.F_VX, 1).view(np.uint32), got.view(np.uint32))
    monkeypatch.setenv("SEISWG_NO_GRAPHS", "1")
    d = gpu_context(p, stf, seiswg)
    d.prepare([0, 1])
    d.advance(40)
    assert d.graph_state()[:2] == (0, 0)
    assert np.array_equal(d.field(seiswg.abi.F_VX, 1).view(np.uint32), got.view(np.uint32))


@pytest.mark.parametrize("phys,k", [("sh", 2), ("sh", 3), ("sh", 4), ("psv", 2), ("psv", 3), ("psv", 4),
                                    ("spin", 2), ("spin", 3)])
@pytest.mark.parametrize("nx,nz", [(45, 131), (30, 40), (61, 150)])
def test_lds_tile_multi_step_launches_match_oracle_bitwise(seiswg, phys, k, nx, nz):
    """K steps per launch on LDS tiles (small grids; sw_set_tiling2 with 1024
    threads and K = 2..4): three shots from random states, sources and
    receivers on edges and duplicated, 11 steps launched directly (launch
    units - pairs of even-K launches, single odd-K launches - then single
    steps), then 29 through the captured blocks (16 steps for K = 2 and 4, 12
    for K = 3, replayed back to back) - every state field and trace bit for
    bit against the FP32 oracle."""
    sh, psv, spin = phys == "sh", phys in ("psv", "spin"), phys == "spin"
    nsteps = 40
    p, stf = random_problem(nx, nz, nsteps + 1, sh, psv, spin, seed=31)
    print(f"case phys={phys} K={k} grid={nx}x{nz}", flush=True)
    c = gpu_context(p, stf, seiswg)
    c.set_tiling2(1024, k, 0)
    runs = [0, -1, 2]
    c.prepare(runs)
    assert c.tiling2()[:2] == (1024, k)
    fields = (STATE["sh"] if sh else []) + (STATE["psv"] if psv else []) + (STATE["spin"] if spin else [])
    rng = np.random.default_rng(9)
    init = []
    for s in range(len(runs)):
        st = {n: rng.standard_normal(nx * nz).astype(np.float32) for n in fields}
        for n, v in st.items():
            c.write_field(getattr(seiswg.abi, "F_" + n), v, shot=s)
        init.append(st)
    c.advance(11)
    c.advance(nsteps - 11)
    gst, ncap, _ = c.graph_state()
    assert gst == 1 and ncap >= 1, (gst, ncap)
    comps = ([O.OBS_Y] if sh or spin else []) + ([O.OBS_X, O.OBS_Z] if psv else []) + \
            ([O.OBS_YI, O.OBS_YC] if spin else [])
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    for s, isrc in enumerate(runs):
        o.clear()
        for n, v in init[s].items():
            o.set_field(O.F[n], v)
        for it in range(nsteps):
            o.step(it, isrc)
        for n in fields:
            got = c.field(getattr(seiswg.abi, "F_" + n), shot=s)
            want = np.asarray(o.field(O.F[n]), np.float64).astype(np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (s, n)
        obs = o.obs()
        for k in comps:
            got = c.obs(s, k)[:, :nsteps]
            want = np.asarray(obs[k][:, :nsteps], np.float64).astype(np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (s, k)


@pytest.mark.parametrize("shape", [(4, 200), (200, 4), (7, 7), (4, 500), (600, 5), (4, 4), (5, 70000), (70000, 5)])
@pytest.mark.parametrize("phys", ["sh", "psv", "spin"])
def test_degenerate_grid_shapes_match_oracle(seiswg, phys, shape):
    """Grids narrower than a stencil or a halo in one or both directions (the
    FD operators are zero near every edge, so such grids mostly scale and
    taper), on the library's default launches (LDS tiles up to 300 000
    shot-cells, marching tiles for the 350 000-cell grids): state and traces
    of two shots bit for bit against the FP32 oracle.
    Grids below 4x4 are refused with an error naming the limit."""
    nx, nz = shape
    sh, psv, spin = phys == "sh", phys in ("psv", "spin"), phys == "spin"
    nt = 21
    rng = np.random.default_rng(nx * 1000 + nz)
    npt = nx * nz
    if spin:
        model = {n: rng.uniform(0.5, 1.5, npt).astype(np.float32)
                 for n in ["rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c"]}
    else:
        model = {"rho": rng.uniform(1.0, 2.0, npt).astype(np.float32),
                 "vp": rng.uniform(1.5, 2.5, npt).astype(np.float32),
                 "vs": rng.uniform(0.5, 1.2, npt).astype(np.float32)}
    cells = [(nx // 2, nz // 2), (0, 0), (nx - 1, nz - 1)]
    src_idx = [i * nz + j for i, j in cells]
    rec_idx = [i * nz + j for i, j in cells + [(nx - 1, 0), (0, nz - 1)]]
    p = O.Problem(nx=nx, nz=nz, nt=nt, dx=float(np.float32(1.7)), dz=float(np.float32(2.3)),
                  dt=float(np.float32(0.05)), sh=sh, psv=psv, spin=spin, abs_left=True, abs_right=True,
                  abs_bottom=True, abs_top=True, abs_width=6, abs_alpha=float(np.float32(0.1)), model=model,
                  src_rows=[], src_idx=src_idx, rec_idx=rec_idx)
    stf = rng.standard_normal((3, len(src_idx), nt)).astype(np.float32)
    print(f"case phys={phys} shape={shape}", flush=True)
    c = gpu_context(p, stf, seiswg)
    runs = [0, 2]
    c.forward_batch(runs)
    comps = ([O.OBS_Y] if sh or spin else []) + ([O.OBS_X, O.OBS_Z] if psv else []) + ([O.OBS_YI] if spin else [])
    fields = (STATE["sh"] if sh else []) + (STATE["psv"] if psv else []) + (STATE["spin"] if spin else [])
    o = O.Oracle(p, 4, stf=stf.astype(np.float64))
    for q, r in enumerate(runs):
        obs, _ = o.forward(r, 0)
        for comp in comps:
            got = c.obs(q, comp)
            want = np.asarray(obs[comp], np.float64).astype(np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (r, comp)
        for name in fields:
            got = c.field(getattr(seiswg.abi, "F_" + name), shot=q)
            want = np.asarray(o.field(O.F[name]), np.float64).astype(np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (r, name)

    with pytest.raises(seiswg.SeiswgError, match="at least 4x4"):
        seiswg.Context(3, 200, 5, 1.0, 1.0, 0.1, sh=sh, psv=psv, spin=spin)
