"""Benchmark of the fused gfx950 time step on the batched multi-shot P-SV
workload (64 shots sharing one 2000x1000 layered-plus-random-field model,
400 receivers, taper left/right/bottom width 40, dt = 1 ms).

A step = one time step of all 64 shots (velocity + stress update, source
injection and receiver recording fused); by default steps run two per launch
(the state of the intermediate step stays on chip).  `value` is
shot-cell-updates per second over the K timed steps (inputs resident in HBM).
`roofline.achieved` is the algorithmic traffic of the cell-updates one launch
performs, 64 * nx * nz * 4 * (2*5 + 3/64) bytes per step (5 state fields
read+written per shot, 3 model fields read once per step), divided by the
kernel's average duration from HIP events around the timed launches;
`roofline.traffic` is the measured HBM traffic of one launch (hardware
counters), which two-step launches keep below the algorithmic bytes.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
METRIC = "cell-updates/s + HBM GB/s % peak; batched gradient shots/s vs plain"


# sw_multi's default merge limit (shot-cells per step below which gradients
# share merged launches): 0, every problem on its own launches
DEFAULT_MERGE_LIMIT = 0


def build_context(seiswg, synth, nt):
    c = synth.PSV_BATCH
    model, geom = synth.psv_batch_problem()
    ctx = seiswg.Context(c["nx"], c["nz"], nt, c["h"], c["h"], c["dt"], sh=False, psv=True, spin=False,
                         abs_left=True, abs_right=True, abs_bottom=True, abs_top=False, abs_width=c["abs_width"],
                         abs_alpha=0.015)
    ctx.set_model(seiswg.abi.MODEL_VPVS, [model["rho"], model["vp"], model["vs"]])
    rows = geom.src_rows()
    dx = dz = c["h"]
    src_idx = [seiswg.grid_index(r[0], r[1], dx, dz, c["nz"]) for r in rows]
    rec_idx = [seiswg.grid_index(x, z, dx, dz, c["nz"]) for x, z in zip(geom.rec_x, geom.rec_z)]
    stf = seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in rows], nt, c["dt"])
    ctx.set_sources(src_idx, stf)
    ctx.set_receivers(rec_idx)
    return ctx, model, geom, src_idx, rec_idx


def single_shot_bench(seiswg, synth, model, geom, src_idx, rec_idx, steps=400):
    """One shot of the 2000x1000 P-SV workload (heuristic tile shapes): its
    per-step working set (5 state fields read and written, 3 model fields:
    ~104 MB) fits the 256 MB Infinity Cache, so the rate is not an HBM
    figure; reported beside the batched one, which streams 5 GB per step."""
    c = synth.PSV_BATCH
    nt = steps + 20
    ctx = seiswg.Context(c["nx"], c["nz"], nt, c["h"], c["h"], c["dt"], sh=False, psv=True, spin=False,
                         abs_left=True, abs_right=True, abs_bottom=True, abs_top=False, abs_width=c["abs_width"],
                         abs_alpha=0.015)
    ctx.set_model(seiswg.abi.MODEL_VPVS, [model["rho"], model["vp"], model["vs"]])
    ctx.set_sources(src_idx, seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in geom.src_rows()], nt, c["dt"]))
    ctx.set_receivers(rec_idx)
    ctx.prepare([len(src_idx) // 2])
    ctx.advance(20)
    t0 = time.perf_counter()
    ctx.advance(steps)
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    cells = c["nx"] * c["nz"]
    out = {"workload": "single-shot P-SV forward, 2000x1000, 400 receivers (cache-resident working set)",
           "us_per_step": wall / steps * 1e6, "cell_updates_per_s": cells * steps / wall,
           "algorithmic_GBps": cells * 4.0 * 13 * steps / wall / 1e9,
           "kernel_avg_us": tm.loop_ms * 1e3 / max(tm.step_launches, 1),
           "tiling": list(ctx.tiling()), "tiling_two_step": list(ctx.tiling2())}
    ctx.close()
    return out


def cpu_baseline(synth, model, geom, src_idx, rec_idx, budget_s=12.0):
    """The oracle's FP32 restatement of the reference scheme (unfused, one
    loop per WGSL kernel, OpenMP over cells), one shot of the same workload,
    timed for about budget_s seconds."""
    from oracle import oracle as O
    c = synth.PSV_BATCH
    p = O.Problem(nx=c["nx"], nz=c["nz"], nt=4000, dx=c["h"], dz=c["h"], dt=c["dt"], sh=False, psv=True,
                  spin=False, abs_left=True, abs_right=True, abs_bottom=True, abs_top=False,
                  abs_width=c["abs_width"], abs_alpha=0.015, model=model, src_rows=geom.src_rows(),
                  src_idx=src_idx, rec_idx=rec_idx)
    o = O.Oracle(p, 4)
    cells = c["nx"] * c["nz"]

    def timed(budget):
        o.clear()
        o.step(0, 0)     # first-touch
        n, shot, it, t0 = 0, 0, 1, time.perf_counter()
        while True:
            o.step(it, shot)
            n += 1
            it += 1
            el = time.perf_counter() - t0
            if el >= budget:
                break
            if it == p.nt:          # next shot of the batch, from a cleared state
                shot, it = shot + 1, 0
                o.clear()
        return n, shot, el

    cores = int(O.lib().or_num_threads())
    n, shot, el = timed(budget_s)
    O.lib().or_set_num_threads(1)
    try:
        n1, _, el1 = timed(budget_s / 4)
    finally:
        O.lib().or_set_num_threads(cores)
    return {"value": cells * n / el, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP32 restatement of the reference kernels on the 2000x1000 P-SV workload: "
                      f"{n} shot-steps ({shot + 1} shot(s) in order, nt=4000 each) in {el:.1f} s",
            "single_core": {"value": cells * n1 / el1, "unit": "cell-updates/s", "cores": 1,
                            "sample": f"the same, {n1} shot-steps in {el1:.1f} s on one thread"}}


def gradient_bench(seiswg, synth, nt_override=0):
    """Checkpointed P-SV adjoint gradient (1000x600, 16 shots, 200 receivers,
    checkpoint every 50 steps, K_rho / K_mu / K_lambda Gaussian-smoothed
    with sigma = 3 cells): shots/s with all shots batched per launch vs the
    plain one-shot-per-launch path on the same GPU, same kernels, each timed
    from the call to the smoothed kernels on the host; the misfits are
    compared bit for bit and the kernels by rel-L2 (the batched adjoint sums
    the shots of a workgroup each step, the plain one shot by shot)."""
    c = synth.PSV_GRADIENT
    init, true, geom = synth.gradient_problem()
    nt = nt_override or c["nt"]

    def ctx_for(model):
        x = seiswg.Context(c["nx"], c["nz"], nt, c["h"], c["h"], c["dt"], sh=False, psv=True, abs_left=True,
                           abs_right=True, abs_bottom=True, abs_top=False, abs_width=c["abs_width"], abs_alpha=0.015)
        x.set_model(seiswg.abi.MODEL_VPVS, [model["rho"], model["vp"], model["vs"]])
        rows = geom.src_rows()
        x.set_sources([seiswg.grid_index(r[0], r[1], c["h"], c["h"], c["nz"]) for r in rows],
                      seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in rows], nt, c["dt"]))
        x.set_receivers([seiswg.grid_index(a, b, c["h"], c["h"], c["nz"]) for a, b in zip(geom.rec_x, geom.rec_z)])
        return x

    import numpy as np
    shots = list(range(c["nshots"]))
    ct = ctx_for(true)
    ct.forward_batch(shots)
    obs = np.stack([np.stack([ct.obs(s, seiswg.abi.OBS_X), ct.obs(s, seiswg.abi.OBS_Z)]) for s in shots])
    ct.close()
    ci = ctx_for(init)
    sig = c["smooth"]
    ci.gradient(shots[:1], obs[:1], ckpt_every=c["ckpt"], smooth=sig)   # warm-up (allocations, code load)
    t0 = time.perf_counter()
    mis_b, kr_b, km_b, kl_b = ci.gradient(shots, obs, ckpt_every=c["ckpt"], smooth=sig)
    t_batched = time.perf_counter() - t0
    t0 = time.perf_counter()
    mis_p, kr_p, km_p, kl_p = ci.gradient(shots, obs, ckpt_every=c["ckpt"], shots_per_launch=1, smooth=sig)
    t_plain = time.perf_counter() - t0
    # kernels: the batched launches sum the shots of a workgroup each step, the
    # plain ones shot by shot, so the two agree to FP32 rounding
    return {"workload": f"checkpointed P-SV adjoint gradient, 1000x600, 16 shots, 200 receivers, nt={nt}, "
                        f"checkpoint every {c['ckpt']}, kernels smoothed with sigma={sig}",
            "shots_per_s_batched": len(shots) / t_batched,
            "shots_per_s_plain": len(shots) / t_plain, "speedup": t_plain / t_batched,
            "s_batched": t_batched, "s_plain": t_plain,
            "misfits_bit_identical": bool(np.array_equal(mis_b, mis_p)),
            "kernels_max_rel_l2_vs_plain": max(rel_l2(a, b) for a, b in ((kr_b, kr_p), (km_b, km_p), (kl_b, kl_p))),
            "misfit_total": float(mis_b.sum())}


def rel_l2(a, b) -> float:
    import numpy as np
    a, b = a.astype(np.float64), b.astype(np.float64)
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / d if d > 0 else float(np.linalg.norm(a))


def small_config_bench(seiswg, synth, which, with_cpu=True):
    """One forward run of the plain SH configuration (200x200, 2000
    steps, launch-bound) or the P-SV + spin configuration (500x500,
    3000 steps, with a stable micro-inertia, see synth.spin_forward_problem): GPU
    wall time of the whole time loop after a warm-up run, and the oracle's
    FP32 CPU restatement of the same run on the host cores."""
    abi = seiswg.abi
    spin = which == "spin"
    q = synth.spin_forward_problem() if spin else synth.sh_forward_problem()
    nx, nz, h, nt, dt, g = q["nx"], q["nz"], q["h"], q["nt"], q["dt"], q["geom"]
    ctx = seiswg.Context(nx, nz, nt, h, h, dt, sh=not spin, psv=spin, spin=spin, abs_left=True, abs_right=True,
                         abs_bottom=True, abs_top=True, abs_width=q["abs_width"], abs_alpha=0.015)
    m = q["model"]
    if spin:
        ctx.set_model(abi.MODEL_LAME, [m[k] for k in ("rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c")])
    else:
        ctx.set_model(abi.MODEL_VPVS, [m["rho"], m["vp"], m["vs"]])
    rows = g.src_rows()
    src_idx = [seiswg.grid_index(r[0], r[1], h, h, nz) for r in rows]
    rec_idx = [seiswg.grid_index(x, z, h, h, nz) for x, z in zip(g.rec_x, g.rec_z)]
    ctx.set_sources(src_idx, seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in rows], nt, dt))
    ctx.set_receivers(rec_idx)
    ctx.forward_batch([-1])
    t0 = time.perf_counter()
    ctx.forward_batch([-1])
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    tiling, tiling2 = list(ctx.tiling()), list(ctx.tiling2())
    ctx.close()
    cells = nx * nz * nt
    out = {"workload": ("plain SH forward, 200x200, nt=2000, 25 receivers, taper on all edges" if not spin else
                        "P-SV + micropolar spin forward, 500x500, nt=3000, j=1e4, 7 receivers"),
           "cell_updates_per_s": cells / wall, "s_total": wall, "us_per_step": wall / nt * 1e6,
           "kernel_loop_ms": tm.loop_ms, "launches": int(tm.step_launches), "tiling": tiling,
           "tiling_multi_step": tiling2}
    if with_cpu:
        from oracle import oracle as O
        p = O.Problem(nx=nx, nz=nz, nt=nt, dx=h, dz=h, dt=dt, sh=not spin, psv=spin, spin=spin, abs_left=True,
                      abs_right=True, abs_bottom=True, abs_top=True, abs_width=q["abs_width"], abs_alpha=0.015,
                      model=m, src_rows=rows, src_idx=src_idx, rec_idx=rec_idx)
        o = O.Oracle(p, 4)
        t0 = time.perf_counter()
        o.forward(-1, 0)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cells / el, "unit": "cell-updates/s", "cores": int(O.lib().or_num_threads()),
                               "kind": "port", "sample": f"oracle FP32 restatement, the full run ({nt} steps) in "
                                                         f"{el:.2f} s"}
    return out


def uneven_bench(seiswg, synth, nt_scale=0.04, budget_gb=200.0, spec=None, only=None, tune=False, workers=0):
    """Uneven batch of 48 independent SH / P-SV gradient problems (grids
    384..3072 x 384..1536, 4-16 shots, 100-300 receivers, nt scaled from
    1500..4000 by nt_scale, checkpoint every 50; `spec` = synth.UNEVEN_SMALL: grids
    128..448, 2-6 shots; `only`: those problem indices): total shots/s through
    sw_multi_gradient (problems below its merge limit share merged launches,
    the others run on their own launches; in waves that fit budget_gb of
    device memory) vs one problem per launch (sw_gradient per problem, its
    shots batched).  tune: tile shapes autotuned once per size class (kind
    and shot-cells to the nearest power of two) on the class's first problem
    and set on every problem of the class, for both paths alike."""
    import math
    import numpy as np
    abi = seiswg.abi
    spec = spec or synth.UNEVEN
    tag = ("uneven batch" if spec is synth.UNEVEN else "small uneven batch") + (" subset" if only else "")

    def progress(msg):
        print(f"[{tag}] {msg}", file=sys.stderr, flush=True)

    probs = synth.uneven_problems(nt_scale, spec=spec, only=set(only) if only else None)
    progress(f"{len(probs)} problems generated")

    def make_ctx(p, model):
        sh = p["kind"] == "sh"
        x = seiswg.Context(p["nx"], p["nz"], p["nt"], p["h"], p["h"], p["dt"], sh=sh, psv=not sh, abs_left=True,
                           abs_right=True, abs_bottom=True, abs_top=False, abs_width=spec["abs_width"],
                           abs_alpha=0.015)
        x.set_model(abi.MODEL_VPVS, [model["rho"], model["vp"], model["vs"]])
        g = p["geom"]
        rows = g.src_rows()
        x.set_sources([seiswg.grid_index(r[0], r[1], p["h"], p["h"], p["nz"]) for r in rows],
                      seiswg.ricker_table([[r[3], r[4], r[5], r[6]] for r in rows], p["nt"], p["dt"]))
        x.set_receivers([seiswg.grid_index(a, b, p["h"], p["h"], p["nz"]) for a, b in zip(g.rec_x, g.rec_z)])
        return x

    comps = {"sh": [abi.OBS_Y], "psv": [abi.OBS_X, abi.OBS_Z]}
    obs, shots = [], []
    for p in probs:
        n = len(p["geom"].src_x)
        c = make_ctx(p, p["true"])
        c.forward_batch(list(range(n)))
        obs.append(np.stack([np.stack([c.obs(s, k) for k in comps[p["kind"]]]) for s in range(n)]))
        shots.append(list(range(n)))
        c.close()
    progress("observed data computed")

    shapes, classes, t_tune = {}, {}, 0.0
    if tune:
        for q, p in enumerate(probs):
            key = (p["kind"], int(round(math.log2(len(shots[q]) * p["nx"] * p["nz"]))))
            classes.setdefault(key, []).append(q)
        t0 = time.perf_counter()
        for key, qs in classes.items():
            p = probs[qs[0]]
            c = make_ctx(p, p["init"])
            c.autotune(len(shots[qs[0]]), steps=8)
            c.prepare(shots[qs[0]])
            shape = (c.tiling(), c.tiling2())
            c.close()
            for q in qs:
                shapes[q] = shape
        t_tune = time.perf_counter() - t0
        progress(f"{len(classes)} size classes tuned in {t_tune:.1f} s")

    def tuned_ctx(q):
        # the class's multi-step shape where the forward runs multi-step
        # launches (the one-step shape, which then serves the adjoint, stays
        # the library's choice per problem), else its one-step shape
        c = make_ctx(probs[q], probs[q]["init"])
        if q in shapes:
            t1, t2 = shapes[q]
            if t2[0]:
                c.set_tiling2(*t2)
            else:
                c.set_tiling(*t1)
        return c

    def mem_bytes(p, n):
        nslot = 3 if p["kind"] == "sh" else 5
        npt = p["nx"] * p["nz"]
        nck = -(-p["nt"] // synth.UNEVEN["ckpt"])
        return 4.0 * npt * n * (4 * nslot + (nck + 1) * nslot + synth.UNEVEN["ckpt"] * nslot + 3) + 4.0 * n * 5 * 300 * p["nt"]

    total_shots = sum(len(x) for x in shots)
    t_plain, ref0 = 0.0, None
    for q, (p, o, sh) in enumerate(zip(probs, obs, shots)):
        c = tuned_ctx(q)
        t0 = time.perf_counter()
        r = c.gradient(sh, o, ckpt_every=synth.UNEVEN["ckpt"])
        t_plain += time.perf_counter() - t0
        if q == 0:
            ref0 = r
        c.close()
        if q % 8 == 7:
            progress(f"plain: {q + 1} problems, {t_plain:.1f} s")
    waves, cur, used = [], [], 0.0
    for q, p in enumerate(probs):
        b = mem_bytes(p, len(shots[q]))
        if cur and used + b > budget_gb * 1e9:
            waves.append(cur)
            cur, used = [], 0.0
        cur.append(q)
        used += b
    waves.append(cur)
    t_batched, got0 = 0.0, None
    for w in waves:
        ctxs = [tuned_ctx(q) for q in w]
        m = seiswg.Multi(ctxs)
        if workers:
            m.set_workers(workers)
        t0 = time.perf_counter()
        res = m.gradient([shots[q] for q in w], [obs[q] for q in w], ckpt_every=synth.UNEVEN["ckpt"])
        t_batched += time.perf_counter() - t0
        if w[0] == 0:
            got0 = res[0]
        m.close()
        for c in ctxs:
            c.close()
        progress(f"batched: wave of {len(w)} problems, {t_batched:.1f} s so far")
    same_misfit = bool(np.array_equal(got0[0], ref0[0]))
    kernel_err = max(rel_l2(a, b) for a, b in zip(got0[1:], ref0[1:]))
    merged = sum(1 for p, sh in zip(probs, shots) if len(sh) * p["nx"] * p["nz"] < DEFAULT_MERGE_LIMIT)
    nts = [p["nt"] for p in probs]
    out = {"workload": f"{tag} of {len(probs)} SH/P-SV gradient problems (grids {spec['nx'][0]}..{spec['nx'][1]} x "
                       f"{spec['nz'][0]}..{spec['nz'][1]}, {spec['shots'][0]}-{spec['shots'][1]} shots, "
                       f"nt = {nt_scale} x 1500..4000, checkpoint every {synth.UNEVEN['ckpt']})",
           "problems": len(probs), "shots": total_shots, "waves": len(waves), "nt_range": [min(nts), max(nts)],
           "cell_updates_forward": int(sum(len(sh) * p["nx"] * p["nz"] * p["nt"] for p, sh in zip(probs, shots))),
           "problems_merged": merged, "problems_own_launches": len(probs) - merged,
           "shots_per_s_batched": total_shots / t_batched, "shots_per_s_plain": total_shots / t_plain,
           "speedup": t_plain / t_batched, "s_batched": t_batched, "s_plain": t_plain,
           "problem0_misfits_bit_identical": same_misfit, "problem0_kernels_rel_l2_vs_plain": kernel_err}
    if only:
        out["problem_indices"] = [p["index"] for p in probs]
        out["grids"] = [[p["kind"], p["nx"], p["nz"], len(sh), p["nt"]] for p, sh in zip(probs, shots)]
    if tune:
        out["autotune"] = {"size_classes": len(classes), "seconds": t_tune,
                           "shapes": {f"{k[0]}~2^{k[1]} shot-cells": [list(shapes[qs[0]][0]), list(shapes[qs[0]][1])]
                                      for k, qs in classes.items()}}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--threads", type=int, default=0, help="tile width (64/128/256); 0 = library default")
    ap.add_argument("--seglen", type=int, default=0)
    ap.add_argument("--group", type=int, default=0, help="shots per workgroup (1, 2, 4); 0 = library default")
    ap.add_argument("--tiling2", type=str, default="",
                    help="shape of the two-step launches as threads,seglen,shots_per_block (no autotuning)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="measure with the heuristic tile shapes instead of autotuned ones")
    ap.add_argument("--steps-per-launch", type=int, default=0,
                    help="P-SV forward steps per launch (1, 2 or 3); 0 = library default (2)")
    ap.add_argument("--reciprocal-division", action="store_true",
                    help="forward-step divisions as multiplications by the rounded reciprocal (not bit-exact)")
    ap.add_argument("--no-reciprocal-line", action="store_true",
                    help="skip the second measurement with reciprocal-multiplication divisions")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gradient", action="store_true", help="skip the batched-vs-plain gradient measurement")
    ap.add_argument("--gradient-nt", type=int, default=0, help="time steps of the gradient workload (0: 3000)")
    ap.add_argument("--gradient-only", action="store_true",
                    help="run only the batched-vs-plain gradient measurement (for profiler runs)")
    ap.add_argument("--uneven", action="store_true", help="also measure the uneven batch of 48 gradient problems")
    ap.add_argument("--no-uneven-subset", action="store_true",
                    help="skip the seeded 8-problem subset of the uneven batch at its full nt")
    ap.add_argument("--uneven-subset-only", action="store_true",
                    help="run only the seeded 8-problem subset of the uneven batch at its full nt")
    ap.add_argument("--uneven-full-only", action="store_true",
                    help="run only the whole uneven batch (48 gradient problems) at its full nt, tiles autotuned "
                         "per size class (about 3 minutes)")
    ap.add_argument("--no-small", action="store_true", help="skip the 200x200 SH and 500x500 spin forward runs")
    ap.add_argument("--small-only", action="store_true",
                    help="only the 200x200 SH, 500x500 spin and single-shot 2000x1000 forward runs (for profiling)")
    ap.add_argument("--uneven-nt-scale", type=float, default=0.04, help="nt of the uneven batch, as a fraction")
    ap.add_argument("--uneven-budget-gb", type=float, default=200.0,
                    help="device memory per wave of the uneven batch's multi-problem calls")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 1 GiB copy kernel (calibrates the HBM counters of a profiler run)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the time loop shards by shots with "
              f"no per-step exchange, so N GPUs would run N independent copies).", file=sys.stderr)
        return 2
    import seiswg_amd as seiswg
    from seiswg_amd import synth
    if a.calibrate_copy:
        seiswg.bench_copy(1 << 30, 5)
    if a.small_only:
        c = synth.PSV_BATCH
        model, geom = synth.psv_batch_problem()
        src_idx = [seiswg.grid_index(r[0], r[1], c["h"], c["h"], c["nz"]) for r in geom.src_rows()]
        rec_idx = [seiswg.grid_index(x, z, c["h"], c["h"], c["nz"]) for x, z in zip(geom.rec_x, geom.rec_z)]
        print(json.dumps({"single_shot_2000x1000": single_shot_bench(seiswg, synth, model, geom, src_idx, rec_idx),
                          "sh_forward_200x200": small_config_bench(seiswg, synth, "sh", not a.no_cpu_baseline),
                          "spin_forward_500x500": small_config_bench(seiswg, synth, "spin", not a.no_cpu_baseline)}),
              flush=True)
        return 0
    if a.gradient_only:
        print(json.dumps({"gradient": gradient_bench(seiswg, synth, a.gradient_nt)}), flush=True)
        return 0
    if a.uneven_full_only:
        print(json.dumps({"uneven_batch_full_nt": uneven_bench(seiswg, synth, 1.0, a.uneven_budget_gb, tune=True)}),
              flush=True)
        return 0
    if a.uneven_subset_only:
        print(json.dumps({"uneven_subset_full_nt": uneven_bench(seiswg, synth, 1.0, a.uneven_budget_gb,
                                                                only=synth.uneven_subset(), tune=True)}), flush=True)
        return 0
    nt = max(synth.PSV_BATCH["nt"], a.warmup + a.steps)
    ctx, model, geom, src_idx, rec_idx = build_context(seiswg, synth, nt)
    nshots = synth.PSV_BATCH["nshots"]
    if a.steps_per_launch:
        ctx.set_steps_per_launch(a.steps_per_launch)
    if a.reciprocal_division:
        ctx.set_division(True)
    ctx.prepare(list(range(nshots)))
    autotune = None
    if a.threads or a.seglen or a.group or a.tiling2:
        if a.threads or a.seglen or a.group:
            t, s, g = ctx.tiling()
            ctx.set_tiling(a.threads or t, a.seglen or s, a.group or g)
        if a.tiling2:
            ctx.set_tiling2(*[int(v) for v in a.tiling2.split(",")])
        ctx.prepare(list(range(nshots)))
    elif not a.no_autotune:
        t0 = time.perf_counter()
        _, _, _, tuned_ms = ctx.autotune(nshots, steps=8)
        autotune = {"seconds": time.perf_counter() - t0, "ms_per_step_while_tuning": tuned_ms}
        ctx.prepare(list(range(nshots)))
    # capture the 16-step graph blocks of both step parities (the timed steps
    # start at the parity the warm-up leaves) outside the timed region; the
    # batch prepared again keeps its buffers, so the blocks are replayed
    ctx.advance(16)
    ctx.advance(1)
    ctx.advance(16)
    ctx.prepare(list(range(nshots)))
    if a.warmup:
        ctx.advance(a.warmup)
    t0 = time.perf_counter()
    ctx.advance(a.steps)          # returns after the stream has drained
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    gstate, gblocks, gmsg = ctx.graph_state()
    cells = synth.PSV_BATCH["nx"] * synth.PSV_BATCH["nz"]
    cu = cells * nshots * a.steps
    avg_launch_s = tm.loop_ms * 1e-3 / max(tm.step_launches, 1)
    steps_per_launch = tm.steps / max(tm.step_launches, 1)
    # bytes per shot-cell-update (5 state fields read + written, 3 model
    # fields per launch shared by the shots) times the updates of one launch
    algo_bytes = steps_per_launch * nshots * cells * 4.0 * (2 * 5 + 3.0 / nshots)
    achieved = algo_bytes / avg_launch_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "psv_2000x1000_traffic.json")
    if os.path.exists(tf):
        # measured HBM bytes per launch of the shape that ran (null when that
        # shape has not been profiled)
        prof = json.load(open(tf))
        shape = ",".join(str(v) for v in (ctx.tiling() if steps_per_launch < 1.5 else ctx.tiling2()))
        key = "one_step" if steps_per_launch < 1.5 else "two_step"
        traffic = prof.get(key, {}).get(shape, {}).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC, "value": cu / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded layered + Gaussian random-field model)",
        "config": {"workload": "batched multi-shot P-SV forward, 2000x1000 grid, dx=10 m, 64 shots, 400 receivers, "
                               "taper L/R/B 40, dt=1 ms", "nx": 2000, "nz": 1000, "shots": nshots,
                   "physics": "psv", "tiling": list(ctx.tiling()), "tiling_two_step": list(ctx.tiling2()),
                   "autotuned": autotune is not None,
                   "division": "reciprocal" if a.reciprocal_division else "exact",
                   "hip_graphs": {"in_use": gstate == 1, "captured_blocks": gblocks,
                                  "steps_in_graph_blocks": a.steps // 16 * 16 if gstate == 1 else 0,
                                  "error": gmsg or None},
                   "parallelism": "1 GPU, shots batched per launch"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBPS, "traffic": traffic,
                     "kernel_avg_ms": avg_launch_s * 1e3, "algorithmic_bytes_per_launch": algo_bytes,
                     "kernel_avg_covers": "the step launch and, for two-step launches of many shot-cells, the "
                                          "receiver kernel after it (HIP events around the timed launches)",
                     "steps_per_launch": steps_per_launch,
                     "hbm_rate_GBps": (traffic / avg_launch_s / 1e9) if traffic else None},
    }
    if autotune is not None:
        out["autotune"] = autotune
    if not a.reciprocal_division and not a.no_reciprocal_line:
        # the same workload and tiles with the forward divisions done as
        # multiplications by the rounded reciprocal (sw_set_division; within
        # 1.5 units per quotient, not bit-exact): reported beside `value`
        ctx.set_division(True)
        ctx.prepare(list(range(nshots)))
        ctx.advance(16)
        ctx.advance(1)
        ctx.advance(16)
        ctx.prepare(list(range(nshots)))
        if a.warmup:
            ctx.advance(a.warmup)
        t0 = time.perf_counter()
        ctx.advance(a.steps)
        wall_r = time.perf_counter() - t0
        tr = ctx.timing()
        launch_r = tr.loop_ms * 1e-3 / max(tr.step_launches, 1)
        out["reciprocal_division"] = {
            "value": cu / wall_r, "unit": "cell-updates/s", "ms_per_step": wall_r / a.steps * 1e3,
            "kernel_avg_ms": launch_r * 1e3, "achieved_GBps": algo_bytes / launch_r / 1e9,
            "frac": algo_bytes / launch_r / 1e9 / PEAK_HBM_GBPS,
            "parity": "traces within rel-L2 1.1e-6 of the FP32 oracle over 1200-2000 steps (exact division: 0); "
                      "1.26e-6 from the reference's forward fixture (exact division: 1.23e-6)",
            "note": "achieved counts the algorithmic bytes of every step a launch advances; a two-step launch "
                    "moves about one step's compulsory bytes (roofline.traffic), so this rate can pass the HBM "
                    "peak"}
        ctx.set_division(False)
    ctx.close()
    if not a.no_gradient:
        out["gradient"] = gradient_bench(seiswg, synth, a.gradient_nt)
    if not a.no_small:
        out["single_shot_2000x1000"] = single_shot_bench(seiswg, synth, model, geom, src_idx, rec_idx)
        out["sh_forward_200x200"] = small_config_bench(seiswg, synth, "sh", not a.no_cpu_baseline)
        out["spin_forward_500x500"] = small_config_bench(seiswg, synth, "spin", not a.no_cpu_baseline)
    if not a.no_uneven_subset:
        out["uneven_subset_full_nt"] = uneven_bench(seiswg, synth, 1.0, a.uneven_budget_gb,
                                                    only=synth.uneven_subset(), tune=True)
    if a.uneven:
        out["uneven_batch"] = uneven_bench(seiswg, synth, a.uneven_nt_scale, a.uneven_budget_gb)
        out["small_uneven_batch"] = uneven_bench(seiswg, synth, 1.0, spec=synth.UNEVEN_SMALL)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(synth, model, geom, src_idx, rec_idx)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
