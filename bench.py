"""Benchmark of the fused gfx950 time step on the batched multi-shot P-SV
workload (64 shots sharing one 2000x1000 layered-plus-random-field model,
400 receivers, taper left/right/bottom width 40, dt = 1 ms).

A step = one time step of all 64 shots = ONE launch of the fused P-SV kernel
(velocity + stress update, source injection and receiver recording fused).
`value` is shot-cell-updates per second over the K timed steps (inputs
resident in HBM).  `roofline.achieved` is the algorithmic HBM traffic of one
launch, 64 * nx * nz * 4 * (2*5 + 3/64) bytes (5 state fields read+written
per shot, 3 model fields read once per launch), divided by the kernel's
average duration from HIP events around the timed launches.

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


def build_context(seiswg, synth, nt):
    c = synth.C2
    model, geom = synth.c2_problem()
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


def cpu_baseline(synth, model, geom, src_idx, rec_idx, budget_s=12.0):
    """The oracle's FP32 restatement of the reference scheme (unfused, one
    loop per WGSL kernel, OpenMP over cells), one shot of the same workload,
    timed for about budget_s seconds."""
    from oracle import oracle as O
    c = synth.C2
    p = O.Problem(nx=c["nx"], nz=c["nz"], nt=4000, dx=c["h"], dz=c["h"], dt=c["dt"], sh=False, psv=True,
                  spin=False, abs_left=True, abs_right=True, abs_bottom=True, abs_top=False,
                  abs_width=c["abs_width"], abs_alpha=0.015, model=model, src_rows=geom.src_rows(),
                  src_idx=src_idx, rec_idx=rec_idx)
    o = O.Oracle(p, 4)
    o.clear()
    o.step(0, 0)     # first-touch
    n, t0 = 0, time.perf_counter()
    while True:
        o.step(1 + n, 0)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 3000:
            break
    cells = c["nx"] * c["nz"]
    return {"value": cells * n / el, "unit": "cell-updates/s", "cores": int(O.lib().or_num_threads()),
            "kind": "port",
            "sample": f"oracle FP32 restatement of the reference kernels, 1 shot of the 2000x1000 P-SV workload, "
                      f"{n} time steps in {el:.1f} s"}


def gradient_bench(seiswg, synth, nt_override=0):
    """Checkpointed P-SV adjoint gradient (1000x600, 16 shots, 200 receivers,
    checkpoint every 50 steps): shots/s with all shots batched per launch vs
    the plain one-shot-per-launch path on the same GPU, same kernels, same
    results (bit-identical kernels are checked)."""
    c = synth.C3
    init, true, geom = synth.c3_problem()
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
    ci.gradient(shots[:1], obs[:1], ckpt_every=c["ckpt"])          # warm-up (allocations, code load)
    t0 = time.perf_counter()
    mis_b, kr_b, km_b, kl_b = ci.gradient(shots, obs, ckpt_every=c["ckpt"])
    t_batched = time.perf_counter() - t0
    t0 = time.perf_counter()
    mis_p, kr_p, km_p, kl_p = ci.gradient(shots, obs, ckpt_every=c["ckpt"], shots_per_launch=1)
    t_plain = time.perf_counter() - t0
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
               for a, b in ((kr_b, kr_p), (km_b, km_p), (kl_b, kl_p)))
    return {"workload": f"checkpointed P-SV adjoint gradient, 1000x600, 16 shots, 200 receivers, nt={nt}, "
                        f"checkpoint every {c['ckpt']}", "shots_per_s_batched": len(shots) / t_batched,
            "shots_per_s_plain": len(shots) / t_plain, "speedup": t_plain / t_batched,
            "s_batched": t_batched, "s_plain": t_plain, "kernels_bit_identical": bool(same),
            "misfit_total": float(mis_b.sum())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--threads", type=int, default=0, help="tile width (64/128/256); 0 = library default")
    ap.add_argument("--seglen", type=int, default=0)
    ap.add_argument("--group", type=int, default=0, help="shots per workgroup (1, 2, 4); 0 = library default")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gradient", action="store_true", help="skip the batched-vs-plain gradient measurement")
    ap.add_argument("--gradient-nt", type=int, default=0, help="time steps of the gradient workload (0: 3000)")
    ap.add_argument("--gradient-only", action="store_true",
                    help="run only the batched-vs-plain gradient measurement (for profiler runs)")
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
    if a.gradient_only:
        print(json.dumps({"gradient": gradient_bench(seiswg, synth, a.gradient_nt)}), flush=True)
        return 0
    nt = max(synth.C2["nt"], a.warmup + a.steps)
    ctx, model, geom, src_idx, rec_idx = build_context(seiswg, synth, nt)
    nshots = synth.C2["nshots"]
    ctx.prepare(list(range(nshots)))
    if a.threads or a.seglen or a.group:
        t, s, g = ctx.tiling()
        ctx.set_tiling(a.threads or t, a.seglen or s, a.group or g)
        ctx.prepare(list(range(nshots)))
    if a.warmup:
        ctx.advance(a.warmup)
    t0 = time.perf_counter()
    ctx.advance(a.steps)          # returns after the stream has drained
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    cells = synth.C2["nx"] * synth.C2["nz"]
    cu = cells * nshots * a.steps
    avg_launch_s = tm.loop_ms * 1e-3 / max(tm.step_launches, 1)
    algo_bytes = nshots * cells * 4.0 * (2 * 5 + 3.0 / nshots)
    achieved = algo_bytes / avg_launch_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "psv_c2_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC, "value": cu / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded layered + Gaussian random-field model)",
        "config": {"workload": "batched multi-shot P-SV forward, 2000x1000 grid, dx=10 m, 64 shots, 400 receivers, "
                               "taper L/R/B 40, dt=1 ms", "nx": 2000, "nz": 1000, "shots": nshots,
                   "physics": "psv", "tiling": list(ctx.tiling()), "parallelism": "1 GPU, shots batched per launch"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBPS, "traffic": traffic,
                     "kernel_avg_ms": avg_launch_s * 1e3, "algorithmic_bytes_per_launch": algo_bytes},
    }
    ctx.close()
    if not a.no_gradient:
        out["gradient"] = gradient_bench(seiswg, synth, a.gradient_nt)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(synth, model, geom, src_idx, rec_idx)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
