/* C ABI of libseiswg_hip.so: the MI355X (gfx950) device module of the
 * seiswg staggered-grid time loop.
 *
 * It replaces the device half of the reference's src/solver.rs: wgpu
 * initialisation, buffer allocation, pipelines and dispatch (Solver::new,
 * solver.rs:578-933), the per-step dispatch lists (solver.rs:999-1030,
 * 1389-1407, 1442-1498), clears/readbacks/uploads (solver.rs:1126-1242) and
 * the smoothing driver (solver.rs:1249-1274).  The host half (config, file
 * formats, workflow drivers) sits above this ABI; see seiswg_host.h.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on error;
 *     sw_last_error() then returns a thread-local message.  -1: bad
 *     arguments, -2: invalid request for this context, -3: HIP error,
 *     -4: a snapshot callback failed, -5: the wavefield diverged (NaN or
 *     infinite values; see sw_set_finite_check).
 *   - Host pointers are borrowed for the duration of the call.  Device memory
 *     belongs to the context.
 *   - One context = one HIP device + one HIP stream; a context must not be
 *     used from two threads at once.  There are no internal host threads.
 *   - All field data are FP32, grid-major with k = i*nz + j (z contiguous),
 *     exactly the reference layout (fd2d.wgsl:140-145).
 */
#ifndef SEISWG_HIP_H
#define SEISWG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW_API __attribute__((visibility("default")))

typedef struct sw_ctx sw_ctx;

/* Grid and solver switches: the reference's Params uniform (fd2d.wgsl:18-49,
 * solver.rs:92-101) minus the per-dispatch fields (it, isrc). */
typedef struct sw_grid {
  int32_t nx, nz, nt;
  float dx, dz, dt;
  int32_t sh, psv, spin;
  int32_t abs_left, abs_right, abs_bottom, abs_top;
  int32_t abs_width;
  float abs_alpha;
} sw_grid;

/* Model kinds of sw_set_model (solver.rs:614-700). */
#define SW_MODEL_VPVS 0 /* fields = {rho, vp, vs}; converted by vps2lm (fd2d.wgsl:188-201) */
#define SW_MODEL_LAME 1 /* fields = {rho, lambda, mu, nu, j, lambda_c, mu_c, nu_c} (spin files) */

/* Field ids of sw_read_field / sw_write_field: the reference's F_* slice ids
 * (fd2d.wgsl:53-105).  Only the materialised ones are accepted: state fields
 * of the enabled physics, model fields, the taper and the kernels. */
#define SW_F_VY 0
#define SW_F_SXY 2
#define SW_F_SZY 3
#define SW_F_VX 7
#define SW_F_VZ 8
#define SW_F_SXX 11
#define SW_F_SZZ 12
#define SW_F_SXZ 13
#define SW_F_VY_C 20
#define SW_F_SYX_C 22
#define SW_F_SYY_C 23
#define SW_F_SYZ_C 24
#define SW_F_LAM 30
#define SW_F_MU 31
#define SW_F_NU 32
#define SW_F_J 33
#define SW_F_LAM_C 34
#define SW_F_MU_C 35
#define SW_F_NU_C 36
#define SW_F_RHO 37
#define SW_F_BOUND 38
#define SW_F_K_LAM 39
#define SW_F_K_MU 40
#define SW_F_K_RHO 41

/* Trace components (fd2d.wgsl:110-114). */
#define SW_OBS_X 0
#define SW_OBS_Y 1
#define SW_OBS_Z 2
#define SW_OBS_YI 3
#define SW_OBS_YC 4

/* Snapshot callback: called after step `it` (it > 0, it % snap_every == 0,
 * solver.rs:1035) once per field written by the reference's save_snapshot
 * (solver.rs:1144-1176: vy | vx, vz | vy_c) with a host copy of the field.
 * Calls come in snapshot order but one snapshot late, while the following
 * steps run on the GPU (the copy overlaps them); `data` is valid during the
 * call only, and the callback must not call back into the context.  A
 * non-zero return aborts the run. */
typedef int (*sw_snapshot_fn)(void* user, int32_t shot, int32_t it, int32_t field, const float* data,
                              int64_t npt);

SW_API const char* sw_last_error(void);

/* Solver::new, device part (solver.rs:578-612, 677-700, 799-809). */
SW_API int sw_create(const sw_grid* grid, int32_t hip_device, sw_ctx** out);
SW_API void sw_destroy(sw_ctx* ctx);

/* Model upload + init pass (solver.rs:677-700, 898-915): copies the fields,
 * runs vps2lm for SW_MODEL_VPVS and builds the absorbing taper (set_bound,
 * fd2d.wgsl:205-235). */
SW_API int sw_set_model(sw_ctx* ctx, int32_t kind, const float* const* fields);

/* Sources: grid indices k = ix*nz + iz (solver.rs:716-718) and the source
 * time functions [3][nsrc][nt] (x, y, z; solver.rs:721-739).  A batch
 * prepared before the call is discarded: sw_prepare must be called again
 * before sw_advance (the run entry points prepare by themselves). */
SW_API int sw_set_sources(sw_ctx* ctx, int32_t nsrc, const int32_t* grid_index, const float* stf);
/* Receivers: grid indices (solver.rs:745-751).  Discards a prepared batch
 * as sw_set_sources does. */
SW_API int sw_set_receivers(sw_ctx* ctx, int32_t nrec, const int32_t* grid_index);

/* One forward run (solver.rs:976-1046): clear wavefields and traces, run nt
 * steps with source isrc active (-1 = all sources, combine_sources).  The
 * state of every shot is checked for NaN / infinite values at each snapshot
 * step and after the last step (no counterpart in the reference, which keeps
 * stepping a diverged field): a diverged run returns -5 naming the step of
 * the first check that failed and of the last one that passed. */
SW_API int sw_forward(sw_ctx* ctx, int32_t isrc, int32_t snap_every, sw_snapshot_fn cb, void* user);
/* nshots forward runs advanced together, one launch per time step for all
 * shots (the reference runs them one after another, solver.rs:964-970).
 * Shot s fires source isrc[s] (-1 = all). */
SW_API int sw_forward_batch(sw_ctx* ctx, int32_t nshots, const int32_t* isrc, int32_t snap_every,
                            sw_snapshot_fn cb, void* user);
/* Traces of shot `shot` of the last run, component comp, [nrec][nt]
 * (readback_obs + slice_obs, solver.rs:1180-1195). */
SW_API int sw_read_obs(sw_ctx* ctx, int32_t shot, int32_t comp, float* out);
/* readback_field (solver.rs:1205-1211) / write_field_slice (1239-1242). */
SW_API int sw_read_field(sw_ctx* ctx, int32_t shot, int32_t field, float* out);
SW_API int sw_write_field(sw_ctx* ctx, int32_t shot, int32_t field, const float* in);

/* Low-level stepping for tests: prepare `nshots` shots (zeroed state and
 * traces, step clock at 0) without running them ... */
SW_API int sw_prepare(sw_ctx* ctx, int32_t nshots, const int32_t* isrc);
/* ... then advance all prepared shots by `nsteps` time steps (no blow-up
 * check: read the fields to inspect them). */
SW_API int sw_advance(sw_ctx* ctx, int32_t nsteps);

/* Divisions of the forward step kernels (no counterpart in the reference,
 * whose WGSL divisions carry the shader language's 2.5-unit error bound):
 * SW_DIV_EXACT (the default) rounds every quotient as IEEE division does,
 * so the steps equal the FP32 restatement of the reference kernels bit for
 * bit; SW_DIV_RECIPROCAL multiplies by the correctly rounded reciprocal
 * (within 1.5 units), fewer instructions per cell.  The gradient's adjoint
 * steps always use the reciprocal-and-correction form. */
#define SW_DIV_EXACT 0
#define SW_DIV_RECIPROCAL 1
SW_API int sw_set_division(sw_ctx* ctx, int32_t mode);

/* Blow-up detection (no counterpart in the reference): on by default.
 * Forward runs scan the state at snapshot steps and after the last step;
 * gradients (sw_gradient, sw_multi_gradient) check their misfits after the
 * forward pass and their kernel fields at the end; sw_adjoint_plain its
 * misfits and K_mu.  A failed check returns -5.  0 turns the checks off. */
SW_API int sw_set_finite_check(sw_ctx* ctx, int32_t on);

/* run_adjoint (solver.rs:1298-1529), SH, reference imaging condition:
 * nruns runs (run r fires isrc[r]); obs_ref = observed uy traces
 * [nruns][nrec][nt]; K_mu is accumulated over runs and then smoothed with
 * sigma = smooth.  misfit[r] receives the run's misfit (may be NULL).
 * Read the result with sw_read_field(ctx, 0, SW_F_K_MU, ...). */
SW_API int sw_adjoint_plain(sw_ctx* ctx, int32_t nruns, const int32_t* isrc, const float* obs_ref,
                            int32_t adj_interval, float smooth, double* misfit);

/* Differentiable gradient (no counterpart in the reference, whose only
 * gradient is the SH imaging kernel of run_adjoint): exact discrete adjoint of
 * the SH or P-SV scheme (spin unsupported) for the objective
 *   J = sum_shots sum_comp sum_rec sum_n (w_n (d^n - obs^n))^2,
 * w_n the misfit taper of solver.rs:441-455, d the recorded traces (SH: uy;
 * P-SV: ux, uz).  obs_ref is [nshots][ncomp][nrec][nt].  The forward is
 * checkpointed every ckpt_every steps and recomputed segment by segment.
 * shots_per_launch shots advance together (0 = all).  The gradients w.r.t.
 * the stored rho, mu, lambda are summed over shots into SW_F_K_RHO,
 * SW_F_K_MU, SW_F_K_LAM (read with sw_read_field); misfit[s] = J of shot s.
 * The adjoint keeps one set of kernel accumulators per workgroup: with
 * several shots per workgroup (sw_set_tiling) their contributions are added
 * in shot order each step, then the sets in a fixed order at the end.
 * Results are deterministic; with one shot per workgroup they are the same
 * bit for bit for any checkpoint interval and shots_per_launch, and other
 * shapes agree with them to FP32 rounding. */
SW_API int sw_gradient(sw_ctx* ctx, int32_t nshots, const int32_t* isrc, const float* obs_ref, int32_t ckpt_every,
                       int32_t shots_per_launch, double* misfit);

/* Several independent problems (each its own context: grid, model, physics,
 * sources, receivers, nt, dt, tile shape) advanced together: every time step
 * is one launch per (physics, tile width) group covering all problems' shots,
 * problems with fewer steps idling once done.  No counterpart in the
 * reference (one problem per process).  All contexts must share one device;
 * the multi object borrows them until destroyed. */
typedef struct sw_multi sw_multi;
SW_API int sw_multi_create(sw_ctx* const* problems, int32_t n, sw_multi** out);
SW_API void sw_multi_destroy(sw_multi* m);
/* Forward runs: problem q fires isrc[q][s] for its shot s (as sw_forward_batch);
 * traces are read per problem with sw_read_obs. */
SW_API int sw_multi_forward(sw_multi* m, const int32_t* nshots, const int32_t* const* isrc);
/* Gradients (as sw_gradient, all shots of a problem in one batch):
 * obs_ref[q] is [nshots[q]][ncomp][nrec][nt] of problem q; misfit[q] may be
 * NULL.  A common checkpoint interval; segments of the merged problems are
 * aligned.  Problems with at least `merge limit` shot-cells per step
 * (nshots * nx * nz; default 0, i.e. all) run on their own launches (see
 * sw_multi_set_workers); smaller ones, if a limit is set, advance together
 * in merged launches, which run one shot per workgroup: their results equal
 * sw_gradient's with one shot per workgroup bit for bit (to FP32 rounding
 * for a context whose tile puts several shots in a workgroup). */
SW_API int sw_multi_gradient(sw_multi* m, const int32_t* nshots, const int32_t* const* isrc,
                             const float* const* obs_ref, int32_t ckpt_every, double* const* misfit);
SW_API int sw_multi_set_merge_limit(sw_multi* m, int64_t shot_cells);
/* Problems on their own launches run concurrently, largest first, each from
 * its own host thread on its context's stream, with their graph capture off:
 * `workers` (default 3) problems that fill the GPU (4 M shot-cells per step
 * or more) at a time, a smaller problem counting for its share of one, on up
 * to 4 * workers threads.  1 runs them one after the other. */
SW_API int sw_multi_set_workers(sw_multi* m, int32_t workers);
/* Elapsed time of the last multi call (HIP events for merged forward runs,
 * host clock for gradients) and merged launches issued. */
SW_API int sw_multi_timing(const sw_multi* m, double* ms, int64_t* launches);

/* Gaussian smoothing of a kernel field in place (smooth_gradient,
 * solver.rs:1249-1274; init_gsum, gaussian_x, gaussian_z). */
SW_API int sw_smooth(sw_ctx* ctx, int32_t field, float sigma);
/* sw_smooth without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc).  The weight table, the 1-D weight sums
 * and the x-pass result live in one fixed device array of the library per
 * GPU (21.2 MB, held until the process ends): grids of at most 5,242,880
 * cells and 16384 points per axis; larger grids are rejected (-2).  The
 * kernel field must exist on the context (after a gradient, an adjoint run
 * or sw_write_field).  Every call on a GPU uses that GPU's array, so calls
 * that overlap on one GPU (other threads or contexts) race; the array holds
 * nothing a call must find cleared.  Results equal sw_smooth's bit for bit. */
SW_API int sw_smooth_noalloc(sw_ctx* ctx, int32_t field, float sigma);

/* Tile shape of the fused step kernels (no counterpart in the reference,
 * which dispatches fixed 64-wide 1-D workgroups): threads per shot (64, 128
 * or 256; they own threads-6 z-columns), rows marched per workgroup, and
 * shots per workgroup of the forward steps (0: automatic; 1; 2 or 4 with
 * threads >= 128; reduced to a divisor of each batch's shot count).  Without
 * a call, each batch gets a heuristic shape sized to fill the GPU.
 * sw_get_tiling reports the shape of the last prepared batch. */
SW_API int sw_set_tiling(sw_ctx* ctx, int32_t threads, int32_t seglen, int32_t shots_per_block);
SW_API int sw_get_tiling(const sw_ctx* ctx, int32_t* threads, int32_t* seglen, int32_t* shots_per_block);
/* Time steps per launch (no counterpart in the reference, which dispatches
 * every half-step separately): 2 (the default) runs the steps of SH or P-SV
 * alone, with or without spin, in launches that each advance several steps,
 * reading the state once and writing it once per launch (the shape set by
 * sw_set_tiling2); 3 runs P-SV alone (no spin) in marching launches of three
 * steps each (tiles of 256, 384 or 512 threads, one shot per workgroup; SH
 * keeps two); 1 runs one step per launch.  Results are the same bit for bit.
 * SH and P-SV together, and the gradient's adjoint, always run one step per
 * launch; its recompute runs two-step marching launches (one step per launch
 * where the batch has LDS tiles or three-step tiles).  A change takes effect
 * at the next prepared batch. */
SW_API int sw_set_steps_per_launch(sw_ctx* ctx, int32_t steps);
/* Tile shape of the multi-step launches (no counterpart in the reference):
 *  - marching tiles, two steps per launch: threads per shot 64, 128, 256,
 *    320 or 512 (they own threads-12 z-columns), `seglen` rows per
 *    workgroup, shots per workgroup (0: automatic; 1; 2 or 4 with threads
 *    >= 128 and at most 1024 threads per workgroup); not with spin;
 *  - marching tiles, three steps per launch (sw_set_steps_per_launch 3):
 *    threads per shot 256, 384 or 512 (they own threads-18 z-columns), one
 *    shot per workgroup;
 *  - LDS tiles (small grids, latency-bound): threads = 1024 and `seglen` =
 *    K, the steps per launch (2..4; with spin 2..3; 0: 4, with spin 3).
 *    Each workgroup loads a 32 x 64-cell region once, advances it K steps
 *    in LDS and stores its (32-6K) x (64-6K)-cell interior; one shot per
 *    workgroup (shots_per_block 0 or 1).
 * Without a call they get a heuristic shape (LDS tiles for batches of at
 * most 300000 shot-cells), or that of sw_set_tiling when it was called.
 * sw_get_tiling2 reports the shape of the last prepared batch (zeros when it
 * runs one step per launch; LDS tiles: 1024, K, 1). */
SW_API int sw_set_tiling2(sw_ctx* ctx, int32_t threads, int32_t seglen, int32_t shots_per_block);
SW_API int sw_get_tiling2(const sw_ctx* ctx, int32_t* threads, int32_t* seglen, int32_t* shots_per_block);
/* Autotuner: times `steps` forward steps of `nshots` shots for a set of
 * tile shapes, keeps the fastest (as sw_set_tiling does) and returns it.
 * When steps run two per launch (P-SV or SH without spin), the two-step
 * shape is tuned too and kept (as sw_set_tiling2 does; read it with
 * sw_get_tiling2); `ms` is the time per step of the configuration that runs
 * (two-step launches where they apply).  Tile shapes never change forward
 * results; a gradient's kernels depend on the adjoint's shots per workgroup
 * at FP32 rounding (see sw_gradient). */
SW_API int sw_autotune(sw_ctx* ctx, int32_t nshots, int32_t steps, int32_t* threads, int32_t* seglen,
                       int32_t* shots_per_block, double* ms);

/* Timing of the last run's time loop, from HIP events on the context stream. */
typedef struct sw_timing {
  double loop_ms;        /* events around all time steps of the last run */
  int64_t steps;         /* time steps in that loop */
  int64_t step_launches; /* fused step-kernel launches in that loop */
  int64_t cell_updates;  /* nx*nz*steps*shots */
} sw_timing;
SW_API int sw_get_timing(const sw_ctx* ctx, sw_timing* out);

/* Captured time-step blocks (no counterpart in the reference): runs of at
 * least 16 steps replay blocks of 16 steps captured once as HIP graphs,
 * keyed by their launch arguments (kept per context, at most 16).  *state:
 * 1 graphs in use, 0 off (environment SEISWG_NO_GRAPHS=1 or a concurrent
 * multi-problem run), -1 off because a capture failed (the context then
 * launches directly; the failure is printed once on stderr and its HIP
 * message copied to msg, up to msg_len bytes, if msg is not NULL).
 * *captured (may be NULL): blocks held. */
SW_API int sw_get_graph_state(const sw_ctx* ctx, int32_t* state, int32_t* captured, char* msg, int32_t msg_len);

/* Device memory of freed buffers is kept for reuse by later buffers of up to
 * the same size (no counterpart in the reference): runs over many problems
 * then do not wait on fresh allocations.  This returns the cached memory of
 * `hip_device` (-1: every device) to the runtime, e.g. before another
 * library allocates; an allocation that fails empties the cache and retries
 * by itself.  Environment SEISWG_NO_MEMORY_CACHE=1 disables the cache. */
SW_API int sw_release_cached_memory(int32_t hip_device, int64_t* bytes_freed);

/* Measurement aid with no counterpart in the reference: `iters` launches of
 * a plain float4 copy kernel moving nbytes (read) + nbytes (write) on
 * `hip_device`; average kernel time in ms from HIP events.  Used to calibrate
 * the HBM read/write counters against a kernel of known traffic. */
SW_API int sw_bench_copy(int32_t hip_device, int64_t nbytes, int32_t iters, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
