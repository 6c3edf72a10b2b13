/* Host half of the seiswg drop-in: the `seiswg <config.ini>` workflow above
 * the device ABI of seiswg_hip.h, exported from the same libseiswg_hip.so.
 *
 * It restates, in C++, the reference's host code: config.rs (INI keys,
 * defaults, path resolution), main.rs (CLI and workflow dispatch) and the
 * host parts of solver.rs (model/table/NPY file formats, grid inference,
 * grid-index rounding, Ricker table, forward and adjoint drivers, output
 * writers).  The error strings of each failure class follow the reference.
 */
#ifndef SEISWG_HOST_H
#define SEISWG_HOST_H

#include <stdint.h>

#include "seiswg_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* `seiswg <config.ini>` (main.rs:33-100): returns 0 on success; on failure a
 * non-zero code with the message in sw_last_error(). */
SW_API int seiswg_run(const char* config_path);

/* Ricker source table (solver.rs:350-356, 721-739).  params[nsrc][4] =
 * (f0, t0, ang_deg, amp); out[3][nsrc][nt] = (x, y, z) components. */
SW_API int seiswg_ricker_table(int32_t nsrc, int32_t nt, float dt, const float* params, float* out);

/* Grid inference from the coordinate files (solver.rs:640-658).  When the
 * reference's rounding rule gives nx*nz != npt, the exact shape is taken from
 * the x coordinates instead and *fallback is set to 1. */
SW_API int seiswg_infer_grid(int64_t npt, const float* x, const float* z, int32_t* nx, int32_t* nz, float* dx,
                             float* dz, int32_t* fallback);

/* Grid index of a point: round(x/dx)*nz + round(z/dz), f32 division,
 * rounding half away from zero, saturating u32 casts, wrapping u32 product
 * (solver.rs:716-718, 748-750). */
SW_API uint32_t seiswg_grid_index(float x, float z, float dx, float dz, uint32_t nz);

/* In-memory workflow: lib.rs:74-180 (`run_simulation`, the reference's
 * embedder entry point) over the Vfs seam of solver.rs:15, 236-266.
 *
 * sw_vfs is a map from path strings to byte strings.  seiswg_vfs_key lists
 * the keys in sorted order; pointers from seiswg_vfs_get and seiswg_vfs_key
 * stay valid until the map is changed or freed. */
typedef struct sw_vfs sw_vfs;
SW_API sw_vfs* seiswg_vfs_new(void);
SW_API void seiswg_vfs_free(sw_vfs* v);
/* copies nbytes from data under key (replacing an entry of that key) */
SW_API int seiswg_vfs_put(sw_vfs* v, const char* key, const void* data, int64_t nbytes);
/* size of the entry and its bytes in *data (if data is not NULL); -1 if absent */
SW_API int64_t seiswg_vfs_get(const sw_vfs* v, const char* key, const void** data);
SW_API int32_t seiswg_vfs_count(const sw_vfs* v);
SW_API const char* seiswg_vfs_key(const sw_vfs* v, int32_t i);

/* Runs the workflow of `config_ini` (INI text; relative paths stay relative,
 * as lib.rs:103 loads it with an empty base) without touching the
 * filesystem.  Inputs: the model files in `files`, keyed by the path the
 * config resolves them to (e.g. "model/proc000000_rho.bin"; a key without a
 * leading "./" also matches); sources_dat / stations_dat, if not NULL, are the
 * text of the tables the config names (else those come from `files` too);
 * for the adjoint (and inversion) workflow, the observed traces in `files`,
 * or, if config_true_ini is not NULL, a forward run of that config first,
 * whose traces join the inputs (lib.rs:106-113).  Outputs go into `out`, keyed as
 * the reference's Solver returns them: each trace file by its path
 * ("output/traces/uy_000000.npy" in .npy format, the raw "vy_000000.npy"
 * etc.); "snapshot/nframes" (u32: snapshot steps over all runs) and
 * "snapshot/<vy|vx|vz|ry>.f32" (the frames, run after run, step after step)
 * when snapshots are on (solver.rs:1104-1117); "output/proc000000_kmu.bin"
 * and "output/gradient_kmu.bin" for the adjoint; the iter_NNN model files and
 * misfit.txt for the inversion; and the scalars of lib.rs's result object as
 * 4-byte little-endian entries "nt", "nrec", "npt", "nx", "nz" (u32) and
 * "dt" (f32).  Returns 0, or non-zero with the message in sw_last_error(). */
SW_API int seiswg_run_simulation(const char* config_ini, const char* sources_dat, const char* stations_dat,
                                 const char* config_true_ini, const sw_vfs* files, sw_vfs* out);

#ifdef __cplusplus
}
#endif
#endif
