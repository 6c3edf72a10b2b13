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

#ifdef __cplusplus
}
#endif
#endif
