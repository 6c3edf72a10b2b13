// gfx950 kernels of the staggered-grid time loop.
//
// One time step of one physics is ONE launch.  Each workgroup owns a z-strip
// of W = NT-6 columns (z is the contiguous axis, k = i*nz + j, fd2d.wgsl:140)
// and marches a segment of rows along x.  x-neighbours stay in registers (a
// 4-row window per stress field, a 4-row window of new velocities), z-
// neighbours are exchanged through LDS.  The velocity update of row m and the
// stress update of row m-2 are fused in the same iteration, so every state
// value is read once and written once per step; the 3-column/3-row halos are
// recomputed, never exchanged.  Reads and writes go to different (ping-pong)
// state buffers, so tiles never race on their halos.
//
// Arithmetic follows fd2d.wgsl operation by operation (no FMA contraction,
// IEEE division), so results equal the reference kernel sequence of
// solver.rs:999-1030 evaluated in FP32.
#include "kernels.h"

#pragma clang fp contract(off)

namespace sw {
namespace {

// fd2d.wgsl:151-164 (backward) and 166-180 (forward) differences, with the
// reference's 1/8 weight on the 3-cell difference.
__device__ __forceinline__ float dminus(float fm2, float fm1, float f0, float fp1, float h8) {
  return (9.0f * (f0 - fm1) - (fp1 - fm2)) / h8;
}
__device__ __forceinline__ float dplus(float fm1, float f0, float fp1, float fp2, float h8) {
  return (9.0f * (fp1 - f0) - (fp2 - fm1)) / h8;
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Map them so that
// consecutive logical tiles (which share halos) land on the same XCD and its L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, p = b >> 3;
  return (x < r) ? x * (q + 1) + p : r * (q + 1) + (x - r) * q + p;
}

__device__ __forceinline__ float ldr(const float* __restrict__ p, int m, int nx, int nz, int j, bool jin,
                                     float dflt = 0.0f) {
  return (jin && m >= 0 && m < nx) ? p[m * nz + j] : dflt;
}

// Advance the step clock once every workgroup of this launch has read it.
__device__ __forceinline__ void advance_clock(int* clock, int it) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int done = atomicAdd(&clock[1], 1);
    if (done == int(gridDim.x) - 1) {
      atomicExch(&clock[1], 0);
      atomicExch(&clock[0], it + 1);
    }
  }
}

// ---------------------------------------------------------------- SH step
// div_sy, stf_dsy, add_vy, div_vy, add_sy_sh, save_obs_y_sh (fd2d.wgsl:242-301)
template <int NT>
__global__ __launch_bounds__(NT) void k_step_sh(const StepArgs a) {
  constexpr int W = NT - 6;
  __shared__ float lds_s[2][NT];
  __shared__ float lds_v[2][NT];

  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tps = a.nstrips * a.nsegs;
  const int shot = logical / tps;
  const int tile = logical - shot * tps;
  const int strip = tile % a.nstrips, seg = tile / a.nstrips;
  const int t = threadIdx.x;
  const int nx = a.nx, nz = a.nz;
  const int j = strip * W - 3 + t;
  const int i0 = seg * a.seglen, i1 = min(i0 + a.seglen, nx);
  const bool jin = j >= 0 && j < nz;
  const bool jst = jin && t >= 3 && t < NT - 3;
  const bool zm_ok = j >= 2 && j < nz - 2;
  const bool zp_ok = j >= 1 && j < nz - 2;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int64_t so = int64_t(shot) * a.shot_stride;
  const float* __restrict__ ivy = a.in[S_VY] + so;
  const float* __restrict__ isxy = a.in[S_SXY] + so;
  const float* __restrict__ iszy = a.in[S_SZY] + so;
  float* __restrict__ ovy = a.out[S_VY] + so;
  float* __restrict__ osxy = a.out[S_SXY] + so;
  float* __restrict__ oszy = a.out[S_SZY] + so;
  const float bzb = jin ? a.bzb[j] : 1.0f, bzt = jin ? a.bzt[j] : 1.0f;
  const float dt = a.dt, h8x = a.h8x, h8z = a.h8z;
  const int tm2 = max(t - 2, 0), tm1 = max(t - 1, 0), tp1 = min(t + 1, NT - 1), tp2 = min(t + 2, NT - 1);

  int ip = a.inj_off[shot * tps + tile];
  const int ie = a.inj_off[shot * tps + tile + 1];
  int rp = 0, re = 0;
  if (a.record) { rp = a.rec_off[tile]; re = a.rec_off[tile + 1]; }
  float* uacc = a.uacc + int64_t(shot) * a.nrec * N_UACC;
  float* obs = a.obs + int64_t(shot) * 5 * a.nrec * a.nt;

  // stress windows hold rows m-2..m+1; prefetch registers hold row m+2
  float sxy[4] = {0, 0, 0, 0}, szy[4] = {0, 0, 0, 0};
  for (int q = 1; q < 4; ++q) {
    sxy[q] = ldr(isxy, i0 - 4 + q, nx, nz, j, jin);
    szy[q] = ldr(iszy, i0 - 4 + q, nx, nz, j, jin);
  }
  float pf_sxy = ldr(isxy, i0, nx, nz, j, jin), pf_szy = ldr(iszy, i0, nx, nz, j, jin);
  float pf_vy = ldr(ivy, i0 - 1, nx, nz, j, jin);
  float pf_rho = ldr(a.rho, i0 - 1, nx, nz, j, jin, 1.0f);
  float pf_mu = ldr(a.mu, i0 - 3, nx, nz, j, jin);
  float vy[4] = {0, 0, 0, 0};        // new velocity, rows m-3..m
  float dvydz[3] = {0, 0, 0};        // d+z of new vy, rows m-2..m

  int par = 0;
  for (int m = i0 - 1; m <= i1 + 1; ++m) {
    sxy[0] = sxy[1]; sxy[1] = sxy[2]; sxy[2] = sxy[3]; sxy[3] = pf_sxy;
    szy[0] = szy[1]; szy[1] = szy[2]; szy[2] = szy[3]; szy[3] = pf_szy;
    const float vy_o = pf_vy, rho = pf_rho, mu = pf_mu;
    pf_sxy = ldr(isxy, m + 2, nx, nz, j, jin);
    pf_szy = ldr(iszy, m + 2, nx, nz, j, jin);
    pf_vy = ldr(ivy, m + 1, nx, nz, j, jin);
    pf_rho = ldr(a.rho, m + 1, nx, nz, j, jin, 1.0f);
    pf_mu = ldr(a.mu, m - 1, nx, nz, j, jin);
    const float bxm = (m >= 0 && m < nx) ? a.bx[m] : 1.0f;

    // velocity half-step at row m
    lds_s[par][t] = szy[2];
    __syncthreads();
    const bool xm_ok = m >= 2 && m < nx - 2;
    float dsy = (xm_ok ? dminus(sxy[0], sxy[1], sxy[2], sxy[3], h8x) : 0.0f) +
                (zm_ok ? dminus(lds_s[par][tm2], lds_s[par][tm1], szy[2], lds_s[par][tp1], h8z) : 0.0f);
    while (ip < ie) {
      const int4 e = a.inj[ip];
      if (e.x > m) break;
      if (e.x == m && e.y == j) dsy = dsy + a.val_y[int64_t(e.z) * a.nt_val + it];
      ++ip;
    }
    const float b = (bxm * bzb) * bzt;
    const float vyn = b * (vy_o + dt * dsy / rho);
    if (jst && m >= i0 && m < i1) ovy[m * nz + j] = vyn;
    vy[0] = vy[1]; vy[1] = vy[2]; vy[2] = vy[3]; vy[3] = vyn;
    lds_v[par][t] = vyn;
    __syncthreads();
    dvydz[0] = dvydz[1]; dvydz[1] = dvydz[2];
    dvydz[2] = zp_ok ? dplus(lds_v[par][tm1], vyn, lds_v[par][tp1], lds_v[par][tp2], h8z) : 0.0f;

    // receivers on row m: uy += vy*dt, trace = uy
    while (rp < re) {
      const int4 e = a.rec[rp];
      if (e.x > m) break;
      if (e.x == m && e.y == j) {
        float* u = uacc + e.z * N_UACC;
        const float uy = u[U_UY] + vyn * dt;
        u[U_UY] = uy;
        obs[(1 * a.nrec + e.z) * a.nt + it] = uy;
      }
      ++rp;
    }

    // stress half-step at row i = m-2
    const int i = m - 2;
    if (i >= i0) {
      const bool xp_ok = i >= 1 && i < nx - 2;
      const float mu_dt = mu * dt;
      const float dvydx = xp_ok ? dplus(vy[0], vy[1], vy[2], vy[3], h8x) : 0.0f;
      const float nsxy = sxy[0] + mu_dt * dvydx;
      const float nszy = szy[0] + mu_dt * dvydz[0];
      if (jst) { osxy[i * nz + j] = nsxy; oszy[i * nz + j] = nszy; }
    }
    par ^= 1;
  }
  if (a.advance) advance_clock(a.clock, it);
}

// ------------------------------------------------------- P-SV (+spin) step
// div_sxz, div_sy_c, div_sxyz_c, stf_dsxz, add_vxz, div_vxz, add_sxz,
// add_vy_c, div_vy_c, add_sy_c, save_obs_x/z, compute_du_grad, save_obs_ry
// (fd2d.wgsl:308-492)
template <int NT, bool SPIN>
__global__ __launch_bounds__(NT) void k_step_psv(const StepArgs a) {
  constexpr int W = NT - 6;
  constexpr int NSL = SPIN ? 4 : 2;   // sxz, szz, syz_c, syy_c
  constexpr int NVL = SPIN ? 3 : 2;   // vx, vz, vy_c
  __shared__ float lds_s[2][NSL][NT];
  __shared__ float lds_v[2][NVL][NT];

  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int tps = a.nstrips * a.nsegs;
  const int shot = logical / tps;
  const int tile = logical - shot * tps;
  const int strip = tile % a.nstrips, seg = tile / a.nstrips;
  const int t = threadIdx.x;
  const int nx = a.nx, nz = a.nz;
  const int j = strip * W - 3 + t;
  const int i0 = seg * a.seglen, i1 = min(i0 + a.seglen, nx);
  const bool jin = j >= 0 && j < nz;
  const bool jst = jin && t >= 3 && t < NT - 3;
  const bool zm_ok = j >= 2 && j < nz - 2;
  const bool zp_ok = j >= 1 && j < nz - 2;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int64_t so = int64_t(shot) * a.shot_stride;
  const float* __restrict__ ivx = a.in[S_VX] + so;
  const float* __restrict__ ivz = a.in[S_VZ] + so;
  const float* __restrict__ isxx = a.in[S_SXX] + so;
  const float* __restrict__ iszz = a.in[S_SZZ] + so;
  const float* __restrict__ isxz = a.in[S_SXZ] + so;
  float* __restrict__ ovx = a.out[S_VX] + so;
  float* __restrict__ ovz = a.out[S_VZ] + so;
  float* __restrict__ osxx = a.out[S_SXX] + so;
  float* __restrict__ oszz = a.out[S_SZZ] + so;
  float* __restrict__ osxz = a.out[S_SXZ] + so;
  const float* __restrict__ ivyc = nullptr;
  const float* __restrict__ isyx = nullptr;
  const float* __restrict__ isyy = nullptr;
  const float* __restrict__ isyz = nullptr;
  float* __restrict__ ovyc = nullptr;
  float* __restrict__ osyx = nullptr;
  float* __restrict__ osyy = nullptr;
  float* __restrict__ osyz = nullptr;
  if constexpr (SPIN) {
    ivyc = a.in[S_VYC] + so; isyx = a.in[S_SYXC] + so; isyy = a.in[S_SYYC] + so; isyz = a.in[S_SYZC] + so;
    ovyc = a.out[S_VYC] + so; osyx = a.out[S_SYXC] + so; osyy = a.out[S_SYYC] + so; osyz = a.out[S_SYZC] + so;
  }
  const float bzb = jin ? a.bzb[j] : 1.0f, bzt = jin ? a.bzt[j] : 1.0f;
  const float dt = a.dt, h8x = a.h8x, h8z = a.h8z;
  const int tm2 = max(t - 2, 0), tm1 = max(t - 1, 0), tp1 = min(t + 1, NT - 1), tp2 = min(t + 2, NT - 1);

  int ip = a.inj_off[shot * tps + tile];
  const int ie = a.inj_off[shot * tps + tile + 1];
  int rp = 0, re = 0;
  if (a.record) { rp = a.rec_off[tile]; re = a.rec_off[tile + 1]; }
  float* uacc = a.uacc + int64_t(shot) * a.nrec * N_UACC;
  float* obs = a.obs + int64_t(shot) * 5 * a.nrec * a.nt;
  const int nobs = a.nrec * a.nt;

  float sxx[4] = {0, 0, 0, 0}, sxz[4] = {0, 0, 0, 0}, szz[4] = {0, 0, 0, 0};
  float syx[4] = {0, 0, 0, 0}, syy[4] = {0, 0, 0, 0}, syz[4] = {0, 0, 0, 0};
  for (int q = 1; q < 4; ++q) {
    const int r = i0 - 4 + q;
    sxx[q] = ldr(isxx, r, nx, nz, j, jin);
    sxz[q] = ldr(isxz, r, nx, nz, j, jin);
    szz[q] = ldr(iszz, r, nx, nz, j, jin);
    if constexpr (SPIN) {
      syx[q] = ldr(isyx, r, nx, nz, j, jin);
      syy[q] = ldr(isyy, r, nx, nz, j, jin);
      syz[q] = ldr(isyz, r, nx, nz, j, jin);
    }
  }
  float pf_sxx = ldr(isxx, i0, nx, nz, j, jin), pf_sxz = ldr(isxz, i0, nx, nz, j, jin),
        pf_szz = ldr(iszz, i0, nx, nz, j, jin);
  float pf_syx = 0, pf_syy = 0, pf_syz = 0, pf_vyc = 0, pf_jj = 1.0f, pf_nu = 0, pf_muc = 0, pf_nuc = 0;
  if constexpr (SPIN) {
    pf_syx = ldr(isyx, i0, nx, nz, j, jin);
    pf_syy = ldr(isyy, i0, nx, nz, j, jin);
    pf_syz = ldr(isyz, i0, nx, nz, j, jin);
    pf_vyc = ldr(ivyc, i0 - 1, nx, nz, j, jin);
    pf_jj = ldr(a.jj, i0 - 1, nx, nz, j, jin, 1.0f);
    pf_nu = ldr(a.nu, i0 - 3, nx, nz, j, jin);
    pf_muc = ldr(a.muc, i0 - 3, nx, nz, j, jin);
    pf_nuc = ldr(a.nuc, i0 - 3, nx, nz, j, jin);
  }
  float pf_vx = ldr(ivx, i0 - 1, nx, nz, j, jin), pf_vz = ldr(ivz, i0 - 1, nx, nz, j, jin);
  float pf_rho = ldr(a.rho, i0 - 1, nx, nz, j, jin, 1.0f);
  float pf_lam = ldr(a.lam, i0 - 3, nx, nz, j, jin), pf_mu = ldr(a.mu, i0 - 3, nx, nz, j, jin);
  float vx[4] = {0, 0, 0, 0}, vz[4] = {0, 0, 0, 0}, vyc[4] = {0, 0, 0, 0};   // rows m-3..m
  float dvxdz[3] = {0, 0, 0}, dvzdz[3] = {0, 0, 0}, dvycdz[3] = {0, 0, 0};   // rows m-2..m

  int par = 0;
  for (int m = i0 - 1; m <= i1 + 1; ++m) {
    sxx[0] = sxx[1]; sxx[1] = sxx[2]; sxx[2] = sxx[3]; sxx[3] = pf_sxx;
    sxz[0] = sxz[1]; sxz[1] = sxz[2]; sxz[2] = sxz[3]; sxz[3] = pf_sxz;
    szz[0] = szz[1]; szz[1] = szz[2]; szz[2] = szz[3]; szz[3] = pf_szz;
    if constexpr (SPIN) {
      syx[0] = syx[1]; syx[1] = syx[2]; syx[2] = syx[3]; syx[3] = pf_syx;
      syy[0] = syy[1]; syy[1] = syy[2]; syy[2] = syy[3]; syy[3] = pf_syy;
      syz[0] = syz[1]; syz[1] = syz[2]; syz[2] = syz[3]; syz[3] = pf_syz;
    }
    const float vx_o = pf_vx, vz_o = pf_vz, rho = pf_rho, lam = pf_lam, mu = pf_mu;
    const float vyc_o = pf_vyc, jj = pf_jj, nu = pf_nu, muc = pf_muc, nuc = pf_nuc;
    pf_sxx = ldr(isxx, m + 2, nx, nz, j, jin);
    pf_sxz = ldr(isxz, m + 2, nx, nz, j, jin);
    pf_szz = ldr(iszz, m + 2, nx, nz, j, jin);
    pf_vx = ldr(ivx, m + 1, nx, nz, j, jin);
    pf_vz = ldr(ivz, m + 1, nx, nz, j, jin);
    pf_rho = ldr(a.rho, m + 1, nx, nz, j, jin, 1.0f);
    pf_lam = ldr(a.lam, m - 1, nx, nz, j, jin);
    pf_mu = ldr(a.mu, m - 1, nx, nz, j, jin);
    if constexpr (SPIN) {
      pf_syx = ldr(isyx, m + 2, nx, nz, j, jin);
      pf_syy = ldr(isyy, m + 2, nx, nz, j, jin);
      pf_syz = ldr(isyz, m + 2, nx, nz, j, jin);
      pf_vyc = ldr(ivyc, m + 1, nx, nz, j, jin);
      pf_jj = ldr(a.jj, m + 1, nx, nz, j, jin, 1.0f);
      pf_nu = ldr(a.nu, m - 1, nx, nz, j, jin);
      pf_muc = ldr(a.muc, m - 1, nx, nz, j, jin);
      pf_nuc = ldr(a.nuc, m - 1, nx, nz, j, jin);
    }
    const float bxm = (m >= 0 && m < nx) ? a.bx[m] : 1.0f;

    // ---- velocity half-step at row m
    lds_s[par][0][t] = sxz[2];
    lds_s[par][1][t] = szz[2];
    if constexpr (SPIN) { lds_s[par][2][t] = syz[2]; lds_s[par][3][t] = syy[2]; }
    __syncthreads();
    const bool xm_ok = m >= 2 && m < nx - 2;
    float dsx = (xm_ok ? dminus(sxx[0], sxx[1], sxx[2], sxx[3], h8x) : 0.0f) +
                (zm_ok ? dminus(lds_s[par][0][tm2], lds_s[par][0][tm1], sxz[2], lds_s[par][0][tp1], h8z) : 0.0f);
    float dsz = (xm_ok ? dminus(sxz[0], sxz[1], sxz[2], sxz[3], h8x) : 0.0f) +
                (zm_ok ? dminus(lds_s[par][1][tm2], lds_s[par][1][tm1], szz[2], lds_s[par][1][tp1], h8z) : 0.0f);
    float dsyc = 0.0f;
    if constexpr (SPIN) {
      dsyc = (xm_ok ? dminus(syx[0], syx[1], syx[2], syx[3], h8x) : 0.0f) +
             (zm_ok ? dminus(lds_s[par][2][tm2], lds_s[par][2][tm1], syz[2], lds_s[par][2][tp1], h8z) : 0.0f);
      dsx = dsx + (zm_ok ? dminus(lds_s[par][3][tm2], lds_s[par][3][tm1], syy[2], lds_s[par][3][tp1], h8z) : 0.0f);
      dsz = dsz + (-(xm_ok ? dminus(syy[0], syy[1], syy[2], syy[3], h8x) : 0.0f));
      dsyc = dsyc + 2.0f * syy[2];
    }
    while (ip < ie) {
      const int4 e = a.inj[ip];
      if (e.x > m) break;
      if (e.x == m && e.y == j) {
        dsx = dsx + a.val_x[int64_t(e.z) * a.nt_val + it];
        dsz = dsz + a.val_z[int64_t(e.z) * a.nt_val + it];
      }
      ++ip;
    }
    const float b = (bxm * bzb) * bzt;
    const float vxn = b * (vx_o + dt * dsx / rho);
    const float vzn = b * (vz_o + dt * dsz / rho);
    float vycn = 0.0f;
    if constexpr (SPIN) vycn = b * (vyc_o + dt * dsyc / jj);
    if (jst && m >= i0 && m < i1) {
      ovx[m * nz + j] = vxn;
      ovz[m * nz + j] = vzn;
      if constexpr (SPIN) ovyc[m * nz + j] = vycn;
    }
    vx[0] = vx[1]; vx[1] = vx[2]; vx[2] = vx[3]; vx[3] = vxn;
    vz[0] = vz[1]; vz[1] = vz[2]; vz[2] = vz[3]; vz[3] = vzn;
    if constexpr (SPIN) { vyc[0] = vyc[1]; vyc[1] = vyc[2]; vyc[2] = vyc[3]; vyc[3] = vycn; }
    lds_v[par][0][t] = vxn;
    lds_v[par][1][t] = vzn;
    if constexpr (SPIN) lds_v[par][2][t] = vycn;
    __syncthreads();
    dvxdz[0] = dvxdz[1]; dvxdz[1] = dvxdz[2];
    dvzdz[0] = dvzdz[1]; dvzdz[1] = dvzdz[2];
    dvxdz[2] = zp_ok ? dplus(lds_v[par][0][tm1], vxn, lds_v[par][0][tp1], lds_v[par][0][tp2], h8z) : 0.0f;
    dvzdz[2] = zp_ok ? dplus(lds_v[par][1][tm1], vzn, lds_v[par][1][tp1], lds_v[par][1][tp2], h8z) : 0.0f;
    if constexpr (SPIN) {
      dvycdz[0] = dvycdz[1]; dvycdz[1] = dvycdz[2];
      dvycdz[2] = zp_ok ? dplus(lds_v[par][2][tm1], vycn, lds_v[par][2][tp1], lds_v[par][2][tp2], h8z) : 0.0f;
    }

    // ---- receivers: u = sum v*dt at the cells the traces need
    if (a.record) {
      if constexpr (!SPIN) {
        while (rp < re) {
          const int4 e = a.rec[rp];
          if (e.x > m) break;
          if (e.x == m && e.y == j) {
            float* u = uacc + e.z * N_UACC;
            const float ux = u[U_UX] + vxn * dt, uz = u[U_UZ] + vzn * dt;
            u[U_UX] = ux; u[U_UZ] = uz;
            obs[0 * nobs + e.z * a.nt + it] = ux;
            obs[2 * nobs + e.z * a.nt + it] = uz;
          }
          ++rp;
        }
      } else {
        // a receiver on row ir needs uz on rows ir-1..ir+2 and ux on columns j-1..j+2
        while (rp < re && a.rec[rp].x + 2 < m) ++rp;
        for (int q = rp; q < re; ++q) {
          const int4 e = a.rec[q];
          if (e.x - 1 > m) break;
          if (e.y == j) {
            float* u = uacc + e.z * N_UACC;
            const int d = m - e.x;
            if (d == -1) {
              u[U_UZ_M1] = u[U_UZ_M1] + vzn * dt;
            } else if (d == 0) {
              u[U_UX] = u[U_UX] + vxn * dt;
              u[U_UZ] = u[U_UZ] + vzn * dt;
              u[U_UYC] = u[U_UYC] + vycn * dt;
              u[U_UX_M1] = u[U_UX_M1] + lds_v[par][0][tm1] * dt;
              u[U_UX_P1] = u[U_UX_P1] + lds_v[par][0][tp1] * dt;
              u[U_UX_P2] = u[U_UX_P2] + lds_v[par][0][tp2] * dt;
            } else if (d == 1) {
              u[U_UZ_P1] = u[U_UZ_P1] + vzn * dt;
            } else {
              const float uz_p2 = u[U_UZ_P2] + vzn * dt;
              u[U_UZ_P2] = uz_p2;
              const float ux = u[U_UX], uz = u[U_UZ], uyc = u[U_UYC];
              const bool xp = e.x >= 1 && e.x < nx - 2;
              const bool zp = e.y >= 1 && e.y < nz - 2;
              const float duzdx = xp ? dplus(u[U_UZ_M1], uz, u[U_UZ_P1], uz_p2, h8x) : 0.0f;
              const float duxdz = zp ? dplus(u[U_UX_M1], ux, u[U_UX_P1], u[U_UX_P2], h8z) : 0.0f;
              const float curl = -(duzdx - duxdz);
              const float iso = uyc - 0.5f * curl;
              const int kt = e.z * a.nt + it;
              obs[0 * nobs + kt] = ux;
              obs[2 * nobs + kt] = uz;
              obs[1 * nobs + kt] = uyc;
              obs[3 * nobs + kt] = iso;
              obs[4 * nobs + kt] = curl;
            }
          }
        }
      }
    }

    // ---- stress half-step at row i = m-2
    const int i = m - 2;
    if (i >= i0) {
      const bool xp_ok = i >= 1 && i < nx - 2;
      const float dvxdx = xp_ok ? dplus(vx[0], vx[1], vx[2], vx[3], h8x) : 0.0f;
      const float dvzdx = xp_ok ? dplus(vz[0], vz[1], vz[2], vz[3], h8x) : 0.0f;
      const float l2m = lam + 2.0f * mu;
      const float nsxx = sxx[0] + dt * (l2m * dvxdx + lam * dvzdz[0]);
      const float nszz = szz[0] + dt * (l2m * dvzdz[0] + lam * dvxdx);
      const float nsxz = sxz[0] + dt * mu * (dvxdz[0] + dvzdx);
      if (jst) {
        osxx[i * nz + j] = nsxx;
        oszz[i * nz + j] = nszz;
        osxz[i * nz + j] = nsxz;
      }
      if constexpr (SPIN) {
        const float dvycdx = xp_ok ? dplus(vyc[0], vyc[1], vyc[2], vyc[3], h8x) : 0.0f;
        const float nsyy = syy[0] + 2.0f * dt * nu * (-vyc[1] - 0.5f * (dvzdx - dvxdz[0]));
        const float cc = dt * (muc + nuc);
        const float nsyx = syx[0] + cc * dvycdx;
        const float nsyz = syz[0] + cc * dvycdz[0];
        if (jst) {
          osyy[i * nz + j] = nsyy;
          osyx[i * nz + j] = nsyx;
          osyz[i * nz + j] = nsyz;
        }
      }
    }
    par ^= 1;
  }
  if (a.advance) advance_clock(a.clock, it);
}

// ------------------------------------------------ plain SH imaging condition
// div_uy, div_fw, interaction_muy (fd2d.wgsl:501-509, 571-588)
__global__ void k_imaging_sh(const float* __restrict__ va, const float* __restrict__ vf, float* __restrict__ kmu,
                             int nx, int nz, float h8x, float h8z, float ndt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int k = i * nz + j;
  const bool xp = i >= 1 && i < nx - 2, zp = j >= 1 && j < nz - 2;
  float ax = 0.0f, az = 0.0f, fx = 0.0f, fz = 0.0f;
  if (xp) {
    ax = dplus(va[k - nz], va[k], va[k + nz], va[k + 2 * nz], h8x);
    fx = dplus(vf[k - nz], vf[k], vf[k + nz], vf[k + 2 * nz], h8x);
  }
  if (zp) {
    az = dplus(va[k - 1], va[k], va[k + 1], va[k + 2], h8z);
    fz = dplus(vf[k - 1], vf[k], vf[k + 1], vf[k + 2], h8z);
  }
  kmu[k] = kmu[k] + (ax * fx + az * fz) * ndt;
}

// ----------------------------------------------------- Gaussian smoothing
// init_gsum, gaussian_x, gaussian_z (fd2d.wgsl:517-554, 592-609).  w[d] is the
// 1-D weight exp(-d*d/(2*sigma*sigma)) tabulated by the host; every sum runs
// over the full axis in the reference's order.
__global__ void k_gsum(const float* __restrict__ sumx, const float* __restrict__ sumz, float* __restrict__ gsum,
                       int nx, int nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  gsum[i * nz + j] = sumx[i] * sumz[j];
}
__global__ void k_gauss_x(const float* __restrict__ w, const float* __restrict__ kin, float* __restrict__ gtmp,
                          int nx, int nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  float s = 0.0f;
  for (int n = 0; n < nx; ++n) s = s + w[abs(i - n)] * kin[n * nz + j];
  gtmp[i * nz + j] = s;
}
__global__ void k_gauss_z(const float* __restrict__ w, const float* __restrict__ gtmp,
                          const float* __restrict__ gsum, float* __restrict__ kout, int nx, int nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  float s = 0.0f;
  const float* row = gtmp + i * nz;
  for (int n = 0; n < nz; ++n) s = s + w[abs(j - n)] * row[n];
  const float g = gsum[i * nz + j];
  kout[i * nz + j] = (g > 1e-30f) ? s / g : 0.0f;
}

// ------------------------------------------ waveform misfit + adjoint source
// compute_waveform_misfit (solver.rs:432-467): adstf time-reversed; the
// squared-misfit partial sums are reduced per block in a fixed order.
__global__ __launch_bounds__(256) void k_misfit(const float* __restrict__ syn, const float* __restrict__ ref,
                                                const float* __restrict__ taper, float* __restrict__ adstf,
                                                double* __restrict__ partial, int nrec, int nt) {
  __shared__ double red[256];
  const int64_t n = int64_t(nrec) * nt;
  const int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x;
  double acc = 0.0;
  if (g < n) {
    const int r = int(g / nt), it = int(g - int64_t(r) * nt);
    const float tp = taper[it];
    const float diff = (syn[g] - ref[g]) * tp;
    acc = double(diff) * double(diff);
    adstf[int64_t(r) * nt + (nt - 1 - it)] = diff * tp * 2.0f;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(256) void k_reduce_sum(const double* __restrict__ partial, int n, double* out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) acc += partial[q];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// vps2lm (fd2d.wgsl:188-201)
__global__ void k_vps2lm(const float* __restrict__ rho, float* __restrict__ lam, float* __restrict__ mu,
                         int64_t npt) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= npt) return;
  const float vp = lam[k], vs = mu[k], r = rho[k];
  mu[k] = r * vs * vs;
  lam[k] = (vp > vs) ? r * (vp * vp - 2.0f * vs * vs) : 0.0f;
}

__global__ void k_fill(float* __restrict__ p, float v, int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// plain streaming copy (float4), the calibration kernel for HBM counters
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n4) {
  for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < n4; k += int64_t(gridDim.x) * 256) dst[k] = src[k];
}

template <int NT>
hipError_t launch_sh_nt(const StepArgs& a, const Tiling& t, hipStream_t s) {
  const int nblk = a.nshots * t.nstrips * t.nsegs;
  hipLaunchKernelGGL(k_step_sh<NT>, dim3(nblk), dim3(NT), 0, s, a);
  return hipGetLastError();
}
template <int NT, bool SPIN>
hipError_t launch_psv_nt(const StepArgs& a, const Tiling& t, hipStream_t s) {
  const int nblk = a.nshots * t.nstrips * t.nsegs;
  hipLaunchKernelGGL((k_step_psv<NT, SPIN>), dim3(nblk), dim3(NT), 0, s, a);
  return hipGetLastError();
}

}  // namespace

bool tiling_supported(int nt) { return nt == 64 || nt == 128 || nt == 256; }

hipError_t launch_step_sh(const StepArgs& a, const Tiling& t, hipStream_t s) {
  switch (t.nt_threads) {
    case 64: return launch_sh_nt<64>(a, t, s);
    case 128: return launch_sh_nt<128>(a, t, s);
    case 256: return launch_sh_nt<256>(a, t, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_step_psv(const StepArgs& a, const Tiling& t, bool spin, hipStream_t s) {
  switch (t.nt_threads) {
    case 64: return spin ? launch_psv_nt<64, true>(a, t, s) : launch_psv_nt<64, false>(a, t, s);
    case 128: return spin ? launch_psv_nt<128, true>(a, t, s) : launch_psv_nt<128, false>(a, t, s);
    case 256: return spin ? launch_psv_nt<256, true>(a, t, s) : launch_psv_nt<256, false>(a, t, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_imaging_sh(const float* va, const float* vf, float* kmu, int nx, int nz, float h8x, float h8z,
                             float ndt, hipStream_t s) {
  hipLaunchKernelGGL(k_imaging_sh, dim3((nz + 255) / 256, nx), dim3(256), 0, s, va, vf, kmu, nx, nz, h8x, h8z, ndt);
  return hipGetLastError();
}
hipError_t launch_gsum(const float* sumx, const float* sumz, float* gsum, int nx, int nz, hipStream_t s) {
  hipLaunchKernelGGL(k_gsum, dim3((nz + 255) / 256, nx), dim3(256), 0, s, sumx, sumz, gsum, nx, nz);
  return hipGetLastError();
}
hipError_t launch_gauss_x(const float* w, const float* k, float* gtmp, int nx, int nz, hipStream_t s) {
  hipLaunchKernelGGL(k_gauss_x, dim3((nz + 255) / 256, nx), dim3(256), 0, s, w, k, gtmp, nx, nz);
  return hipGetLastError();
}
hipError_t launch_gauss_z(const float* w, const float* gtmp, const float* gsum, float* k, int nx, int nz,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_gauss_z, dim3((nz + 255) / 256, nx), dim3(256), 0, s, w, gtmp, gsum, k, nx, nz);
  return hipGetLastError();
}
hipError_t launch_misfit(const float* syn, const float* ref, const float* taper, float* adstf, double* partial,
                         int nrec, int nt, hipStream_t s, int* nblocks_out) {
  const int64_t n = int64_t(nrec) * nt;
  const int nb = int((n + 255) / 256);
  *nblocks_out = nb;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_misfit, dim3(nb), dim3(256), 0, s, syn, ref, taper, adstf, partial, nrec, nt);
  return hipGetLastError();
}
hipError_t launch_reduce_sum(const double* partial, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, s, partial, n, out);
  return hipGetLastError();
}
hipError_t launch_vps2lm(const float* rho, float* lam, float* mu, int64_t npt, hipStream_t s) {
  hipLaunchKernelGGL(k_vps2lm, dim3(unsigned((npt + 255) / 256)), dim3(256), 0, s, rho, lam, mu, npt);
  return hipGetLastError();
}
hipError_t launch_fill(float* p, float v, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_copy(const float* src, float* dst, int64_t n, hipStream_t s) {
  const int64_t n4 = n / 4;
  const int64_t nb = std::min<int64_t>((n4 + 255) / 256, 65536);
  hipLaunchKernelGGL(k_copy, dim3(unsigned(nb)), dim3(256), 0, s, reinterpret_cast<const float4*>(src),
                     reinterpret_cast<float4*>(dst), n4);
  return hipGetLastError();
}

}  // namespace sw
