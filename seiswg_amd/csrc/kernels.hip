// gfx950 kernels of the staggered-grid time loop.
//
// One time step of one physics is ONE launch.  Each workgroup owns a z-strip
// of W = NT-6 columns (z is the contiguous axis, k = i*nz + j, fd2d.wgsl:140)
// and marches a segment of rows along x.  x-neighbours stay in registers (a
// 4-row ring per stress field, a 4-row ring of new velocities), z-neighbours
// are exchanged through LDS.  The velocity update of row m and the stress
// update of row m-2 are fused in the same iteration, so every state value is
// read once and written once per step; the 3-column/3-row halos are
// recomputed, never exchanged.  Reads and writes go to different (ping-pong)
// state buffers, so tiles never race on their halos.  All field traffic uses
// buffer loads/stores: the per-thread column offset is a constant VGPR and the
// row/field offset a scalar, so no vector address arithmetic is spent per load.
//
// Arithmetic follows fd2d.wgsl operation by operation (no FMA contraction,
// IEEE-exact divisions), so results equal the reference kernel sequence of
// solver.rs:999-1030 evaluated in FP32.
#include <type_traits>

#include "kernels.h"

#pragma clang fp contract(off)

namespace sw {
namespace {

// Exact division x / y with the correctly rounded reciprocal r = RN(1/y)
// known in advance (Markstein's correction): q0 = x*r, e = q0*y - x (exact in
// one FMA), q = q0 - e*r.  For y in [2^-20, 2^20] (checked on the host, which
// clears `fast` otherwise) and x zero or 2^-100 <= |x| <= 2^100, q equals the
// IEEE quotient bit for bit, signed zeros included (verified exhaustively on
// whole binades at the range ends and in between).  Outside that range -
// numerators so small that the quotient is subnormal, where the correction
// can miss by one unit - the lanes concerned take the IEEE division.  3 VALU
// plus a 3-VALU range test instead of the ~10 of the generic sequence.
struct Div {
  float y, r;
  bool fast;
};
__device__ __forceinline__ bool markstein_range(float x) {
  const float ax = __builtin_fabsf(x);
  return __builtin_amdgcn_fmed3f(ax, 0x1p-100f, 0x1p100f) == ax || x == 0.0f;
}
__device__ __forceinline__ float div_exact(float x, Div d) {
  const float q0 = x * d.r;
  const float e = __builtin_fmaf(q0, d.y, -x);
  const float q = __builtin_fmaf(-e, d.r, q0);
  if (__builtin_expect(d.fast && markstein_range(x), 1)) return q;
  return x / d.y;
}

// fd2d.wgsl:166-180 forward difference, with the reference's 1/8 weight on
// the 3-cell difference; d.y = 8*h, d.r = 1/(8*h).
__device__ __forceinline__ float dplus(float fm1, float f0, float fp1, float fp2, Div d) {
  return div_exact(9.0f * (fp1 - f0) - (fp2 - fm1), d);
}
// numerators of the backward (fd2d.wgsl:151-164) and forward differences
__device__ __forceinline__ float nminus(float fm2, float fm1, float f0, float fp1) {
  return 9.0f * (f0 - fm1) - (fp1 - fm2);
}
__device__ __forceinline__ float nplus(float fm1, float f0, float fp1, float fp2) {
  return 9.0f * (fp1 - f0) - (fp2 - fm1);
}
// v where an FD operator is active, +0 elsewhere (the operators' edge rule),
// as a select of a value computed on every lane rather than a branch around it
__device__ __forceinline__ float keep(bool active, float v) { return active ? v : 0.0f; }

// Division by a constant y = 8h (Brisebarre, Muller and Raina, "Accelerating
// correctly rounded floating-point division when the divisor is known in
// advance", IEEE Trans. Computers 53(8), 2004): with 1/y split as zh + zl
// (zh = RN(1/y), or RD(1/y) where that keeps zl = RN(1/y - zh) from being
// negative, so that x = -0 gives -0), q = RN(x zh + RN(x zl)) - one
// multiplication and one FMA.  For most divisors this equals RN(x / y) for
// every x; the host checks the divisor at hand on every significand of one
// binade (const_div), which covers every x whose products x zh, x zl and
// quotient stay normal (scaling by a power of two commutes with rounding
// there), and passes in `lo` the smallest numerator magnitude T for which
// that holds.  r = RN(1/y) serves the reciprocal-and-correction form
// (dplus) and FD_RCP.
struct CDiv {
  float y, r, zh, zl;
  uint32_t lo;     // (bits(T) << 1) - 1
  bool fast;       // y in [2^-20, 2^20] and the two-operation form exact for it
  __device__ operator Div() const { return Div{y, r, fast}; }
};

// Range test of a group of numerators: true when every x is zero or
// T <= |x| <= 2^100.  (bits(x) << 1) - 1 drops the sign and wraps zero to
// the maximum, so the minimum over the group is >= (bits(T) << 1) - 1
// exactly when no nonzero |x| lies below T (a shift-add per numerator and a
// three-input minimum per two); the maximum of |x| (absolute values are
// free operand modifiers, a three-input maximum per two) bounds the other
// end and catches infinities.  NaN numerators give NaN on either path.
template <int N>
__device__ __forceinline__ bool div_in_range(const float (&x)[N], uint32_t lo_bits) {
  uint32_t lo = 0xffffffffu;
  float hi = 0.0f;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    lo = min(lo, (__builtin_bit_cast(uint32_t, x[k]) << 1) - 1u);
    hi = __builtin_fmaxf(hi, __builtin_fabsf(x[k]));
  }
  return lo >= lo_bits && hi <= 0x1p100f;
}
// the threshold of the reciprocal-and-correction form: 2^-100
constexpr uint32_t kMarksteinLo = ((127u - 100u) << 24) - 1u;

// A group of divisions q[k] = x[k] / d[k].y sharing one range test: the fast
// form for all of them, and the IEEE division for the lanes where a
// numerator is outside the range or a divisor outside the fast set, so a
// group costs one branch instead of one per division.
// mode (from div_mode): 1 fast form with the range test, 0 IEEE division,
// 2 x * RN(1/y) (FD_RCP)
__device__ __forceinline__ int div_mode(int bits, int which) {
  return (bits & FD_RCP) ? 2 : ((bits & which) == which ? 1 : 0);
}
// divisors rho and J: reciprocal-and-correction
template <int N>
__device__ __forceinline__ void div_group(float (&q)[N], const float (&x)[N], const Div (&d)[N], int mode) {
  if (mode == 2) {
#pragma unroll
    for (int k = 0; k < N; ++k) q[k] = x[k] * d[k].r;
    return;
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const float q0 = x[k] * d[k].r;
    const float e = __builtin_fmaf(q0, d[k].y, -x[k]);
    q[k] = __builtin_fmaf(-e, d[k].r, q0);
  }
  if (__builtin_expect(!(mode == 1 && div_in_range(x, kMarksteinLo)), 0)) {
#pragma unroll
    for (int k = 0; k < N; ++k) q[k] = x[k] / d[k].y;
  }
}
// divisors 8 dx, 8 dz: the two-operation form
template <int N>
__device__ __forceinline__ void div_group(float (&q)[N], const float (&x)[N], const CDiv (&d)[N], int mode) {
  if (mode == 2) {
#pragma unroll
    for (int k = 0; k < N; ++k) q[k] = x[k] * d[k].r;
    return;
  }
#pragma unroll
  for (int k = 0; k < N; ++k) q[k] = __builtin_fmaf(x[k], d[k].zh, x[k] * d[k].zl);
  if (__builtin_expect(!(mode == 1 && div_in_range(x, d[0].lo)), 0)) {
#pragma unroll
    for (int k = 0; k < N; ++k) q[k] = x[k] / d[k].y;
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Map them so that
// consecutive logical tiles (which share halos) land on the same XCD and its L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, p = b >> 3;
  return (x < r) ? x * (q + 1) + p : r * (q + 1) + (x - r) * q + p;
}

// Raw buffer resource over `bytes` bytes (gfx9 dword3: 32-bit raw access).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, int(bytes), 0x00020000);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), 0));
}
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, int(voff), int(soff), 0);
}
// store of the next state or of aux fields.  When the launch's state
// footprint far exceeds the caches (nt = 1, chosen by the host) the store is
// non-temporal, so it does not displace the model rows the other shots of
// the launch still read (two-step P-SV, 64 shots on 2000x1000: 1.1 % faster;
// 16-shot gradient on 1000x600: 3 % faster); a footprint that stays in the
// Infinity Cache keeps ordinary stores (one shot of 2000x1000: 5 % slower
// with non-temporal ones).
__device__ __forceinline__ void bstore_out(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, int nt) {
  if (nt) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, int(voff), int(soff), 2);
  else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, int(voff), int(soff), 0);
}

// Advance the step clock once every workgroup of this launch has read it:
// every workgroup read clock[0] at its start, before it counts itself here,
// so the last one to count may overwrite it.  Device-scope atomics order
// the count among workgroups; the next launch sees the new clock because a
// kernel boundary makes this launch's writes visible (no fence needed).
__device__ __forceinline__ void advance_clock(int* clock, int it, int dir, int nblk) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const int done = atomicAdd(&clock[1], 1);
    if (done == nblk - 1) {
      atomicExch(&clock[1], 0);
      atomicExch(&clock[0], it + dir);
    }
  }
}

template <int R>
using Phase = std::integral_constant<int, R>;

// A register copy the compiler cannot fold away.  Stress rows stay in the
// register ring for four rows after leaving the prefetch FIFO; by their last
// use more than 63 vector-memory operations (the counter's range) have been
// issued after their load, and the compiler would then wait for nearly all
// outstanding prefetches.  Copying a row as it enters the ring waits for its
// load once, at a point where the count is still exact.
__device__ __forceinline__ float settle(float x) {
  float y;
  asm("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
}

// Uniform reads of data a launch never writes (per-tile event lists, source
// and adjoint-source values) through the scalar cache (constant address
// space).  Their waits count scalar loads only: a vector load here would make
// the row that reads an event wait for every row prefetched before it (in
// the adjoint, whose sources sit at the receivers, that cost 4 % of the
// gradient).  A divergent index falls back to a vector load.
__device__ __forceinline__ int4 list_entry(const int4* list, int q) {
  const auto* c = (const __attribute__((address_space(4))) int*)list;
  return make_int4(c[4 * q], c[4 * q + 1], c[4 * q + 2], c[4 * q + 3]);
}
__device__ __forceinline__ float uniform_value(const float* p, int64_t k) {
  return ((const __attribute__((address_space(4))) float*)p)[k];
}

// Row of CSR entry q of a per-tile list (INT_MAX past the end), a scalar: the
// per-row compare against it needs no vector wait.
__device__ __forceinline__ int next_row(const int4* list, int q, int end) {
  return __builtin_amdgcn_readfirstlane(q < end ? list_entry(list, q).x : 0x7fffffff);
}

// Uniform read through the scalar cache (constant address space): the x
// taper factor of a row.  Scalar loads leave the vector-memory counter to
// the prefetched rows (at most 63 vector loads and stores can be pending).
__device__ __forceinline__ float taper_x(const float* bx, int row, int nx) {
  const auto* c = (const __attribute__((address_space(4))) float*)bx;
  return c[min(max(row, 0), nx - 1)];
}

// Geometry shared by the step kernels.  Rows are clamped to the grid and the
// column to [0, nz): loads never branch and never leave the slice; values read
// for rows/columns outside the grid only feed halo lanes whose results are
// discarded or zeroed by the edge rules of the FD operators.
struct TileGeom {
  int shot, tile, t, lane, base, j, i0, i1;
  bool jin, jst, zm_ok, zp_ok;
  uint32_t jc4, rowb;
  int nx;
  __device__ __forceinline__ uint32_t row(int m) const { return uint32_t(min(max(m, 0), nx - 1)) * rowb; }
};

// A workgroup of S*NT threads runs the same tile of S consecutive shots, one
// NT-thread group per shot (S divides the shot count): the groups share the
// barriers, so their model-field loads of a row come close together in time
// and all but the first are served by the caches.
// H is the halo width in columns: 3 for one step per launch, 6 for two.
template <int NT, int S = 1, int H = 3>
__device__ __forceinline__ TileGeom tile_geom(const StepArgs& a, int logical) {
  constexpr int W = NT - 2 * H;
  TileGeom g;
  const int tps = a.nstrips * a.nsegs;
  int grp;
  if (a.tile_major) {
    // the shot groups of one tile run side by side (on one XCD after
    // xcd_remap), so they share its model rows in that XCD's L2
    const int ngrp = a.nshots / S;
    g.tile = logical / ngrp;
    grp = logical - g.tile * ngrp;
  } else {
    grp = logical / tps;
    g.tile = logical - grp * tps;
  }
  const int sub = S == 1 ? 0 : __builtin_amdgcn_readfirstlane(int(threadIdx.x) / NT);   // NT is a multiple of 64
  g.shot = grp * S + sub;
  g.base = sub * NT;
  g.lane = int(threadIdx.x);
  const int strip = g.tile % a.nstrips, seg = g.tile / a.nstrips;
  g.t = g.lane - g.base;
  g.nx = a.nx;
  g.j = strip * W - H + g.t;
  g.i0 = seg * a.seglen;
  g.i1 = min(g.i0 + a.seglen, a.nx);
  g.jin = g.j >= 0 && g.j < a.nz;
  g.jst = g.jin && g.t >= H && g.t < NT - H;
  g.zm_ok = g.j >= 2 && g.j < a.nz - 2;
  g.zp_ok = g.j >= 1 && g.j < a.nz - 2;
  g.jc4 = 4u * uint32_t(min(max(g.j, 0), a.nz - 1));
  g.rowb = 4u * uint32_t(a.nz);
  return g;
}

// ---------------------------------------------------------------- SH step
// div_sy, stf_dsy, add_vy, div_vy, add_sy_sh, save_obs_y_sh (fd2d.wgsl:242-301)
template <int NT, int S, bool AUX>
__device__ __forceinline__ void step_sh(const StepArgs& a, int logical) {
  __shared__ float lds_s[2][S * NT];
  __shared__ float lds_v[2][S * NT];
  const TileGeom g = tile_geom<NT, S>(a, logical);
  const int t = g.t, lane = g.lane, j = g.j, nx = a.nx, i0 = g.i0, i1 = g.i1;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (it < a.it_lo || it >= a.it_hi) return;
  const int64_t so = int64_t(g.shot) * a.shot_floats;
  const uint32_t sbytes = 4u * uint32_t(a.shot_floats);
  const auto rin = make_rsrc(a.st[it & 1] + so, sbytes);
  const auto rout = make_rsrc(a.st[(it + 1) & 1] + so, sbytes);
  const uint32_t npt4 = 4u * uint32_t(a.npt);
  const auto rmod = make_rsrc(a.model, uint32_t(MF_N) * npt4);
  const uint32_t o_vy = 4u * a.slot_off[S_VY], o_sxy = 4u * a.slot_off[S_SXY], o_szy = 4u * a.slot_off[S_SZY];
  const uint32_t m_rho = MF_RHO * npt4, m_mu = MF_MU * npt4, m_irho = MF_IRHO * npt4;
  const uint32_t jc4 = g.jc4;
  const int jc = int(jc4 >> 2);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const float dt = a.dt;
  const CDiv Dx{a.h8x, a.r8x, a.z8x, a.l8x, a.c_lo, (a.fast_div & FD_CONST) != 0},
             Dz{a.h8z, a.r8z, a.z8z, a.l8z, a.c_lo, (a.fast_div & FD_CONST) != 0};
  const int fast_c = div_mode(a.fast_div, FD_CONST);
  const int fast_rho = div_mode(a.fast_div, FD_RHO);
  const int tm2 = g.base + max(t - 2, 0), tm1 = g.base + max(t - 1, 0), tp1 = g.base + min(t + 1, NT - 1),
            tp2 = g.base + min(t + 2, NT - 1);
  const int tps = a.nstrips * a.nsegs;
  int ip = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile + 1]);
  int inj_row = next_row(a.inj, ip, ie);     // row of the next injection entry
  int rp = 0, re = 0;
  if (a.record) {
    rp = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile]);
    re = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile + 1]);
  }
  int rec_row = next_row(a.rec, rp, re);     // row of the next receiver entry
  float* uacc = a.uacc + int64_t(g.shot) * a.nrec * N_UACC;
  float* obs = a.obs + int64_t(g.shot) * 5 * a.nrec * a.nt;
  // forward recompute of the gradient: the new velocity v^{it+1} into its
  // time slot (it + 1) % aux_period
  const auto raux =
      make_rsrc(AUX ? a.aux + (int64_t((it + 1) % a.aux_period) * a.nshots + g.shot) * NAUX_SH * int64_t(a.npt)
                    : a.model,
                AUX ? uint32_t(NAUX_SH) * npt4 : 0u);

  // rings: stress row r in slot (r - (i0-3)) & 3; new-velocity row r in slot
  // (r - (i0-4)) & 3; d+z of new vy, row r, in slot (r - (i0-3)) & 3
  float sxy[4], szy[4], vy[4] = {0, 0, 0, 0}, dvydz[4] = {0, 0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    sxy[q] = settle(bload(rin, jc4, o_sxy + g.row(i0 - 3 + q)));
    szy[q] = settle(bload(rin, jc4, o_szy + g.row(i0 - 3 + q)));
  }
  sxy[3] = szy[3] = 0.0f;
  // D-deep prefetch FIFO, slot (m - (i0-1)) % D: stress row m+1, velocity and
  // density row m, shear modulus row m-2 of iteration m.  Four shots per
  // workgroup leave one workgroup per CU, with registers to spare for D = 4.
  constexpr int D = S >= 4 ? 4 : 2;
  float pf_sxy[D], pf_szy[D], pf_vy[D], pf_rho[D], pf_irho[D], pf_mu[D];
  float bx_next = taper_x(a.bx, i0 - 1, nx);     // taper factor of the next velocity row
  for (int q = 0; q < D; ++q) {
    pf_sxy[q] = bload(rin, jc4, o_sxy + g.row(i0 + q));
    pf_szy[q] = bload(rin, jc4, o_szy + g.row(i0 + q));
    pf_vy[q] = bload(rin, jc4, o_vy + g.row(i0 - 1 + q));
    pf_rho[q] = bload(rmod, jc4, m_rho + g.row(i0 - 1 + q));
    pf_irho[q] = bload(rmod, jc4, m_irho + g.row(i0 - 1 + q));
    pf_mu[q] = bload(rmod, jc4, m_mu + g.row(i0 - 3 + q));
  }
  // One barrier per row: iteration m reads the z-neighbours of stress row m
  // from lds_s[m & 1] (written by iteration m-1, or here for the first row),
  // writes new-velocity row m to lds_v[m & 1] and stress row m+1 to
  // lds_s[(m+1) & 1], then synchronises once before reading velocity
  // neighbours.  Parity double-buffering keeps every write two barriers away
  // from the reads of the previous contents.
  lds_s[0][lane] = szy[2];
  __syncthreads();

  auto body = [&](auto phase, int m) __attribute__((always_inline)) {
    constexpr int R = decltype(phase)::value;
    constexpr int s0 = R & 3, s1 = (R + 1) & 3, s2 = (R + 2) & 3, s3 = (R + 3) & 3, par = R & 1;
    constexpr int f = R % D;
    sxy[s3] = settle(pf_sxy[f]);       // row m+1
    szy[s3] = settle(pf_szy[f]);
    const float vy_o = pf_vy[f], rho = pf_rho[f], irho = pf_irho[f], mu = pf_mu[f], bxm = bx_next;
    bx_next = taper_x(a.bx, m + 1, nx);
    {
      const uint32_t r3 = g.row(m + D + 1), r2 = g.row(m + D), r0 = g.row(m + D - 2);
      pf_sxy[f] = bload(rin, jc4, o_sxy + r3);
      pf_szy[f] = bload(rin, jc4, o_szy + r3);
      pf_vy[f] = bload(rin, jc4, o_vy + r2);
      pf_rho[f] = bload(rmod, jc4, m_rho + r2);
      pf_irho[f] = bload(rmod, jc4, m_irho + r2);
      pf_mu[f] = bload(rmod, jc4, m_mu + r0);
    }

    // velocity half-step at row m
    const bool xm_ok = m >= 2 && m < nx - 2;
    float qa[2];
    {
      const float xa[2] = {xm_ok ? nminus(sxy[s0], sxy[s1], sxy[s2], sxy[s3]) : 0.0f,
                           g.zm_ok ? nminus(lds_s[par][tm2], lds_s[par][tm1], szy[s2], lds_s[par][tp1]) : 0.0f};
      div_group(qa, xa, {Dx, Dz}, fast_c);
    }
    float dsy = qa[0] + qa[1];
    if (m == inj_row) {
      for (; ip < ie; ++ip) {
        const int4 e = list_entry(a.inj, ip);
        if (e.x != m) break;
        if (e.y == j) dsy = dsy + uniform_value(a.val_y, int64_t(e.z) * a.nt_val + it);
      }
      inj_row = next_row(a.inj, ip, ie);
    }
    const float b = (bxm * bzb) * bzt;
    float qb[1];
    {
      const float xb[1] = {dt * dsy};
      div_group(qb, xb, {Div{rho, irho, true}}, fast_rho);
    }
    const float vyn = b * (vy_o + qb[0]);
    if (g.jst && m >= i0 && m < i1) {
      bstore_out(vyn, rout, jc4, o_vy + uint32_t(m) * g.rowb, a.stream_out);
      if constexpr (AUX) bstore_out(vyn, raux, jc4, A_VY * npt4 + uint32_t(m) * g.rowb, a.stream_out);
    }
    vy[s3] = vyn;              // new-velocity ring: rows m-3..m in slots s0..s3
    lds_v[par][lane] = vyn;
    lds_s[par ^ 1][lane] = szy[s3];
    __syncthreads();
    // d+z of the new vy at row m and d+x at row m-2 (for the stress half-step)
    const int i = m - 2;
    const bool xp_ok = i >= 1 && i < nx - 2;
    float qc[2];
    {
      const float xc[2] = {g.zp_ok ? nplus(lds_v[par][tm1], vyn, lds_v[par][tp1], lds_v[par][tp2]) : 0.0f,
                           xp_ok ? nplus(vy[s0], vy[s1], vy[s2], vy[s3]) : 0.0f};
      div_group(qc, xc, {Dz, Dx}, fast_c);
    }
    dvydz[s2] = qc[0];

    // receivers on row m: uy += vy*dt, trace = uy
    if (m == rec_row) {
      for (; rp < re; ++rp) {
        const int4 e = list_entry(a.rec, rp);
        if (e.x != m) break;
        if (e.y == j) {
          float* u = uacc + e.z * N_UACC;
          const float uy = u[U_UY] + vyn * dt;
          u[U_UY] = uy;
          obs[(1 * a.nrec + e.z) * a.nt + it] = uy;
        }
      }
      rec_row = next_row(a.rec, rp, re);
    }

    // stress half-step at row i = m-2
    if (i >= i0) {
      const float mu_dt = mu * dt;
      const float dvydx = qc[1];
      const float nsxy = sxy[s0] + mu_dt * dvydx;
      const float nszy = szy[s0] + mu_dt * dvydz[s0];
      if (g.jst) {
        bstore_out(nsxy, rout, jc4, o_sxy + uint32_t(i) * g.rowb, a.stream_out);
        bstore_out(nszy, rout, jc4, o_szy + uint32_t(i) * g.rowb, a.stream_out);
      }
    }
  };

  // The first four rows run before the loop, which then repeats 8 phases:
  // a FIFO slot refilled in phase k is read in phase k+D, so the registers
  // carried around the loop are the same every trip (no copies of rows still
  // in flight at the back-edge), and on entering the loop every slot already
  // holds a row loaded in that steady pattern, so the compiler's wait counts
  // for the loop body are those of the steady state rather than those of the
  // prologue's load order.
  const int mlast = i1 + 1;
  int m = i0 - 1;
  body(Phase<0>{}, m);
  if (++m <= mlast) {
    body(Phase<1>{}, m);
    if (++m <= mlast) {
      body(Phase<2>{}, m);
      if (++m <= mlast) {
        body(Phase<3>{}, m);
        if (++m <= mlast) {
          for (;;) {
            body(Phase<4>{}, m);
            if (++m > mlast) break;
            body(Phase<5>{}, m);
            if (++m > mlast) break;
            body(Phase<6>{}, m);
            if (++m > mlast) break;
            body(Phase<7>{}, m);
            if (++m > mlast) break;
            body(Phase<0>{}, m);
            if (++m > mlast) break;
            body(Phase<1>{}, m);
            if (++m > mlast) break;
            body(Phase<2>{}, m);
            if (++m > mlast) break;
            body(Phase<3>{}, m);
            if (++m > mlast) break;
          }
        }
      }
    }
  }
  if (a.advance) advance_clock(a.clock, it, 1, a.nblocks);
}

// ------------------------------------------------------- P-SV (+spin) step
// div_sxz, div_sy_c, div_sxyz_c, stf_dsxz, add_vxz, div_vxz, add_sxz,
// add_vy_c, div_vy_c, add_sy_c, save_obs_x/z, compute_du_grad, save_obs_ry
// (fd2d.wgsl:308-492)
template <int NT, int S, bool SPIN, bool AUX>
__device__ __forceinline__ void step_psv(const StepArgs& a, int logical) {
  constexpr int NSL = SPIN ? 4 : 2;   // sxz, szz, syz_c, syy_c
  constexpr int NVL = SPIN ? 3 : 2;   // vx, vz, vy_c
  __shared__ float lds_s[2][NSL][S * NT];
  __shared__ float lds_v[2][NVL][S * NT];
  const TileGeom g = tile_geom<NT, S>(a, logical);
  const int t = g.t, lane = g.lane, j = g.j, nx = a.nx, nz = a.nz, i0 = g.i0, i1 = g.i1;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (it < a.it_lo || it >= a.it_hi) return;
  const int64_t so = int64_t(g.shot) * a.shot_floats;
  const uint32_t sbytes = 4u * uint32_t(a.shot_floats);
  const auto rin = make_rsrc(a.st[it & 1] + so, sbytes);
  const auto rout = make_rsrc(a.st[(it + 1) & 1] + so, sbytes);
  const uint32_t npt4 = 4u * uint32_t(a.npt);
  const auto rmod = make_rsrc(a.model, uint32_t(MF_N) * npt4);
  const uint32_t o_vx = 4u * a.slot_off[S_VX], o_vz = 4u * a.slot_off[S_VZ];
  const uint32_t o_sxx = 4u * a.slot_off[S_SXX], o_szz = 4u * a.slot_off[S_SZZ], o_sxz = 4u * a.slot_off[S_SXZ];
  uint32_t o_vyc = 0, o_syx = 0, o_syy = 0, o_syz = 0;
  if constexpr (SPIN) {
    o_vyc = 4u * a.slot_off[S_VYC]; o_syx = 4u * a.slot_off[S_SYXC];
    o_syy = 4u * a.slot_off[S_SYYC]; o_syz = 4u * a.slot_off[S_SYZC];
  }
  const uint32_t m_rho = MF_RHO * npt4, m_lam = MF_LAM * npt4, m_mu = MF_MU * npt4;
  const uint32_t m_nu = MF_NU * npt4, m_j = MF_J * npt4, m_muc = MF_MU_C * npt4, m_nuc = MF_NU_C * npt4;
  const uint32_t m_irho = MF_IRHO * npt4, m_ij = MF_IJ * npt4;
  const uint32_t jc4 = g.jc4;
  const int jc = int(jc4 >> 2);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const float dt = a.dt;
  const CDiv Dx{a.h8x, a.r8x, a.z8x, a.l8x, a.c_lo, (a.fast_div & FD_CONST) != 0},
             Dz{a.h8z, a.r8z, a.z8z, a.l8z, a.c_lo, (a.fast_div & FD_CONST) != 0};
  const int fast_c = div_mode(a.fast_div, FD_CONST);
  const int fast_rho = div_mode(a.fast_div, FD_RHO), fast_rj = div_mode(a.fast_div, FD_RHO | FD_J);
  const int tm2 = g.base + max(t - 2, 0), tm1 = g.base + max(t - 1, 0), tp1 = g.base + min(t + 1, NT - 1),
            tp2 = g.base + min(t + 2, NT - 1);
  const int tps = a.nstrips * a.nsegs;
  int ip = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile + 1]);
  int inj_row = next_row(a.inj, ip, ie);
  int rp = 0, re = 0;
  if (a.record) {
    rp = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile]);
    re = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile + 1]);
  }
  int rec_row = next_row(a.rec, rp, re);   // first row of the pending receivers
  float* uacc = a.uacc + int64_t(g.shot) * a.nrec * N_UACC;
  float* obs = a.obs + int64_t(g.shot) * 5 * a.nrec * a.nt;
  const int nobs = a.nrec * a.nt;
  // forward recompute of the gradient: the new velocities v^{it+1} into their
  // time slot (it + 1) % aux_period
  const auto raux =
      make_rsrc(AUX ? a.aux + (int64_t((it + 1) % a.aux_period) * a.nshots + g.shot) * NAUX_PSV * int64_t(a.npt)
                    : a.model,
                AUX ? uint32_t(NAUX_PSV) * npt4 : 0u);

  // rings as in the SH kernel
  float sxx[4], sxz[4], szz[4], syx[4], syy[4], syz[4];
  float vx[4] = {0, 0, 0, 0}, vz[4] = {0, 0, 0, 0}, vyc[4] = {0, 0, 0, 0};
  float dvxdz[4] = {0, 0, 0, 0}, dvzdz[4] = {0, 0, 0, 0}, dvycdz[4] = {0, 0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    const uint32_t r = g.row(i0 - 3 + q);
    sxx[q] = settle(bload(rin, jc4, o_sxx + r));
    sxz[q] = settle(bload(rin, jc4, o_sxz + r));
    szz[q] = settle(bload(rin, jc4, o_szz + r));
    if constexpr (SPIN) {
      syx[q] = settle(bload(rin, jc4, o_syx + r));
      syy[q] = settle(bload(rin, jc4, o_syy + r));
      syz[q] = settle(bload(rin, jc4, o_syz + r));
    }
  }
  sxx[3] = sxz[3] = szz[3] = syx[3] = syy[3] = syz[3] = 0.0f;
  if constexpr (!SPIN) syx[0] = syx[1] = syx[2] = syy[0] = syy[1] = syy[2] = syz[0] = syz[1] = syz[2] = 0.0f;
  // D-deep prefetch FIFO, slot (m - (i0-1)) % D: stress rows m+1, velocity /
  // density / J rows m, moduli rows m-2 of iteration m
  constexpr int D = S >= 4 ? 4 : 2;
  float pf_sxx[D], pf_sxz[D], pf_szz[D], pf_vx[D], pf_vz[D], pf_rho[D], pf_irho[D], pf_lam[D], pf_mu[D];
  float bx_next = taper_x(a.bx, i0 - 1, nx);     // taper factor of the next velocity row
  float pf_syx[D] = {}, pf_syy[D] = {}, pf_syz[D] = {}, pf_vyc[D] = {}, pf_jj[D] = {}, pf_ij[D] = {}, pf_nu[D] = {},
        pf_muc[D] = {}, pf_nuc[D] = {};
  for (int q = 0; q < D; ++q) {
    const uint32_t r1 = g.row(i0 + q), r0 = g.row(i0 - 1 + q), rm2 = g.row(i0 - 3 + q);
    pf_sxx[q] = bload(rin, jc4, o_sxx + r1);
    pf_sxz[q] = bload(rin, jc4, o_sxz + r1);
    pf_szz[q] = bload(rin, jc4, o_szz + r1);
    pf_vx[q] = bload(rin, jc4, o_vx + r0);
    pf_vz[q] = bload(rin, jc4, o_vz + r0);
    pf_rho[q] = bload(rmod, jc4, m_rho + r0);
    pf_irho[q] = bload(rmod, jc4, m_irho + r0);
    pf_lam[q] = bload(rmod, jc4, m_lam + rm2);
    pf_mu[q] = bload(rmod, jc4, m_mu + rm2);
    if constexpr (SPIN) {
      pf_syx[q] = bload(rin, jc4, o_syx + r1);
      pf_syy[q] = bload(rin, jc4, o_syy + r1);
      pf_syz[q] = bload(rin, jc4, o_syz + r1);
      pf_vyc[q] = bload(rin, jc4, o_vyc + r0);
      pf_jj[q] = bload(rmod, jc4, m_j + r0);
      pf_ij[q] = bload(rmod, jc4, m_ij + r0);
      pf_nu[q] = bload(rmod, jc4, m_nu + rm2);
      pf_muc[q] = bload(rmod, jc4, m_muc + rm2);
      pf_nuc[q] = bload(rmod, jc4, m_nuc + rm2);
    }
  }

  // one barrier per row, as in the SH step: stress row i0-1 goes to lds_s[0]
  lds_s[0][0][lane] = sxz[2];
  lds_s[0][1][lane] = szz[2];
  if constexpr (SPIN) { lds_s[0][2][lane] = syz[2]; lds_s[0][3][lane] = syy[2]; }
  __syncthreads();

  auto body = [&](auto phase, int m) __attribute__((always_inline)) {
    constexpr int R = decltype(phase)::value;
    constexpr int s0 = R & 3, s1 = (R + 1) & 3, s2 = (R + 2) & 3, s3 = (R + 3) & 3, par = R & 1;
    constexpr int f = R % D;
    sxx[s3] = settle(pf_sxx[f]); sxz[s3] = settle(pf_sxz[f]); szz[s3] = settle(pf_szz[f]);     // row m+1
    if constexpr (SPIN) { syx[s3] = settle(pf_syx[f]); syy[s3] = settle(pf_syy[f]); syz[s3] = settle(pf_syz[f]); }
    const float vx_o = pf_vx[f], vz_o = pf_vz[f], rho = pf_rho[f], irho = pf_irho[f], lam = pf_lam[f], mu = pf_mu[f];
    const float bxm = bx_next;
    bx_next = taper_x(a.bx, m + 1, nx);
    const float vyc_o = pf_vyc[f], jj = pf_jj[f], ij = pf_ij[f], nu = pf_nu[f], muc = pf_muc[f], nuc = pf_nuc[f];
    {
      const uint32_t r3 = g.row(m + D + 1), r2 = g.row(m + D), r0 = g.row(m + D - 2);
      pf_sxx[f] = bload(rin, jc4, o_sxx + r3);
      pf_sxz[f] = bload(rin, jc4, o_sxz + r3);
      pf_szz[f] = bload(rin, jc4, o_szz + r3);
      pf_vx[f] = bload(rin, jc4, o_vx + r2);
      pf_vz[f] = bload(rin, jc4, o_vz + r2);
      pf_rho[f] = bload(rmod, jc4, m_rho + r2);
      pf_irho[f] = bload(rmod, jc4, m_irho + r2);
      pf_lam[f] = bload(rmod, jc4, m_lam + r0);
      pf_mu[f] = bload(rmod, jc4, m_mu + r0);
      if constexpr (SPIN) {
        pf_syx[f] = bload(rin, jc4, o_syx + r3);
        pf_syy[f] = bload(rin, jc4, o_syy + r3);
        pf_syz[f] = bload(rin, jc4, o_syz + r3);
        pf_vyc[f] = bload(rin, jc4, o_vyc + r2);
        pf_jj[f] = bload(rmod, jc4, m_j + r2);
        pf_ij[f] = bload(rmod, jc4, m_ij + r2);
        pf_nu[f] = bload(rmod, jc4, m_nu + r0);
        pf_muc[f] = bload(rmod, jc4, m_muc + r0);
        pf_nuc[f] = bload(rmod, jc4, m_nuc + r0);
      }
    }

    // ---- velocity half-step at row m
    const bool xm_ok = m >= 2 && m < nx - 2;
    const bool zm_ok = g.zm_ok;
    // D- of the stresses: one group of 4 (8 with spin) divisions
    constexpr int NA = SPIN ? 8 : 4;
    float qa[NA];
    {
      float xa[NA];
      xa[0] = keep(xm_ok, nminus(sxx[s0], sxx[s1], sxx[s2], sxx[s3]));
      xa[1] = keep(zm_ok, nminus(lds_s[par][0][tm2], lds_s[par][0][tm1], sxz[s2], lds_s[par][0][tp1]));
      xa[2] = keep(xm_ok, nminus(sxz[s0], sxz[s1], sxz[s2], sxz[s3]));
      xa[3] = keep(zm_ok, nminus(lds_s[par][1][tm2], lds_s[par][1][tm1], szz[s2], lds_s[par][1][tp1]));
      if constexpr (SPIN) {
        xa[4] = keep(xm_ok, nminus(syx[s0], syx[s1], syx[s2], syx[s3]));
        xa[5] = keep(zm_ok, nminus(lds_s[par][2][tm2], lds_s[par][2][tm1], syz[s2], lds_s[par][2][tp1]));
        xa[6] = keep(zm_ok, nminus(lds_s[par][3][tm2], lds_s[par][3][tm1], syy[s2], lds_s[par][3][tp1]));
        xa[7] = keep(xm_ok, nminus(syy[s0], syy[s1], syy[s2], syy[s3]));
        div_group(qa, xa, {Dx, Dz, Dx, Dz, Dx, Dz, Dz, Dx}, fast_c);
      } else {
        div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
      }
    }
    float dsx = qa[0] + qa[1];
    float dsz = qa[2] + qa[3];
    float dsyc = 0.0f;
    if constexpr (SPIN) {
      dsyc = qa[4] + qa[5];
      dsx = dsx + qa[6];
      dsz = dsz + (-qa[7]);
      dsyc = dsyc + 2.0f * syy[s2];
    }
    if (m == inj_row) {
      for (; ip < ie; ++ip) {
        const int4 e = list_entry(a.inj, ip);
        if (e.x != m) break;
        if (e.y == j) {
          dsx = dsx + uniform_value(a.val_x, int64_t(e.z) * a.nt_val + it);
          dsz = dsz + uniform_value(a.val_z, int64_t(e.z) * a.nt_val + it);
        }
      }
      inj_row = next_row(a.inj, ip, ie);
    }
    const float b = (bxm * bzb) * bzt;
    // divisions by rho (and J)
    constexpr int NB = SPIN ? 3 : 2;
    float qb[NB];
    if constexpr (SPIN) {
      const float xb[3] = {dt * dsx, dt * dsz, dt * dsyc};
      div_group(qb, xb, {Div{rho, irho, true}, Div{rho, irho, true}, Div{jj, ij, true}}, fast_rj);
    } else {
      const float xb[2] = {dt * dsx, dt * dsz};
      div_group(qb, xb, {Div{rho, irho, true}, Div{rho, irho, true}}, fast_rho);
    }
    const float vxn = b * (vx_o + qb[0]);
    const float vzn = b * (vz_o + qb[1]);
    float vycn = 0.0f;
    if constexpr (SPIN) vycn = b * (vyc_o + qb[2]);
    if (g.jst && m >= i0 && m < i1) {
      const uint32_t ro = uint32_t(m) * g.rowb;
      bstore_out(vxn, rout, jc4, o_vx + ro, a.stream_out);
      bstore_out(vzn, rout, jc4, o_vz + ro, a.stream_out);
      if constexpr (SPIN) bstore_out(vycn, rout, jc4, o_vyc + ro, a.stream_out);
      if constexpr (AUX) {
        bstore_out(vxn, raux, jc4, A_VX * npt4 + ro, a.stream_out);
        bstore_out(vzn, raux, jc4, A_VZ * npt4 + ro, a.stream_out);
      }
    }
    vx[s3] = vxn;
    vz[s3] = vzn;
    if constexpr (SPIN) vyc[s3] = vycn;
    lds_v[par][0][lane] = vxn;
    lds_v[par][1][lane] = vzn;
    if constexpr (SPIN) lds_v[par][2][lane] = vycn;
    lds_s[par ^ 1][0][lane] = sxz[s3];
    lds_s[par ^ 1][1][lane] = szz[s3];
    if constexpr (SPIN) { lds_s[par ^ 1][2][lane] = syz[s3]; lds_s[par ^ 1][3][lane] = syy[s3]; }
    __syncthreads();
    // D+ of the new velocities: z at row m, x at row i = m-2 (stress half-step)
    const bool zp_ok = g.zp_ok;
    const int i = m - 2;
    const bool xp_ok = i >= 1 && i < nx - 2;
    constexpr int NC = SPIN ? 6 : 4;
    float qc[NC];
    {
      float xc[NC];
      xc[0] = keep(zp_ok, nplus(lds_v[par][0][tm1], vxn, lds_v[par][0][tp1], lds_v[par][0][tp2]));
      xc[1] = keep(zp_ok, nplus(lds_v[par][1][tm1], vzn, lds_v[par][1][tp1], lds_v[par][1][tp2]));
      xc[2] = keep(xp_ok, nplus(vx[s0], vx[s1], vx[s2], vx[s3]));
      xc[3] = keep(xp_ok, nplus(vz[s0], vz[s1], vz[s2], vz[s3]));
      if constexpr (SPIN) {
        xc[4] = keep(zp_ok, nplus(lds_v[par][2][tm1], vycn, lds_v[par][2][tp1], lds_v[par][2][tp2]));
        xc[5] = keep(xp_ok, nplus(vyc[s0], vyc[s1], vyc[s2], vyc[s3]));
        div_group(qc, xc, {Dz, Dz, Dx, Dx, Dz, Dx}, fast_c);
      } else {
        div_group(qc, xc, {Dz, Dz, Dx, Dx}, fast_c);
      }
    }
    dvxdz[s2] = qc[0];
    dvzdz[s2] = qc[1];
    if constexpr (SPIN) dvycdz[s2] = qc[4];

    // ---- receivers: u = sum v*dt at the cells the traces need
    if constexpr (!SPIN) {
      if (m == rec_row) {
        for (; rp < re; ++rp) {
          const int4 e = list_entry(a.rec, rp);
          if (e.x != m) break;
          if (e.y == j) {
            float* u = uacc + e.z * N_UACC;
            const float ux = u[U_UX] + vxn * dt, uz = u[U_UZ] + vzn * dt;
            u[U_UX] = ux;
            u[U_UZ] = uz;
            obs[0 * nobs + e.z * a.nt + it] = ux;
            obs[2 * nobs + e.z * a.nt + it] = uz;
          }
        }
        rec_row = next_row(a.rec, rp, re);
      }
    } else {
      // a receiver on row ir needs uz on rows ir-1..ir+2 and ux on columns j-1..j+2
      if (m >= rec_row - 1) {
        for (int q = rp; q < re; ++q) {
          const int4 e = list_entry(a.rec, q);
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
              const float duzdx = xp ? dplus(u[U_UZ_M1], uz, u[U_UZ_P1], uz_p2, Dx) : 0.0f;
              const float duxdz = zp ? dplus(u[U_UX_M1], ux, u[U_UX_P1], u[U_UX_P2], Dz) : 0.0f;
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
        while (rp < re && list_entry(a.rec, rp).x + 2 <= m) ++rp;
        rec_row = next_row(a.rec, rp, re);
      }
    }

    // ---- stress half-step at row i = m-2
    if (i >= i0) {
      const float dvxdx = qc[2];
      const float dvzdx = qc[3];
      const float l2m = lam + 2.0f * mu;
      const float nsxx = sxx[s0] + dt * (l2m * dvxdx + lam * dvzdz[s0]);
      const float nszz = szz[s0] + dt * (l2m * dvzdz[s0] + lam * dvxdx);
      const float nsxz = sxz[s0] + dt * mu * (dvxdz[s0] + dvzdx);
      const uint32_t ro = uint32_t(i) * g.rowb;
      if (g.jst) {
        bstore_out(nsxx, rout, jc4, o_sxx + ro, a.stream_out);
        bstore_out(nszz, rout, jc4, o_szz + ro, a.stream_out);
        bstore_out(nsxz, rout, jc4, o_sxz + ro, a.stream_out);
      }
      if constexpr (SPIN) {
        const float dvycdx = qc[5];
        const float nsyy = syy[s0] + 2.0f * dt * nu * (-vyc[s1] - 0.5f * (dvzdx - dvxdz[s0]));
        const float cc = dt * (muc + nuc);
        const float nsyx = syx[s0] + cc * dvycdx;
        const float nsyz = syz[s0] + cc * dvycdz[s0];
        if (g.jst) {
          bstore_out(nsyy, rout, jc4, o_syy + ro, a.stream_out);
          bstore_out(nsyx, rout, jc4, o_syx + ro, a.stream_out);
          bstore_out(nsyz, rout, jc4, o_syz + ro, a.stream_out);
        }
      }
    }
  };

  // The first four rows run before the loop, which then repeats 8 phases:
  // a FIFO slot refilled in phase k is read in phase k+D, so the registers
  // carried around the loop are the same every trip (no copies of rows still
  // in flight at the back-edge), and on entering the loop every slot already
  // holds a row loaded in that steady pattern, so the compiler's wait counts
  // for the loop body are those of the steady state rather than those of the
  // prologue's load order.
  const int mlast = i1 + 1;
  int m = i0 - 1;
  body(Phase<0>{}, m);
  if (++m <= mlast) {
    body(Phase<1>{}, m);
    if (++m <= mlast) {
      body(Phase<2>{}, m);
      if (++m <= mlast) {
        body(Phase<3>{}, m);
        if (++m <= mlast) {
          for (;;) {
            body(Phase<4>{}, m);
            if (++m > mlast) break;
            body(Phase<5>{}, m);
            if (++m > mlast) break;
            body(Phase<6>{}, m);
            if (++m > mlast) break;
            body(Phase<7>{}, m);
            if (++m > mlast) break;
            body(Phase<0>{}, m);
            if (++m > mlast) break;
            body(Phase<1>{}, m);
            if (++m > mlast) break;
            body(Phase<2>{}, m);
            if (++m > mlast) break;
            body(Phase<3>{}, m);
            if (++m > mlast) break;
          }
        }
      }
    }
  }
  if (a.advance) advance_clock(a.clock, it, 1, a.nblocks);
}

// ------------------------------------------- two P-SV steps per launch
// Steps it and it+1 of P-SV without spin in one march.  A second pipeline
// stage trails the first by three rows and takes the stresses of step it+1
// from registers and LDS as the first stage produces them, so the state of
// step it+1 never goes to memory: each workgroup reads the state of step it
// once and writes that of step it+2 once, half the HBM traffic of two
// one-step launches.  The price is a 6-column halo (W = NT-12) and 4-5 rows of
// recomputation at each end of a segment.  Iteration m runs
//   stage 1 (step it):   velocity row m | barrier | stress row m-2
//   stage 2 (step it+1): velocity row m-3, D+z of velocity row m-4, stress row m-5
// Stage-2 z-neighbours go through LDS written after the barrier of one
// iteration and read after the barrier of the next.  Every cell's arithmetic
// is that of step_psv, operation for operation, so the result equals two
// one-step launches bit for bit.  The host passes the buffer to read in st[0]
// and the one to write in st[1].  Receivers are recorded in the march (REC)
// for launches of a few million shot-cells; larger launches leave them to
// k_record2 after the launch: the per-row receiver checks, cursors and
// displacement loads inside the march cost 4.5 % of the launch time on the
// 64-shot 2000x1000 workload, against a few microseconds for the extra launch.
template <int NT, int S, bool AUX, bool REC>
__device__ __forceinline__ void step_psv2(const StepArgs& a, int logical) {
  __shared__ float lds_s[2][2][S * NT];    // stage 1: sxz, szz of step it
  __shared__ float lds_v[2][2][S * NT];    // stage 1: vx, vz of step it+1
  __shared__ float lds_s2[2][2][S * NT];   // stage 2: sxz, szz of step it+1
  __shared__ float lds_v2[2][2][S * NT];   // stage 2: vx, vz of step it+2
  const TileGeom g = tile_geom<NT, S, 6>(a, logical);
  const int t = g.t, lane = g.lane, j = g.j, nx = a.nx, i0 = g.i0, i1 = g.i1;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int64_t so = int64_t(g.shot) * a.shot_floats;
  const uint32_t sbytes = 4u * uint32_t(a.shot_floats);
  const auto rin = make_rsrc(a.st[0] + so, sbytes);
  const auto rout = make_rsrc(a.st[1] + so, sbytes);
  const uint32_t npt4 = 4u * uint32_t(a.npt);
  const auto rmod = make_rsrc(a.model, uint32_t(MF_N) * npt4);
  // P-SV alone (the only physics two-step launches run): slots vx, vz, sxx,
  // szz, sxz in that order (the host checks), offsets as multiples of npt
  const uint32_t o_vx = 0u, o_vz = npt4;
  const uint32_t o_sxx = 2u * npt4, o_szz = 3u * npt4, o_sxz = 4u * npt4;
  const uint32_t m_rho = MF_RHO * npt4, m_lam = MF_LAM * npt4, m_mu = MF_MU * npt4, m_irho = MF_IRHO * npt4;
  const uint32_t jc4 = g.jc4;
  const int jc = int(jc4 >> 2);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const float dt = a.dt;
  const CDiv Dx{a.h8x, a.r8x, a.z8x, a.l8x, a.c_lo, (a.fast_div & FD_CONST) != 0},
             Dz{a.h8z, a.r8z, a.z8z, a.l8z, a.c_lo, (a.fast_div & FD_CONST) != 0};
  const int fast_c = div_mode(a.fast_div, FD_CONST);
  const int fast_rho = div_mode(a.fast_div, FD_RHO);
  const int tm2 = g.base + max(t - 2, 0), tm1 = g.base + max(t - 1, 0), tp1 = g.base + min(t + 1, NT - 1),
            tp2 = g.base + min(t + 2, NT - 1);
  const int tps = a.nstrips * a.nsegs;
  // one injection list per tile, walked by a cursor per stage
  int ip = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile + 1]);
  int ip2 = ip;
  int inj_row = next_row(a.inj, ip, ie);
  int inj_row2 = inj_row;
  // receivers recorded in the march (REC; otherwise k_record2 after the launch)
  [[maybe_unused]] int rp = 0, re = 0;
  if (REC && a.record) {
    rp = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile]);
    re = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile + 1]);
  }
  [[maybe_unused]] int rp2 = rp;
  [[maybe_unused]] int rec_row = REC ? next_row(a.rec, rp, re) : 0;
  [[maybe_unused]] int rec_row2 = rec_row;
  [[maybe_unused]] float* uacc = a.uacc + int64_t(g.shot) * a.nrec * N_UACC;
  [[maybe_unused]] float* obs = a.obs + int64_t(g.shot) * 5 * a.nrec * a.nt;
  [[maybe_unused]] const int nobs = a.nrec * a.nt;
  const int m0 = i0 - 4;   // first velocity row of stage 1
  // forward recompute of the gradient: the velocities v^{it+1} (stage 1) and
  // v^{it+2} (stage 2) into their time slots (step % aux_period)
  auto aux_rsrc = [&](int step) {
    return make_rsrc(
        AUX ? a.aux + (int64_t(step % a.aux_period) * a.nshots + g.shot) * NAUX_PSV * int64_t(a.npt) : a.model,
        AUX ? uint32_t(NAUX_PSV) * npt4 : 0u);
  };
  const auto raux1 = aux_rsrc(it + 1), raux2 = aux_rsrc(it + 2);

  // stage 1 rings: stresses of step it (rows m-2..m+1), new velocities
  // (rows m-3..m), D+z of the new velocities (row m at slot s2, used at s0)
  float sxx[4], sxz[4], szz[4];
  float vx[4] = {0, 0, 0, 0}, vz[4] = {0, 0, 0, 0};
  float dvxdz[4] = {0, 0, 0, 0}, dvzdz[4] = {0, 0, 0, 0};
  // stage 2 rings: stresses of step it+1 (rows m-5..m-2, row m-2 at s0),
  // velocities of step it+2 (rows m-6..m-3, row m-3 at s0), D+z of those
  // velocities (row m-4 at [par], row m-5 at [par^1]), and the model values /
  // taper factor stage 1 used three rows earlier (written at s0, read at s1)
  float t1xx[4] = {0, 0, 0, 0}, t1xz[4] = {0, 0, 0, 0}, t1zz[4] = {0, 0, 0, 0};
  float u2x[4] = {0, 0, 0, 0}, u2z[4] = {0, 0, 0, 0};
  float dz2x[2] = {0, 0}, dz2z[2] = {0, 0};
  float drho[4] = {1, 1, 1, 1}, dirho[4] = {1, 1, 1, 1}, dlam[4] = {0, 0, 0, 0}, dmu[4] = {0, 0, 0, 0};
  float dbx[4] = {0, 0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    const uint32_t r = g.row(m0 - 2 + q);
    sxx[q] = settle(bload(rin, jc4, o_sxx + r));
    sxz[q] = settle(bload(rin, jc4, o_sxz + r));
    szz[q] = settle(bload(rin, jc4, o_szz + r));
  }
  sxx[3] = sxz[3] = szz[3] = 0.0f;
  // D-deep prefetch FIFO, slot (m - m0) % D: stress rows m+1, velocity /
  // density rows m, moduli rows m-2 of iteration m
  // 2-deep for the forward (111 VGPRs + 9: still four waves per SIMD), 1-deep
  // for the recompute, whose velocity stores leave no registers for a second row
  constexpr int D = AUX ? 1 : 2;
  float pf_sxx[D], pf_sxz[D], pf_szz[D], pf_vx[D], pf_vz[D], pf_rho[D], pf_irho[D], pf_lam[D], pf_mu[D];
  float bx_next = taper_x(a.bx, m0, nx);
  // clamped row offsets, slot (row - m0) & 3: each iteration computes the
  // newest (m+D+1) and finds rows m+D and m+D-2 among the earlier ones
  uint32_t rofs[4];
  for (int k = D - 2; k <= D; ++k) rofs[k & 3] = g.row(m0 + k);
  for (int q = 0; q < D; ++q) {
    const uint32_t r1 = g.row(m0 + 1 + q), r0 = g.row(m0 + q), rm2 = g.row(m0 - 2 + q);
    pf_sxx[q] = bload(rin, jc4, o_sxx + r1);
    pf_sxz[q] = bload(rin, jc4, o_sxz + r1);
    pf_szz[q] = bload(rin, jc4, o_szz + r1);
    pf_vx[q] = bload(rin, jc4, o_vx + r0);
    pf_vz[q] = bload(rin, jc4, o_vz + r0);
    pf_rho[q] = bload(rmod, jc4, m_rho + r0);
    pf_irho[q] = bload(rmod, jc4, m_irho + r0);
    pf_lam[q] = bload(rmod, jc4, m_lam + rm2);
    pf_mu[q] = bload(rmod, jc4, m_mu + rm2);
  }
  lds_s[0][0][lane] = sxz[2];
  lds_s[0][1][lane] = szz[2];
  __syncthreads();

  const bool zm_ok = g.zm_ok, zp_ok = g.zp_ok;
  auto body = [&](auto phase, int m) __attribute__((always_inline)) {
    constexpr int R = decltype(phase)::value;
    constexpr int s0 = R & 3, s1 = (R + 1) & 3, s2 = (R + 2) & 3, s3 = (R + 3) & 3, par = R & 1;
    constexpr int f = R % D;
    sxx[s3] = settle(pf_sxx[f]); sxz[s3] = settle(pf_sxz[f]); szz[s3] = settle(pf_szz[f]);     // row m+1
    const float vx_o = pf_vx[f], vz_o = pf_vz[f], rho = pf_rho[f], irho = pf_irho[f], lam = pf_lam[f], mu = pf_mu[f];
    const float bxm = bx_next;
    bx_next = taper_x(a.bx, m + 1, nx);
    {
      constexpr int k1 = (R + D + 1) & 3, k0 = (R + D) & 3, km2 = (R + D + 2) & 3;
      rofs[k1] = g.row(m + D + 1);
      const uint32_t r3 = rofs[k1], r2 = rofs[k0], r0 = rofs[km2];
      pf_sxx[f] = bload(rin, jc4, o_sxx + r3);
      pf_sxz[f] = bload(rin, jc4, o_sxz + r3);
      pf_szz[f] = bload(rin, jc4, o_szz + r3);
      pf_vx[f] = bload(rin, jc4, o_vx + r2);
      pf_vz[f] = bload(rin, jc4, o_vz + r2);
      pf_rho[f] = bload(rmod, jc4, m_rho + r2);
      pf_irho[f] = bload(rmod, jc4, m_irho + r2);
      pf_lam[f] = bload(rmod, jc4, m_lam + r0);
      pf_mu[f] = bload(rmod, jc4, m_mu + r0);
    }

    // ---- stage 1: velocity of step it+1 at row m (not stored)
    {
      const bool xm_ok = m >= 2 && m < nx - 2;
      float qa[4];
      {
        const float n0 = nminus(sxx[s0], sxx[s1], sxx[s2], sxx[s3]);
        const float n1 = nminus(lds_s[par][0][tm2], lds_s[par][0][tm1], sxz[s2], lds_s[par][0][tp1]);
        const float n2 = nminus(sxz[s0], sxz[s1], sxz[s2], sxz[s3]);
        const float n3 = nminus(lds_s[par][1][tm2], lds_s[par][1][tm1], szz[s2], lds_s[par][1][tp1]);
        const float xa[4] = {keep(xm_ok, n0), keep(zm_ok, n1), keep(xm_ok, n2), keep(zm_ok, n3)};
        div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
      }
      float dsx = qa[0] + qa[1];
      float dsz = qa[2] + qa[3];
      if (m == inj_row) {
        for (; ip < ie; ++ip) {
          const int4 e = list_entry(a.inj, ip);
          if (e.x != m) break;
          if (e.y == j) {
            dsx = dsx + uniform_value(a.val_x, int64_t(e.z) * a.nt_val + it);
            dsz = dsz + uniform_value(a.val_z, int64_t(e.z) * a.nt_val + it);
          }
        }
        inj_row = next_row(a.inj, ip, ie);
      }
      const float b = (bxm * bzb) * bzt;
      float qb[2];
      const float xb[2] = {dt * dsx, dt * dsz};
      div_group(qb, xb, {Div{rho, irho, true}, Div{rho, irho, true}}, fast_rho);
      vx[s3] = b * (vx_o + qb[0]);
      vz[s3] = b * (vz_o + qb[1]);
      if constexpr (AUX) {
        if (g.jst && m >= i0 && m < i1) {
          const uint32_t ro = uint32_t(m) * g.rowb;
          bstore_out(vx[s3], raux1, jc4, A_VX * npt4 + ro, a.stream_out);
          bstore_out(vz[s3], raux1, jc4, A_VZ * npt4 + ro, a.stream_out);
        }
      }
    }
    lds_v[par][0][lane] = vx[s3];
    lds_v[par][1][lane] = vz[s3];
    lds_s[par ^ 1][0][lane] = sxz[s3];
    lds_s[par ^ 1][1][lane] = szz[s3];
    __syncthreads();

    // ---- stage 1: D+ of the new velocities (z at row m, x at row m-2)
    {
      const int i = m - 2;
      const bool xp_ok = i >= 1 && i < nx - 2;
      float qc[4];
      {
        const float n0 = nplus(lds_v[par][0][tm1], vx[s3], lds_v[par][0][tp1], lds_v[par][0][tp2]);
        const float n1 = nplus(lds_v[par][1][tm1], vz[s3], lds_v[par][1][tp1], lds_v[par][1][tp2]);
        const float n2 = nplus(vx[s0], vx[s1], vx[s2], vx[s3]);
        const float n3 = nplus(vz[s0], vz[s1], vz[s2], vz[s3]);
        const float xc[4] = {keep(zp_ok, n0), keep(zp_ok, n1), keep(xp_ok, n2), keep(xp_ok, n3)};
        div_group(qc, xc, {Dz, Dz, Dx, Dx}, fast_c);
      }
      dvxdz[s2] = qc[0];
      dvzdz[s2] = qc[1];
      // receivers of step it (row m)
      if constexpr (REC) {
        if (m == rec_row) {
          for (; rp < re; ++rp) {
            const int4 e = list_entry(a.rec, rp);
            if (e.x != m) break;
            if (e.y == j) {
              float* u = uacc + e.z * N_UACC;
              const float ux = u[U_UX] + vx[s3] * dt, uz = u[U_UZ] + vz[s3] * dt;
              u[U_UX] = ux;
              u[U_UZ] = uz;
              obs[0 * nobs + e.z * a.nt + it] = ux;
              obs[2 * nobs + e.z * a.nt + it] = uz;
            }
          }
          rec_row = next_row(a.rec, rp, re);
        }
      }
      // stress of step it+1 at row m-2: into the stage-2 ring and LDS
      const float dvxdx = qc[2];
      const float dvzdx = qc[3];
      const float l2m = lam + 2.0f * mu;
      t1xx[s0] = sxx[s0] + dt * (l2m * dvxdx + lam * dvzdz[s0]);
      t1zz[s0] = szz[s0] + dt * (l2m * dvzdz[s0] + lam * dvxdx);
      t1xz[s0] = sxz[s0] + dt * mu * (dvxdz[s0] + dvzdx);
      lds_s2[par][0][lane] = t1xz[s0];
      lds_s2[par][1][lane] = t1zz[s0];
    }
    drho[s0] = rho;
    dirho[s0] = irho;
    dlam[s0] = lam;
    dmu[s0] = mu;
    dbx[s0] = bxm;

    // ---- stage 2: velocity of step it+2 at row m-3
    {
      const int m2 = m - 3;
      const bool xm_ok = m2 >= 2 && m2 < nx - 2;
      float qa[4];
      {
        const float n0 = nminus(t1xx[s1], t1xx[s2], t1xx[s3], t1xx[s0]);
        const float n1 = nminus(lds_s2[par ^ 1][0][tm2], lds_s2[par ^ 1][0][tm1], t1xz[s3], lds_s2[par ^ 1][0][tp1]);
        const float n2 = nminus(t1xz[s1], t1xz[s2], t1xz[s3], t1xz[s0]);
        const float n3 = nminus(lds_s2[par ^ 1][1][tm2], lds_s2[par ^ 1][1][tm1], t1zz[s3], lds_s2[par ^ 1][1][tp1]);
        const float xa[4] = {keep(xm_ok, n0), keep(zm_ok, n1), keep(xm_ok, n2), keep(zm_ok, n3)};
        div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
      }
      float dsx = qa[0] + qa[1];
      float dsz = qa[2] + qa[3];
      if (m2 == inj_row2) {
        for (; ip2 < ie; ++ip2) {
          const int4 e = list_entry(a.inj, ip2);
          if (e.x != m2) break;
          if (e.y == j) {
            dsx = dsx + uniform_value(a.val_x, int64_t(e.z) * a.nt_val + it + 1);
            dsz = dsz + uniform_value(a.val_z, int64_t(e.z) * a.nt_val + it + 1);
          }
        }
        inj_row2 = next_row(a.inj, ip2, ie);
      }
      const float b = (dbx[s1] * bzb) * bzt;
      float qb[2];
      const float xb[2] = {dt * dsx, dt * dsz};
      div_group(qb, xb, {Div{drho[s1], dirho[s1], true}, Div{drho[s1], dirho[s1], true}}, fast_rho);
      const float vxn = b * (vx[s0] + qb[0]);
      const float vzn = b * (vz[s0] + qb[1]);
      if (g.jst && m2 >= i0 && m2 < i1) {
        const uint32_t ro = uint32_t(m2) * g.rowb;
        bstore_out(vxn, rout, jc4, o_vx + ro, a.stream_out);
        bstore_out(vzn, rout, jc4, o_vz + ro, a.stream_out);
        if constexpr (AUX) {
          bstore_out(vxn, raux2, jc4, A_VX * npt4 + ro, a.stream_out);
          bstore_out(vzn, raux2, jc4, A_VZ * npt4 + ro, a.stream_out);
        }
      }
      u2x[s0] = vxn;
      u2z[s0] = vzn;
      // receivers of step it+1 (row m-3)
      if constexpr (REC) {
        if (m2 == rec_row2) {
          for (; rp2 < re; ++rp2) {
            const int4 e = list_entry(a.rec, rp2);
            if (e.x != m2) break;
            if (e.y == j) {
              float* u = uacc + e.z * N_UACC;
              const float ux = u[U_UX] + vxn * dt, uz = u[U_UZ] + vzn * dt;
              u[U_UX] = ux;
              u[U_UZ] = uz;
              obs[0 * nobs + e.z * a.nt + it + 1] = ux;
              obs[2 * nobs + e.z * a.nt + it + 1] = uz;
            }
          }
          rec_row2 = next_row(a.rec, rp2, re);
        }
      }
    }

    // ---- stage 2: D+z of velocity row m-4 (LDS of the previous iteration),
    // D+x at row m-5, stress of step it+2 at row m-5
    {
      const int i2 = m - 5;
      const bool xp_ok = i2 >= 1 && i2 < nx - 2;
      float qc[4];
      {
        const float n0 = nplus(lds_v2[par ^ 1][0][tm1], u2x[s3], lds_v2[par ^ 1][0][tp1], lds_v2[par ^ 1][0][tp2]);
        const float n1 = nplus(lds_v2[par ^ 1][1][tm1], u2z[s3], lds_v2[par ^ 1][1][tp1], lds_v2[par ^ 1][1][tp2]);
        const float n2 = nplus(u2x[s1], u2x[s2], u2x[s3], u2x[s0]);
        const float n3 = nplus(u2z[s1], u2z[s2], u2z[s3], u2z[s0]);
        const float xc[4] = {keep(zp_ok, n0), keep(zp_ok, n1), keep(xp_ok, n2), keep(xp_ok, n3)};
        div_group(qc, xc, {Dz, Dz, Dx, Dx}, fast_c);
      }
      dz2x[par] = qc[0];
      dz2z[par] = qc[1];
      const float lam2 = dlam[s1], mu2 = dmu[s1];
      const float dvxdx = qc[2];
      const float dvzdx = qc[3];
      const float l2m = lam2 + 2.0f * mu2;
      const float nsxx = t1xx[s1] + dt * (l2m * dvxdx + lam2 * dz2z[par ^ 1]);
      const float nszz = t1zz[s1] + dt * (l2m * dz2z[par ^ 1] + lam2 * dvxdx);
      const float nsxz = t1xz[s1] + dt * mu2 * (dz2x[par ^ 1] + dvzdx);
      if (g.jst && i2 >= i0 && i2 < i1) {
        const uint32_t ro = uint32_t(i2) * g.rowb;
        bstore_out(nsxx, rout, jc4, o_sxx + ro, a.stream_out);
        bstore_out(nszz, rout, jc4, o_szz + ro, a.stream_out);
        bstore_out(nsxz, rout, jc4, o_sxz + ro, a.stream_out);
      }
    }
    lds_v2[par][0][lane] = u2x[s0];
    lds_v2[par][1][lane] = u2z[s0];
  };

  // four peeled rows, then 4 phases per loop trip (with a 1-deep FIFO every
  // ring repeats with period 4; half the code of an 8-phase trip)
  static_assert(D == 1 || D == 2, "the 4-phase loop needs prefetch slots that repeat with a period dividing 4");
  const int mlast = i1 + 4;
  int m = m0;
  body(Phase<0>{}, m);
  if (++m <= mlast) {
    body(Phase<1>{}, m);
    if (++m <= mlast) {
      body(Phase<2>{}, m);
      if (++m <= mlast) {
        body(Phase<3>{}, m);
        if (++m <= mlast) {
          for (;;) {
            body(Phase<4>{}, m);
            if (++m > mlast) break;
            body(Phase<5>{}, m);
            if (++m > mlast) break;
            body(Phase<6>{}, m);
            if (++m > mlast) break;
            body(Phase<7>{}, m);
            if (++m > mlast) break;
          }
        }
      }
    }
  }
  if (a.advance) advance_clock(a.clock, it, 2, a.nblocks);
}

// ----------------------------------------- three P-SV steps per launch
// Steps it, it+1 and it+2 of P-SV without spin in one march: the two stages
// of step_psv2 and a third trailing the second by three rows, so neither the
// state of step it+1 nor that of step it+2 goes to memory: a workgroup reads
// the state of step it once and writes that of step it+3 once (a third of
// the state traffic of three one-step launches).  The price is a 9-column
// halo (W = NT-18) and 7 + 8 recomputed rows per segment.  Iteration m runs
//   stage 1 (step it):   velocity row m | barrier | stress row m-2
//   stage 2 (step it+1): velocity row m-3, D+z of velocity row m-4, stress row m-5
//   stage 3 (step it+2): velocity row m-6, D+z of velocity row m-7, stress row m-8
// Stage-2 and stage-3 z-neighbours go through LDS written after the barrier
// of one iteration and read after the barrier of the next; model values and
// taper factors ride along in register rings (stage 3 takes what stage 2
// used three rows earlier).  Every cell's arithmetic is that of step_psv,
// operation for operation, so the result equals three one-step launches bit
// for bit.  Reads st[0], writes st[1] (the buffer of the other parity).  No
// gradient variant: the recompute of the gradient runs two-step launches.
template <int NT, int S>
__device__ __forceinline__ void step_psv3(const StepArgs& a, int logical) {
  __shared__ float lds_s[2][2][S * NT];    // stage 1: sxz, szz of step it
  __shared__ float lds_v[2][2][S * NT];    // stage 1: vx, vz of step it+1
  __shared__ float lds_s2[2][2][S * NT];   // stage 2: sxz, szz of step it+1
  __shared__ float lds_v2[2][2][S * NT];   // stage 2: vx, vz of step it+2
  __shared__ float lds_s3[2][2][S * NT];   // stage 3: sxz, szz of step it+2
  __shared__ float lds_v3[2][2][S * NT];   // stage 3: vx, vz of step it+3
  const TileGeom g = tile_geom<NT, S, 9>(a, logical);
  const int t = g.t, lane = g.lane, j = g.j, nx = a.nx, i0 = g.i0, i1 = g.i1;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int64_t so = int64_t(g.shot) * a.shot_floats;
  const uint32_t sbytes = 4u * uint32_t(a.shot_floats);
  const auto rin = make_rsrc(a.st[0] + so, sbytes);
  const auto rout = make_rsrc(a.st[1] + so, sbytes);
  const uint32_t npt4 = 4u * uint32_t(a.npt);
  const auto rmod = make_rsrc(a.model, uint32_t(MF_N) * npt4);
  // P-SV alone: slots vx, vz, sxx, szz, sxz in that order (the host checks)
  const uint32_t o_vx = 0u, o_vz = npt4;
  const uint32_t o_sxx = 2u * npt4, o_szz = 3u * npt4, o_sxz = 4u * npt4;
  const uint32_t m_rho = MF_RHO * npt4, m_lam = MF_LAM * npt4, m_mu = MF_MU * npt4, m_irho = MF_IRHO * npt4;
  const uint32_t jc4 = g.jc4;
  const int jc = int(jc4 >> 2);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const float dt = a.dt;
  const CDiv Dx{a.h8x, a.r8x, a.z8x, a.l8x, a.c_lo, (a.fast_div & FD_CONST) != 0},
             Dz{a.h8z, a.r8z, a.z8z, a.l8z, a.c_lo, (a.fast_div & FD_CONST) != 0};
  const int fast_c = div_mode(a.fast_div, FD_CONST);
  const int fast_rho = div_mode(a.fast_div, FD_RHO);
  const int tm2 = g.base + max(t - 2, 0), tm1 = g.base + max(t - 1, 0), tp1 = g.base + min(t + 1, NT - 1),
            tp2 = g.base + min(t + 2, NT - 1);
  const int tps = a.nstrips * a.nsegs;
  // one injection list per tile, walked by a cursor per stage
  int ip = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile + 1]);
  int ip2 = ip, ip3 = ip;
  int inj_row = next_row(a.inj, ip, ie);
  int inj_row2 = inj_row, inj_row3 = inj_row;
  int rp = 0, re = 0;
  if (a.record) {
    rp = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile]);
    re = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile + 1]);
  }
  int rp2 = rp, rp3 = rp;
  int rec_row = next_row(a.rec, rp, re);
  int rec_row2 = rec_row, rec_row3 = rec_row;
  float* uacc = a.uacc + int64_t(g.shot) * a.nrec * N_UACC;
  float* obs = a.obs + int64_t(g.shot) * 5 * a.nrec * a.nt;
  const int nobs = a.nrec * a.nt;
  const int m0 = i0 - 7;   // first velocity row of stage 1

  // stage 1 rings: stresses of step it (rows m-2..m+1), new velocities
  // (rows m-3..m), D+z of the new velocities (row m at slot s2, used at s0)
  float sxx[4], sxz[4], szz[4];
  float vx[4] = {0, 0, 0, 0}, vz[4] = {0, 0, 0, 0};
  float dvxdz[4] = {0, 0, 0, 0}, dvzdz[4] = {0, 0, 0, 0};
  // stage 2 rings: stresses of step it+1 (row m-2 at s0), velocities of step
  // it+2 (row m-3 at s0), D+z of those velocities (row m-4 at [par], m-5 at
  // [par^1]), model values / taper factor stage 1 used (written at s0, read at s1)
  float t1xx[4] = {0, 0, 0, 0}, t1xz[4] = {0, 0, 0, 0}, t1zz[4] = {0, 0, 0, 0};
  float u2x[4] = {0, 0, 0, 0}, u2z[4] = {0, 0, 0, 0};
  float dz2x[2] = {0, 0}, dz2z[2] = {0, 0};
  float drho[4] = {1, 1, 1, 1}, dirho[4] = {1, 1, 1, 1}, dlam[4] = {0, 0, 0, 0}, dmu[4] = {0, 0, 0, 0};
  float dbx[4] = {0, 0, 0, 0};
  // stage 3 rings: stresses of step it+2 (row m-5 at s0), velocities of step
  // it+3 (row m-6 at s0), D+z of those (row m-7 at [par], m-8 at [par^1]),
  // the model values / taper factor stage 2 used (written at s0, read at s1)
  float t2xx[4] = {0, 0, 0, 0}, t2xz[4] = {0, 0, 0, 0}, t2zz[4] = {0, 0, 0, 0};
  float u3x[4] = {0, 0, 0, 0}, u3z[4] = {0, 0, 0, 0};
  float dz3x[2] = {0, 0}, dz3z[2] = {0, 0};
  float erho[4] = {1, 1, 1, 1}, eirho[4] = {1, 1, 1, 1}, elam[4] = {0, 0, 0, 0}, emu[4] = {0, 0, 0, 0};
  float ebx[4] = {0, 0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    const uint32_t r = g.row(m0 - 2 + q);
    sxx[q] = settle(bload(rin, jc4, o_sxx + r));
    sxz[q] = settle(bload(rin, jc4, o_sxz + r));
    szz[q] = settle(bload(rin, jc4, o_szz + r));
  }
  sxx[3] = sxz[3] = szz[3] = 0.0f;
  // 1-deep prefetch: stress rows m+1, velocity / density rows m, moduli rows
  // m-2 of iteration m
  float pf_sxx, pf_sxz, pf_szz, pf_vx, pf_vz, pf_rho, pf_irho, pf_lam, pf_mu;
  float bx_next = taper_x(a.bx, m0, nx);
  // clamped row offsets, slot (row - m0) & 3
  uint32_t rofs[4];
  for (int k = -1; k <= 1; ++k) rofs[k & 3] = g.row(m0 + k);
  {
    const uint32_t r1 = g.row(m0 + 1), r0 = g.row(m0), rm2 = g.row(m0 - 2);
    pf_sxx = bload(rin, jc4, o_sxx + r1);
    pf_sxz = bload(rin, jc4, o_sxz + r1);
    pf_szz = bload(rin, jc4, o_szz + r1);
    pf_vx = bload(rin, jc4, o_vx + r0);
    pf_vz = bload(rin, jc4, o_vz + r0);
    pf_rho = bload(rmod, jc4, m_rho + r0);
    pf_irho = bload(rmod, jc4, m_irho + r0);
    pf_lam = bload(rmod, jc4, m_lam + rm2);
    pf_mu = bload(rmod, jc4, m_mu + rm2);
  }
  lds_s[0][0][lane] = sxz[2];
  lds_s[0][1][lane] = szz[2];
  __syncthreads();

  const bool zm_ok = g.zm_ok, zp_ok = g.zp_ok;
  auto body = [&](auto phase, int m) __attribute__((always_inline)) {
    constexpr int R = decltype(phase)::value;
    constexpr int s0 = R & 3, s1 = (R + 1) & 3, s2 = (R + 2) & 3, s3 = (R + 3) & 3, par = R & 1;
    sxx[s3] = settle(pf_sxx); sxz[s3] = settle(pf_sxz); szz[s3] = settle(pf_szz);     // row m+1
    const float vx_o = pf_vx, vz_o = pf_vz, rho = pf_rho, irho = pf_irho, lam = pf_lam, mu = pf_mu;
    const float bxm = bx_next;
    bx_next = taper_x(a.bx, m + 1, nx);
    {
      constexpr int k1 = (R + 2) & 3, k0 = (R + 1) & 3, km2 = (R + 3) & 3;
      rofs[k1] = g.row(m + 2);
      const uint32_t r3 = rofs[k1], r2 = rofs[k0], r0 = rofs[km2];
      pf_sxx = bload(rin, jc4, o_sxx + r3);
      pf_sxz = bload(rin, jc4, o_sxz + r3);
      pf_szz = bload(rin, jc4, o_szz + r3);
      pf_vx = bload(rin, jc4, o_vx + r2);
      pf_vz = bload(rin, jc4, o_vz + r2);
      pf_rho = bload(rmod, jc4, m_rho + r2);
      pf_irho = bload(rmod, jc4, m_irho + r2);
      pf_lam = bload(rmod, jc4, m_lam + r0);
      pf_mu = bload(rmod, jc4, m_mu + r0);
    }

    // ---- stage 1: velocity of step it+1 at row m (not stored)
    {
      const bool xm_ok = m >= 2 && m < nx - 2;
      float qa[4];
      {
        const float n0 = nminus(sxx[s0], sxx[s1], sxx[s2], sxx[s3]);
        const float n1 = nminus(lds_s[par][0][tm2], lds_s[par][0][tm1], sxz[s2], lds_s[par][0][tp1]);
        const float n2 = nminus(sxz[s0], sxz[s1], sxz[s2], sxz[s3]);
        const float n3 = nminus(lds_s[par][1][tm2], lds_s[par][1][tm1], szz[s2], lds_s[par][1][tp1]);
        const float xa[4] = {keep(xm_ok, n0), keep(zm_ok, n1), keep(xm_ok, n2), keep(zm_ok, n3)};
        div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
      }
      float dsx = qa[0] + qa[1];
      float dsz = qa[2] + qa[3];
      if (m == inj_row) {
        for (; ip < ie; ++ip) {
          const int4 e = list_entry(a.inj, ip);
          if (e.x != m) break;
          if (e.y == j) {
            dsx = dsx + uniform_value(a.val_x, int64_t(e.z) * a.nt_val + it);
            dsz = dsz + uniform_value(a.val_z, int64_t(e.z) * a.nt_val + it);
          }
        }
        inj_row = next_row(a.inj, ip, ie);
      }
      const float b = (bxm * bzb) * bzt;
      float qb[2];
      const float xb[2] = {dt * dsx, dt * dsz};
      div_group(qb, xb, {Div{rho, irho, true}, Div{rho, irho, true}}, fast_rho);
      vx[s3] = b * (vx_o + qb[0]);
      vz[s3] = b * (vz_o + qb[1]);
    }
    lds_v[par][0][lane] = vx[s3];
    lds_v[par][1][lane] = vz[s3];
    lds_s[par ^ 1][0][lane] = sxz[s3];
    lds_s[par ^ 1][1][lane] = szz[s3];
    __syncthreads();

    // ---- stage 1: D+ of the new velocities (z at row m, x at row m-2)
    {
      const int i = m - 2;
      const bool xp_ok = i >= 1 && i < nx - 2;
      float qc[4];
      {
        const float n0 = nplus(lds_v[par][0][tm1], vx[s3], lds_v[par][0][tp1], lds_v[par][0][tp2]);
        const float n1 = nplus(lds_v[par][1][tm1], vz[s3], lds_v[par][1][tp1], lds_v[par][1][tp2]);
        const float n2 = nplus(vx[s0], vx[s1], vx[s2], vx[s3]);
        const float n3 = nplus(vz[s0], vz[s1], vz[s2], vz[s3]);
        const float xc[4] = {keep(zp_ok, n0), keep(zp_ok, n1), keep(xp_ok, n2), keep(xp_ok, n3)};
        div_group(qc, xc, {Dz, Dz, Dx, Dx}, fast_c);
      }
      dvxdz[s2] = qc[0];
      dvzdz[s2] = qc[1];
      // receivers of step it (row m)
      if (m == rec_row) {
        for (; rp < re; ++rp) {
          const int4 e = list_entry(a.rec, rp);
          if (e.x != m) break;
          if (e.y == j) {
            float* u = uacc + e.z * N_UACC;
            const float ux = u[U_UX] + vx[s3] * dt, uz = u[U_UZ] + vz[s3] * dt;
            u[U_UX] = ux;
            u[U_UZ] = uz;
            obs[0 * nobs + e.z * a.nt + it] = ux;
            obs[2 * nobs + e.z * a.nt + it] = uz;
          }
        }
        rec_row = next_row(a.rec, rp, re);
      }
      // stress of step it+1 at row m-2: into the stage-2 ring and LDS
      const float dvxdx = qc[2];
      const float dvzdx = qc[3];
      const float l2m = lam + 2.0f * mu;
      t1xx[s0] = sxx[s0] + dt * (l2m * dvxdx + lam * dvzdz[s0]);
      t1zz[s0] = szz[s0] + dt * (l2m * dvzdz[s0] + lam * dvxdx);
      t1xz[s0] = sxz[s0] + dt * mu * (dvxdz[s0] + dvzdx);
      lds_s2[par][0][lane] = t1xz[s0];
      lds_s2[par][1][lane] = t1zz[s0];
    }
    // the model values stage 2 used three rows ago go on to stage 3
    erho[s0] = drho[s1];
    eirho[s0] = dirho[s1];
    elam[s0] = dlam[s1];
    emu[s0] = dmu[s1];
    ebx[s0] = dbx[s1];
    drho[s0] = rho;
    dirho[s0] = irho;
    dlam[s0] = lam;
    dmu[s0] = mu;
    dbx[s0] = bxm;

    // ---- stage 2: velocity of step it+2 at row m-3 (not stored)
    {
      const int m2 = m - 3;
      const bool xm_ok = m2 >= 2 && m2 < nx - 2;
      float qa[4];
      {
        const float n0 = nminus(t1xx[s1], t1xx[s2], t1xx[s3], t1xx[s0]);
        const float n1 = nminus(lds_s2[par ^ 1][0][tm2], lds_s2[par ^ 1][0][tm1], t1xz[s3], lds_s2[par ^ 1][0][tp1]);
        const float n2 = nminus(t1xz[s1], t1xz[s2], t1xz[s3], t1xz[s0]);
        const float n3 = nminus(lds_s2[par ^ 1][1][tm2], lds_s2[par ^ 1][1][tm1], t1zz[s3], lds_s2[par ^ 1][1][tp1]);
        const float xa[4] = {keep(xm_ok, n0), keep(zm_ok, n1), keep(xm_ok, n2), keep(zm_ok, n3)};
        div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
      }
      float dsx = qa[0] + qa[1];
      float dsz = qa[2] + qa[3];
      if (m2 == inj_row2) {
        for (; ip2 < ie; ++ip2) {
          const int4 e = list_entry(a.inj, ip2);
          if (e.x != m2) break;
          if (e.y == j) {
            dsx = dsx + uniform_value(a.val_x, int64_t(e.z) * a.nt_val + it + 1);
            dsz = dsz + uniform_value(a.val_z, int64_t(e.z) * a.nt_val + it + 1);
          }
        }
        inj_row2 = next_row(a.inj, ip2, ie);
      }
      const float b = (dbx[s1] * bzb) * bzt;
      float qb[2];
      const float xb[2] = {dt * dsx, dt * dsz};
      div_group(qb, xb, {Div{drho[s1], dirho[s1], true}, Div{drho[s1], dirho[s1], true}}, fast_rho);
      const float vxn = b * (vx[s0] + qb[0]);
      const float vzn = b * (vz[s0] + qb[1]);
      u2x[s0] = vxn;
      u2z[s0] = vzn;
      // receivers of step it+1 (row m-3)
      if (m2 == rec_row2) {
        for (; rp2 < re; ++rp2) {
          const int4 e = list_entry(a.rec, rp2);
          if (e.x != m2) break;
          if (e.y == j) {
            float* u = uacc + e.z * N_UACC;
            const float ux = u[U_UX] + vxn * dt, uz = u[U_UZ] + vzn * dt;
            u[U_UX] = ux;
            u[U_UZ] = uz;
            obs[0 * nobs + e.z * a.nt + it + 1] = ux;
            obs[2 * nobs + e.z * a.nt + it + 1] = uz;
          }
        }
        rec_row2 = next_row(a.rec, rp2, re);
      }
    }

    // ---- stage 2: D+z of velocity row m-4, D+x at row m-5, stress of step
    // it+2 at row m-5: into the stage-3 ring and LDS
    {
      const int i2 = m - 5;
      const bool xp_ok = i2 >= 1 && i2 < nx - 2;
      float qc[4];
      {
        const float n0 = nplus(lds_v2[par ^ 1][0][tm1], u2x[s3], lds_v2[par ^ 1][0][tp1], lds_v2[par ^ 1][0][tp2]);
        const float n1 = nplus(lds_v2[par ^ 1][1][tm1], u2z[s3], lds_v2[par ^ 1][1][tp1], lds_v2[par ^ 1][1][tp2]);
        const float n2 = nplus(u2x[s1], u2x[s2], u2x[s3], u2x[s0]);
        const float n3 = nplus(u2z[s1], u2z[s2], u2z[s3], u2z[s0]);
        const float xc[4] = {keep(zp_ok, n0), keep(zp_ok, n1), keep(xp_ok, n2), keep(xp_ok, n3)};
        div_group(qc, xc, {Dz, Dz, Dx, Dx}, fast_c);
      }
      dz2x[par] = qc[0];
      dz2z[par] = qc[1];
      const float lam2 = dlam[s1], mu2 = dmu[s1];
      const float dvxdx = qc[2];
      const float dvzdx = qc[3];
      const float l2m = lam2 + 2.0f * mu2;
      t2xx[s0] = t1xx[s1] + dt * (l2m * dvxdx + lam2 * dz2z[par ^ 1]);
      t2zz[s0] = t1zz[s1] + dt * (l2m * dz2z[par ^ 1] + lam2 * dvxdx);
      t2xz[s0] = t1xz[s1] + dt * mu2 * (dz2x[par ^ 1] + dvzdx);
      lds_s3[par][0][lane] = t2xz[s0];
      lds_s3[par][1][lane] = t2zz[s0];
    }
    lds_v2[par][0][lane] = u2x[s0];
    lds_v2[par][1][lane] = u2z[s0];

    // ---- stage 3: velocity of step it+3 at row m-6
    {
      const int m3 = m - 6;
      const bool xm_ok = m3 >= 2 && m3 < nx - 2;
      float qa[4];
      {
        const float n0 = nminus(t2xx[s1], t2xx[s2], t2xx[s3], t2xx[s0]);
        const float n1 = nminus(lds_s3[par ^ 1][0][tm2], lds_s3[par ^ 1][0][tm1], t2xz[s3], lds_s3[par ^ 1][0][tp1]);
        const float n2 = nminus(t2xz[s1], t2xz[s2], t2xz[s3], t2xz[s0]);
        const float n3 = nminus(lds_s3[par ^ 1][1][tm2], lds_s3[par ^ 1][1][tm1], t2zz[s3], lds_s3[par ^ 1][1][tp1]);
        const float xa[4] = {keep(xm_ok, n0), keep(zm_ok, n1), keep(xm_ok, n2), keep(zm_ok, n3)};
        div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
      }
      float dsx = qa[0] + qa[1];
      float dsz = qa[2] + qa[3];
      if (m3 == inj_row3) {
        for (; ip3 < ie; ++ip3) {
          const int4 e = list_entry(a.inj, ip3);
          if (e.x != m3) break;
          if (e.y == j) {
            dsx = dsx + uniform_value(a.val_x, int64_t(e.z) * a.nt_val + it + 2);
            dsz = dsz + uniform_value(a.val_z, int64_t(e.z) * a.nt_val + it + 2);
          }
        }
        inj_row3 = next_row(a.inj, ip3, ie);
      }
      const float b = (ebx[s1] * bzb) * bzt;
      float qb[2];
      const float xb[2] = {dt * dsx, dt * dsz};
      div_group(qb, xb, {Div{erho[s1], eirho[s1], true}, Div{erho[s1], eirho[s1], true}}, fast_rho);
      const float vxn = b * (u2x[s1] + qb[0]);
      const float vzn = b * (u2z[s1] + qb[1]);
      if (g.jst && m3 >= i0 && m3 < i1) {
        const uint32_t ro = uint32_t(m3) * g.rowb;
        bstore_out(vxn, rout, jc4, o_vx + ro, a.stream_out);
        bstore_out(vzn, rout, jc4, o_vz + ro, a.stream_out);
      }
      u3x[s0] = vxn;
      u3z[s0] = vzn;
      // receivers of step it+2 (row m-6)
      if (m3 == rec_row3) {
        for (; rp3 < re; ++rp3) {
          const int4 e = list_entry(a.rec, rp3);
          if (e.x != m3) break;
          if (e.y == j) {
            float* u = uacc + e.z * N_UACC;
            const float ux = u[U_UX] + vxn * dt, uz = u[U_UZ] + vzn * dt;
            u[U_UX] = ux;
            u[U_UZ] = uz;
            obs[0 * nobs + e.z * a.nt + it + 2] = ux;
            obs[2 * nobs + e.z * a.nt + it + 2] = uz;
          }
        }
        rec_row3 = next_row(a.rec, rp3, re);
      }
    }

    // ---- stage 3: D+z of velocity row m-7 (LDS of the previous iteration),
    // D+x at row m-8, stress of step it+3 at row m-8
    {
      const int i3 = m - 8;
      const bool xp_ok = i3 >= 1 && i3 < nx - 2;
      float qc[4];
      {
        const float n0 = nplus(lds_v3[par ^ 1][0][tm1], u3x[s3], lds_v3[par ^ 1][0][tp1], lds_v3[par ^ 1][0][tp2]);
        const float n1 = nplus(lds_v3[par ^ 1][1][tm1], u3z[s3], lds_v3[par ^ 1][1][tp1], lds_v3[par ^ 1][1][tp2]);
        const float n2 = nplus(u3x[s1], u3x[s2], u3x[s3], u3x[s0]);
        const float n3 = nplus(u3z[s1], u3z[s2], u3z[s3], u3z[s0]);
        const float xc[4] = {keep(zp_ok, n0), keep(zp_ok, n1), keep(xp_ok, n2), keep(xp_ok, n3)};
        div_group(qc, xc, {Dz, Dz, Dx, Dx}, fast_c);
      }
      dz3x[par] = qc[0];
      dz3z[par] = qc[1];
      const float lam3 = elam[s1], mu3 = emu[s1];
      const float dvxdx = qc[2];
      const float dvzdx = qc[3];
      const float l2m = lam3 + 2.0f * mu3;
      const float nsxx = t2xx[s1] + dt * (l2m * dvxdx + lam3 * dz3z[par ^ 1]);
      const float nszz = t2zz[s1] + dt * (l2m * dz3z[par ^ 1] + lam3 * dvxdx);
      const float nsxz = t2xz[s1] + dt * mu3 * (dz3x[par ^ 1] + dvzdx);
      if (g.jst && i3 >= i0 && i3 < i1) {
        const uint32_t ro = uint32_t(i3) * g.rowb;
        bstore_out(nsxx, rout, jc4, o_sxx + ro, a.stream_out);
        bstore_out(nszz, rout, jc4, o_szz + ro, a.stream_out);
        bstore_out(nsxz, rout, jc4, o_sxz + ro, a.stream_out);
      }
    }
    lds_v3[par][0][lane] = u3x[s0];
    lds_v3[par][1][lane] = u3z[s0];
  };

  // four peeled rows, then 4 phases per loop trip (every ring repeats with
  // period 4)
  const int mlast = i1 + 7;
  int m = m0;
  body(Phase<0>{}, m);
  if (++m <= mlast) {
    body(Phase<1>{}, m);
    if (++m <= mlast) {
      body(Phase<2>{}, m);
      if (++m <= mlast) {
        body(Phase<3>{}, m);
        if (++m <= mlast) {
          for (;;) {
            body(Phase<4>{}, m);
            if (++m > mlast) break;
            body(Phase<5>{}, m);
            if (++m > mlast) break;
            body(Phase<6>{}, m);
            if (++m > mlast) break;
            body(Phase<7>{}, m);
            if (++m > mlast) break;
          }
        }
      }
    }
  }
  if (a.advance) advance_clock(a.clock, it, 3, a.nblocks);
}

// ------------------------------------------------- two SH steps per launch
// Steps it and it+1 of SH in one march, pipelined as step_psv2 (stage 2
// trails stage 1 by three rows; the state of step it+1 stays in registers
// and LDS); every cell's arithmetic is that of step_sh.  The host passes the
// buffer to read in st[0] and the one to write in st[1].
template <int NT, int S, bool AUX, bool REC>
__device__ __forceinline__ void step_sh2(const StepArgs& a, int logical) {
  __shared__ float lds_s[2][S * NT];    // stage 1: szy of step it
  __shared__ float lds_v[2][S * NT];    // stage 1: vy of step it+1
  __shared__ float lds_s2[2][S * NT];   // stage 2: szy of step it+1
  __shared__ float lds_v2[2][S * NT];   // stage 2: vy of step it+2
  const TileGeom g = tile_geom<NT, S, 6>(a, logical);
  const int t = g.t, lane = g.lane, j = g.j, nx = a.nx, i0 = g.i0, i1 = g.i1;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int64_t so = int64_t(g.shot) * a.shot_floats;
  const uint32_t sbytes = 4u * uint32_t(a.shot_floats);
  const auto rin = make_rsrc(a.st[0] + so, sbytes);
  const auto rout = make_rsrc(a.st[1] + so, sbytes);
  const uint32_t npt4 = 4u * uint32_t(a.npt);
  const auto rmod = make_rsrc(a.model, uint32_t(MF_N) * npt4);
  // SH alone (the only physics two-step launches run): slots vy, sxy, szy in
  // that order (the host checks), offsets as multiples of npt
  const uint32_t o_vy = 0u, o_sxy = npt4, o_szy = 2u * npt4;
  const uint32_t m_rho = MF_RHO * npt4, m_mu = MF_MU * npt4, m_irho = MF_IRHO * npt4;
  const uint32_t jc4 = g.jc4;
  const int jc = int(jc4 >> 2);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const float dt = a.dt;
  const CDiv Dx{a.h8x, a.r8x, a.z8x, a.l8x, a.c_lo, (a.fast_div & FD_CONST) != 0},
             Dz{a.h8z, a.r8z, a.z8z, a.l8z, a.c_lo, (a.fast_div & FD_CONST) != 0};
  const int fast_c = div_mode(a.fast_div, FD_CONST);
  const int fast_rho = div_mode(a.fast_div, FD_RHO);
  const int tm2 = g.base + max(t - 2, 0), tm1 = g.base + max(t - 1, 0), tp1 = g.base + min(t + 1, NT - 1),
            tp2 = g.base + min(t + 2, NT - 1);
  const int tps = a.nstrips * a.nsegs;
  int ip = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile + 1]);
  int ip2 = ip;
  int inj_row = next_row(a.inj, ip, ie);
  int inj_row2 = inj_row;
  // receivers recorded in the march (REC; otherwise k_record2 after the launch)
  [[maybe_unused]] int rp = 0, re = 0;
  if (REC && a.record) {
    rp = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile]);
    re = __builtin_amdgcn_readfirstlane(a.rec_off[g.tile + 1]);
  }
  [[maybe_unused]] int rp2 = rp;
  [[maybe_unused]] int rec_row = REC ? next_row(a.rec, rp, re) : 0;
  [[maybe_unused]] int rec_row2 = rec_row;
  [[maybe_unused]] float* uacc = a.uacc + int64_t(g.shot) * a.nrec * N_UACC;
  [[maybe_unused]] float* obs = a.obs + int64_t(g.shot) * 5 * a.nrec * a.nt;
  const int m0 = i0 - 4;   // first velocity row of stage 1
  // forward recompute of the gradient: v^{it+1} and v^{it+2} into their time slots
  auto aux_rsrc = [&](int step) {
    return make_rsrc(
        AUX ? a.aux + (int64_t(step % a.aux_period) * a.nshots + g.shot) * NAUX_SH * int64_t(a.npt) : a.model,
        AUX ? uint32_t(NAUX_SH) * npt4 : 0u);
  };
  const auto raux1 = aux_rsrc(it + 1), raux2 = aux_rsrc(it + 2);

  // stage 1 rings: stresses of step it (rows m-2..m+1), new vy (rows m-3..m),
  // D+z of the new vy (row m at s2, used at s0); stage 2 rings: stresses of
  // step it+1 (rows m-5..m-2, row m-2 at s0), vy of step it+2 (rows m-6..m-3,
  // row m-3 at s0), its D+z (row m-4 at [par], row m-5 at [par^1]); model
  // values and taper factor of three rows back (written at s0, read at s1)
  float sxy[4], szy[4], vy[4] = {0, 0, 0, 0}, dvydz[4] = {0, 0, 0, 0};
  float t1xy[4] = {0, 0, 0, 0}, t1zy[4] = {0, 0, 0, 0}, u2[4] = {0, 0, 0, 0}, dz2[2] = {0, 0};
  float drho[4] = {1, 1, 1, 1}, dirho[4] = {1, 1, 1, 1}, dmu[4] = {0, 0, 0, 0}, dbx[4] = {0, 0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    sxy[q] = settle(bload(rin, jc4, o_sxy + g.row(m0 - 2 + q)));
    szy[q] = settle(bload(rin, jc4, o_szy + g.row(m0 - 2 + q)));
  }
  sxy[3] = szy[3] = 0.0f;
  constexpr int D = 2;
  float pf_sxy[D], pf_szy[D], pf_vy[D], pf_rho[D], pf_irho[D], pf_mu[D];
  float bx_next = taper_x(a.bx, m0, nx);
  uint32_t rofs[4];
  for (int k = D - 2; k <= D; ++k) rofs[k & 3] = g.row(m0 + k);
  for (int q = 0; q < D; ++q) {
    pf_sxy[q] = bload(rin, jc4, o_sxy + g.row(m0 + 1 + q));
    pf_szy[q] = bload(rin, jc4, o_szy + g.row(m0 + 1 + q));
    pf_vy[q] = bload(rin, jc4, o_vy + g.row(m0 + q));
    pf_rho[q] = bload(rmod, jc4, m_rho + g.row(m0 + q));
    pf_irho[q] = bload(rmod, jc4, m_irho + g.row(m0 + q));
    pf_mu[q] = bload(rmod, jc4, m_mu + g.row(m0 - 2 + q));
  }
  lds_s[0][lane] = szy[2];
  __syncthreads();

  const bool zm_ok = g.zm_ok, zp_ok = g.zp_ok;
  auto body = [&](auto phase, int m) __attribute__((always_inline)) {
    constexpr int R = decltype(phase)::value;
    constexpr int s0 = R & 3, s1 = (R + 1) & 3, s2 = (R + 2) & 3, s3 = (R + 3) & 3, par = R & 1;
    constexpr int f = R % D;
    sxy[s3] = settle(pf_sxy[f]);       // row m+1
    szy[s3] = settle(pf_szy[f]);
    const float vy_o = pf_vy[f], rho = pf_rho[f], irho = pf_irho[f], mu = pf_mu[f], bxm = bx_next;
    bx_next = taper_x(a.bx, m + 1, nx);
    {
      constexpr int k1 = (R + D + 1) & 3, k0 = (R + D) & 3, km2 = (R + D + 2) & 3;
      rofs[k1] = g.row(m + D + 1);
      const uint32_t r3 = rofs[k1], r2 = rofs[k0], r0 = rofs[km2];
      pf_sxy[f] = bload(rin, jc4, o_sxy + r3);
      pf_szy[f] = bload(rin, jc4, o_szy + r3);
      pf_vy[f] = bload(rin, jc4, o_vy + r2);
      pf_rho[f] = bload(rmod, jc4, m_rho + r2);
      pf_irho[f] = bload(rmod, jc4, m_irho + r2);
      pf_mu[f] = bload(rmod, jc4, m_mu + r0);
    }

    // ---- stage 1: vy of step it+1 at row m (not stored)
    {
      const bool xm_ok = m >= 2 && m < nx - 2;
      float qa[2];
      {
        const float n0 = nminus(sxy[s0], sxy[s1], sxy[s2], sxy[s3]);
        const float n1 = nminus(lds_s[par][tm2], lds_s[par][tm1], szy[s2], lds_s[par][tp1]);
        const float xa[2] = {keep(xm_ok, n0), keep(zm_ok, n1)};
        div_group(qa, xa, {Dx, Dz}, fast_c);
      }
      float dsy = qa[0] + qa[1];
      if (m == inj_row) {
        for (; ip < ie; ++ip) {
          const int4 e = list_entry(a.inj, ip);
          if (e.x != m) break;
          if (e.y == j) dsy = dsy + uniform_value(a.val_y, int64_t(e.z) * a.nt_val + it);
        }
        inj_row = next_row(a.inj, ip, ie);
      }
      const float b = (bxm * bzb) * bzt;
      float qb[1];
      {
        const float xb[1] = {dt * dsy};
        div_group(qb, xb, {Div{rho, irho, true}}, fast_rho);
      }
      vy[s3] = b * (vy_o + qb[0]);
      if constexpr (AUX) {
        if (g.jst && m >= i0 && m < i1) bstore_out(vy[s3], raux1, jc4, A_VY * npt4 + uint32_t(m) * g.rowb, a.stream_out);
      }
    }
    lds_v[par][lane] = vy[s3];
    lds_s[par ^ 1][lane] = szy[s3];
    __syncthreads();

    // ---- stage 1: D+ of the new vy (z at row m, x at row m-2), stress of
    // step it+1 at row m-2 into the stage-2 ring and LDS
    {
      const int i = m - 2;
      const bool xp_ok = i >= 1 && i < nx - 2;
      float qc[2];
      {
        const float n0 = nplus(lds_v[par][tm1], vy[s3], lds_v[par][tp1], lds_v[par][tp2]);
        const float n1 = nplus(vy[s0], vy[s1], vy[s2], vy[s3]);
        const float xc[2] = {keep(zp_ok, n0), keep(xp_ok, n1)};
        div_group(qc, xc, {Dz, Dx}, fast_c);
      }
      dvydz[s2] = qc[0];
      if constexpr (REC) {
        if (m == rec_row) {
          for (; rp < re; ++rp) {
            const int4 e = list_entry(a.rec, rp);
            if (e.x != m) break;
            if (e.y == j) {
              float* u = uacc + e.z * N_UACC;
              const float uy = u[U_UY] + vy[s3] * dt;
              u[U_UY] = uy;
              obs[(1 * a.nrec + e.z) * a.nt + it] = uy;
            }
          }
          rec_row = next_row(a.rec, rp, re);
        }
      }
      const float mu_dt = mu * dt;
      const float dvydx = qc[1];
      t1xy[s0] = sxy[s0] + mu_dt * dvydx;
      t1zy[s0] = szy[s0] + mu_dt * dvydz[s0];
      lds_s2[par][lane] = t1zy[s0];
    }
    drho[s0] = rho;
    dirho[s0] = irho;
    dmu[s0] = mu;
    dbx[s0] = bxm;

    // ---- stage 2: vy of step it+2 at row m-3
    {
      const int m2 = m - 3;
      const bool xm_ok = m2 >= 2 && m2 < nx - 2;
      float qa[2];
      {
        const float n0 = nminus(t1xy[s1], t1xy[s2], t1xy[s3], t1xy[s0]);
        const float n1 = nminus(lds_s2[par ^ 1][tm2], lds_s2[par ^ 1][tm1], t1zy[s3], lds_s2[par ^ 1][tp1]);
        const float xa[2] = {keep(xm_ok, n0), keep(zm_ok, n1)};
        div_group(qa, xa, {Dx, Dz}, fast_c);
      }
      float dsy = qa[0] + qa[1];
      if (m2 == inj_row2) {
        for (; ip2 < ie; ++ip2) {
          const int4 e = list_entry(a.inj, ip2);
          if (e.x != m2) break;
          if (e.y == j) dsy = dsy + uniform_value(a.val_y, int64_t(e.z) * a.nt_val + it + 1);
        }
        inj_row2 = next_row(a.inj, ip2, ie);
      }
      const float b = (dbx[s1] * bzb) * bzt;
      float qb[1];
      {
        const float xb[1] = {dt * dsy};
        div_group(qb, xb, {Div{drho[s1], dirho[s1], true}}, fast_rho);
      }
      const float vyn = b * (vy[s0] + qb[0]);
      if (g.jst && m2 >= i0 && m2 < i1) {
        bstore_out(vyn, rout, jc4, o_vy + uint32_t(m2) * g.rowb, a.stream_out);
        if constexpr (AUX) bstore_out(vyn, raux2, jc4, A_VY * npt4 + uint32_t(m2) * g.rowb, a.stream_out);
      }
      u2[s0] = vyn;
      if constexpr (REC) {
        if (m2 == rec_row2) {
          for (; rp2 < re; ++rp2) {
            const int4 e = list_entry(a.rec, rp2);
            if (e.x != m2) break;
            if (e.y == j) {
              float* u = uacc + e.z * N_UACC;
              const float uy = u[U_UY] + vyn * dt;
              u[U_UY] = uy;
              obs[(1 * a.nrec + e.z) * a.nt + it + 1] = uy;
            }
          }
          rec_row2 = next_row(a.rec, rp2, re);
        }
      }
    }

    // ---- stage 2: D+z of vy row m-4 (LDS of the previous iteration), D+x at
    // row m-5, stress of step it+2 at row m-5
    {
      const int i2 = m - 5;
      const bool xp_ok = i2 >= 1 && i2 < nx - 2;
      float qc[2];
      {
        const float n0 = nplus(lds_v2[par ^ 1][tm1], u2[s3], lds_v2[par ^ 1][tp1], lds_v2[par ^ 1][tp2]);
        const float n1 = nplus(u2[s1], u2[s2], u2[s3], u2[s0]);
        const float xc[2] = {keep(zp_ok, n0), keep(xp_ok, n1)};
        div_group(qc, xc, {Dz, Dx}, fast_c);
      }
      dz2[par] = qc[0];
      const float mu_dt = dmu[s1] * dt;
      const float dvydx = qc[1];
      const float nsxy = t1xy[s1] + mu_dt * dvydx;
      const float nszy = t1zy[s1] + mu_dt * dz2[par ^ 1];
      if (g.jst && i2 >= i0 && i2 < i1) {
        bstore_out(nsxy, rout, jc4, o_sxy + uint32_t(i2) * g.rowb, a.stream_out);
        bstore_out(nszy, rout, jc4, o_szy + uint32_t(i2) * g.rowb, a.stream_out);
      }
    }
    lds_v2[par][lane] = u2[s0];
  };

  const int mlast = i1 + 4;
  int m = m0;
  body(Phase<0>{}, m);
  if (++m <= mlast) {
    body(Phase<1>{}, m);
    if (++m <= mlast) {
      body(Phase<2>{}, m);
      if (++m <= mlast) {
        body(Phase<3>{}, m);
        if (++m <= mlast) {
          for (;;) {
            body(Phase<4>{}, m);
            if (++m > mlast) break;
            body(Phase<5>{}, m);
            if (++m > mlast) break;
            body(Phase<6>{}, m);
            if (++m > mlast) break;
            body(Phase<7>{}, m);
            if (++m > mlast) break;
            body(Phase<0>{}, m);
            if (++m > mlast) break;
            body(Phase<1>{}, m);
            if (++m > mlast) break;
            body(Phase<2>{}, m);
            if (++m > mlast) break;
            body(Phase<3>{}, m);
            if (++m > mlast) break;
          }
        }
      }
    }
  }
  if (a.advance) advance_clock(a.clock, it, 2, a.nblocks);
}

// ------------------------ small grids: K steps per launch on LDS tiles
// Grids too small to give every CU several marching workgroups (up to a few
// hundred thousand shot-cells) are latency-bound: a marching workgroup waits
// on its row loads about once per row.  Here a workgroup of 1024 threads
// loads a 32 x 64 region of every state field into LDS at once - its tile
// of (32 - 2H) x (64 - 2H) cells plus an H = 3K-cell halo on each side, the
// reach of K steps - and advances the region K steps in LDS (velocity
// half-step | barrier | receivers, stress half-step | barrier, K times),
// then writes the tile back.  Halo cells are recomputed, never exchanged: a
// step shrinks the region's valid part by 3 cells on each side (the velocity
// reads stresses 2 before / 1 after, the stress reads new velocities 1
// before / 2 after), so after K steps the tile is exact; cells outside the
// grid stay 0 and are never computed (the FD edge rules mask them).  Every
// cell's arithmetic is that of step_sh / step_psv operation for operation,
// so results equal one-step launches bit for bit.  Thread t owns the cells
// (t / 64, t % 64) and (t / 64 + 16, t % 64) of the region: a wavefront is
// one region row.  Reads st[0], writes st[1]: with K = 2 the host runs
// launches in pairs through a spare buffer (as the marching two-step
// launches), with K = 3 st[1] is the buffer of the other parity.
constexpr int TL_R = 32, TL_C = 64;
enum TilePhys : int { TP_SH = 0, TP_PSV, TP_SPIN };

template <int PH, int K>
__device__ __forceinline__ void tile2(const StepArgs& a, int logical) {
  constexpr int TL_H = 3 * K, TL_TR = TL_R - 2 * TL_H, TL_TC = TL_C - 2 * TL_H;
  constexpr bool SH = PH == TP_SH, SPIN = PH == TP_SPIN;
  constexpr int NF = SH ? 3 : SPIN ? 9 : 5;
  // LDS fields: SH vy, sxy, szy; P-SV vx, vz, sxx, szz, sxz; spin + vy_c,
  // syx_c, syy_c, syz_c (the state slot order)
  enum { L_VY = 0, L_SXY = 1, L_SZY = 2 };
  enum { L_VX = 0, L_VZ, L_SXX, L_SZZ, L_SXZ, L_VYC, L_SYX, L_SYY, L_SYZ };
  // 2 cells of padding around the region: every neighbour read is a constant
  // offset from the cell (padding cells only feed cells outside the valid part)
  __shared__ float L[NF][TL_R + 4][TL_C + 4];
  const int tps = a.nstrips * a.nsegs;
  const int shot = logical / tps, tile = logical - shot * tps;
  const int strip = tile % a.nstrips, seg = tile / a.nstrips;
  const int nx = a.nx, nz = a.nz;
  const int r0 = seg * TL_TR - TL_H, c0 = strip * TL_TC - TL_H;   // grid row / column of region cell (0, 0)
  const int tid = threadIdx.x, cc = tid & (TL_C - 1), rr0 = tid >> 6;
  const int j = c0 + cc;
  const bool jin = j >= 0 && j < nz;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const float* __restrict__ s_in = a.st[0] + int64_t(shot) * a.shot_floats;
  float* __restrict__ s_out = a.st[1] + int64_t(shot) * a.shot_floats;
  const int npt = a.npt;
  const float dt = a.dt;
  const CDiv Dx{a.h8x, a.r8x, a.z8x, a.l8x, a.c_lo, (a.fast_div & FD_CONST) != 0},
             Dz{a.h8z, a.r8z, a.z8z, a.l8z, a.c_lo, (a.fast_div & FD_CONST) != 0};
  const int fast_c = div_mode(a.fast_div, FD_CONST);
  const int fast_rho = div_mode(a.fast_div, FD_RHO), fast_rj = div_mode(a.fast_div, FD_RHO | FD_J);
  const int jc = min(max(j, 0), nz - 1);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const bool zm_ok = j >= 2 && j < nz - 2, zp_ok = j >= 1 && j < nz - 2;
  int slot[NF];
  if constexpr (SH) {
    slot[L_VY] = a.slot_off[S_VY]; slot[L_SXY] = a.slot_off[S_SXY]; slot[L_SZY] = a.slot_off[S_SZY];
  } else {
    slot[L_VX] = a.slot_off[S_VX]; slot[L_VZ] = a.slot_off[S_VZ]; slot[L_SXX] = a.slot_off[S_SXX];
    slot[L_SZZ] = a.slot_off[S_SZZ]; slot[L_SXZ] = a.slot_off[S_SXZ];
    if constexpr (SPIN) {
      slot[L_VYC] = a.slot_off[S_VYC]; slot[L_SYX] = a.slot_off[S_SYXC];
      slot[L_SYY] = a.slot_off[S_SYYC]; slot[L_SYZ] = a.slot_off[S_SYZC];
    }
  }
  // the padding (and cells outside the grid) hold zeros: they feed the
  // outermost cells only, with finite values, so the divisions keep their
  // fast path there
  {
    float* flat = &L[0][0][0];
    for (int q = tid; q < NF * (TL_R + 4) * (TL_C + 4); q += kLdsTileThreads) flat[q] = 0.0f;
  }
  __syncthreads();
  // the two cells of this thread: region rows rr0 and rr0 + 16
  constexpr int NC = 2;
  int rr[NC], ii[NC];
  bool in[NC];
  float rho[NC], irho[NC], lam[NC], mu[NC], nu[NC], jj[NC], ij[NC], muc[NC], nuc[NC], bb[NC];
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    rr[q] = rr0 + 16 * q;
    ii[q] = r0 + rr[q];
    in[q] = jin && ii[q] >= 0 && ii[q] < nx;
    const int k = in[q] ? ii[q] * nz + j : 0;
#pragma unroll
    for (int f = 0; f < NF; ++f) L[f][rr[q] + 2][cc + 2] = in[q] ? s_in[slot[f] + k] : 0.0f;
    rho[q] = a.model[MF_RHO * npt + k];
    irho[q] = a.model[MF_IRHO * npt + k];
    mu[q] = a.model[MF_MU * npt + k];
    lam[q] = SH ? 0.0f : a.model[MF_LAM * npt + k];
    nu[q] = SPIN ? a.model[MF_NU * npt + k] : 0.0f;
    jj[q] = SPIN ? a.model[MF_J * npt + k] : 1.0f;
    ij[q] = SPIN ? a.model[MF_IJ * npt + k] : 1.0f;
    muc[q] = SPIN ? a.model[MF_MU_C * npt + k] : 0.0f;
    nuc[q] = SPIN ? a.model[MF_NU_C * npt + k] : 0.0f;
    bb[q] = (taper_x(a.bx, ii[q], nx) * bzb) * bzt;
  }
  // lists of this tile: injections anywhere in the region, receivers in the tile
  const int ip0 = __builtin_amdgcn_readfirstlane(a.inj_off[shot * tps + tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[shot * tps + tile + 1]);
  int rp0 = 0, re = 0;
  if (a.record) {
    rp0 = __builtin_amdgcn_readfirstlane(a.rec_off[tile]);
    re = __builtin_amdgcn_readfirstlane(a.rec_off[tile + 1]);
  }
  float* uacc = a.uacc + int64_t(shot) * a.nrec * N_UACC;
  float* obs = a.obs + int64_t(shot) * 5 * a.nrec * a.nt;
  const int nobs = a.nrec * a.nt;
  // neighbour reads clamp to the region: values past its edge only reach
  // cells outside the valid part, which are discarded
  auto at = [&](int f, int r, int c) -> float { return L[f][r + 2][c + 2]; };
  const int cm2 = cc - 2, cm1 = cc - 1, cp1 = cc + 1, cp2 = cc + 2;
  __syncthreads();

  for (int s = 0; s < K; ++s) {
    const int its = it + s;
    // ---- velocity half-step at every cell of the region (in place)
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      if (!in[q]) continue;
      const int r = rr[q], i = ii[q];
      const bool xm_ok = i >= 2 && i < nx - 2;
      if constexpr (SH) {
        float qa[2];
        const float xa[2] = {xm_ok ? nminus(at(L_SXY, r - 2, cc), at(L_SXY, r - 1, cc), at(L_SXY, r, cc),
                                            at(L_SXY, r + 1, cc))
                                   : 0.0f,
                             zm_ok ? nminus(at(L_SZY, r, cm2), at(L_SZY, r, cm1), at(L_SZY, r, cc), at(L_SZY, r, cp1))
                                   : 0.0f};
        div_group(qa, xa, {Dx, Dz}, fast_c);
        float dsy = qa[0] + qa[1];
        for (int p = ip0; p < ie; ++p) {
          const int4 e = list_entry(a.inj, p);
          if (e.x == i && e.y == j) dsy = dsy + uniform_value(a.val_y, int64_t(e.z) * a.nt_val + its);
        }
        float qb[1];
        const float xb[1] = {dt * dsy};
        div_group(qb, xb, {Div{rho[q], irho[q], true}}, fast_rho);
        L[L_VY][r + 2][cc + 2] = bb[q] * (L[L_VY][r + 2][cc + 2] + qb[0]);
      } else {
        constexpr int NA = SPIN ? 8 : 4;
        float qa[NA], xa[NA];
        xa[0] = keep(xm_ok, nminus(at(L_SXX, r - 2, cc), at(L_SXX, r - 1, cc), at(L_SXX, r, cc), at(L_SXX, r + 1, cc)));
        xa[1] = keep(zm_ok, nminus(at(L_SXZ, r, cm2), at(L_SXZ, r, cm1), at(L_SXZ, r, cc), at(L_SXZ, r, cp1)));
        xa[2] = keep(xm_ok, nminus(at(L_SXZ, r - 2, cc), at(L_SXZ, r - 1, cc), at(L_SXZ, r, cc), at(L_SXZ, r + 1, cc)));
        xa[3] = keep(zm_ok, nminus(at(L_SZZ, r, cm2), at(L_SZZ, r, cm1), at(L_SZZ, r, cc), at(L_SZZ, r, cp1)));
        if constexpr (SPIN) {
          xa[4] = keep(xm_ok, nminus(at(L_SYX, r - 2, cc), at(L_SYX, r - 1, cc), at(L_SYX, r, cc), at(L_SYX, r + 1, cc)));
          xa[5] = keep(zm_ok, nminus(at(L_SYZ, r, cm2), at(L_SYZ, r, cm1), at(L_SYZ, r, cc), at(L_SYZ, r, cp1)));
          xa[6] = keep(zm_ok, nminus(at(L_SYY, r, cm2), at(L_SYY, r, cm1), at(L_SYY, r, cc), at(L_SYY, r, cp1)));
          xa[7] = keep(xm_ok, nminus(at(L_SYY, r - 2, cc), at(L_SYY, r - 1, cc), at(L_SYY, r, cc), at(L_SYY, r + 1, cc)));
          div_group(qa, xa, {Dx, Dz, Dx, Dz, Dx, Dz, Dz, Dx}, fast_c);
        } else {
          div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
        }
        float dsx = qa[0] + qa[1];
        float dsz = qa[2] + qa[3];
        float dsyc = 0.0f;
        if constexpr (SPIN) {
          dsyc = qa[4] + qa[5];
          dsx = dsx + qa[6];
          dsz = dsz + (-qa[7]);
          dsyc = dsyc + 2.0f * L[L_SYY][r + 2][cc + 2];
        }
        for (int p = ip0; p < ie; ++p) {
          const int4 e = list_entry(a.inj, p);
          if (e.x == i && e.y == j) {
            dsx = dsx + uniform_value(a.val_x, int64_t(e.z) * a.nt_val + its);
            dsz = dsz + uniform_value(a.val_z, int64_t(e.z) * a.nt_val + its);
          }
        }
        constexpr int NB = SPIN ? 3 : 2;
        float qb[NB];
        if constexpr (SPIN) {
          const float xb[3] = {dt * dsx, dt * dsz, dt * dsyc};
          div_group(qb, xb, {Div{rho[q], irho[q], true}, Div{rho[q], irho[q], true}, Div{jj[q], ij[q], true}}, fast_rj);
          L[L_VYC][r + 2][cc + 2] = bb[q] * (L[L_VYC][r + 2][cc + 2] + qb[2]);
        } else {
          const float xb[2] = {dt * dsx, dt * dsz};
          div_group(qb, xb, {Div{rho[q], irho[q], true}, Div{rho[q], irho[q], true}}, fast_rho);
        }
        L[L_VX][r + 2][cc + 2] = bb[q] * (L[L_VX][r + 2][cc + 2] + qb[0]);
        L[L_VZ][r + 2][cc + 2] = bb[q] * (L[L_VZ][r + 2][cc + 2] + qb[1]);
      }
    }
    __syncthreads();

    // ---- receivers of step its: u = sum v dt at the cells the traces need
    // (each receiver belongs to one tile; its owner thread reads the new
    // velocities of the neighbouring cells from LDS)
    for (int p = rp0; p < re; ++p) {
      const int4 e = list_entry(a.rec, p);
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        if (!(in[q] && e.x == ii[q] && e.y == j)) continue;
        float* u = uacc + e.z * N_UACC;
        const int kt = e.z * a.nt + its;
        const int r = rr[q];
        if constexpr (SH) {
          const float uy = u[U_UY] + L[L_VY][r + 2][cc + 2] * dt;
          u[U_UY] = uy;
          obs[1 * nobs + kt] = uy;
        } else if constexpr (!SPIN) {
          const float ux = u[U_UX] + L[L_VX][r + 2][cc + 2] * dt, uz = u[U_UZ] + L[L_VZ][r + 2][cc + 2] * dt;
          u[U_UX] = ux;
          u[U_UZ] = uz;
          obs[0 * nobs + kt] = ux;
          obs[2 * nobs + kt] = uz;
        } else {
          u[U_UZ_M1] = u[U_UZ_M1] + at(L_VZ, r - 1, cc) * dt;
          u[U_UX] = u[U_UX] + L[L_VX][r + 2][cc + 2] * dt;
          u[U_UZ] = u[U_UZ] + L[L_VZ][r + 2][cc + 2] * dt;
          u[U_UYC] = u[U_UYC] + L[L_VYC][r + 2][cc + 2] * dt;
          u[U_UX_M1] = u[U_UX_M1] + at(L_VX, r, cm1) * dt;
          u[U_UX_P1] = u[U_UX_P1] + at(L_VX, r, cp1) * dt;
          u[U_UX_P2] = u[U_UX_P2] + at(L_VX, r, cp2) * dt;
          u[U_UZ_P1] = u[U_UZ_P1] + at(L_VZ, r + 1, cc) * dt;
          const float uz_p2 = u[U_UZ_P2] + at(L_VZ, r + 2, cc) * dt;
          u[U_UZ_P2] = uz_p2;
          const float ux = u[U_UX], uz = u[U_UZ], uyc = u[U_UYC];
          const bool xp = e.x >= 1 && e.x < nx - 2;
          const bool zp = e.y >= 1 && e.y < nz - 2;
          const float duzdx = xp ? dplus(u[U_UZ_M1], uz, u[U_UZ_P1], uz_p2, Dx) : 0.0f;
          const float duxdz = zp ? dplus(u[U_UX_M1], ux, u[U_UX_P1], u[U_UX_P2], Dz) : 0.0f;
          const float curl = -(duzdx - duxdz);
          const float iso = uyc - 0.5f * curl;
          obs[0 * nobs + kt] = ux;
          obs[2 * nobs + kt] = uz;
          obs[1 * nobs + kt] = uyc;
          obs[3 * nobs + kt] = iso;
          obs[4 * nobs + kt] = curl;
        }
      }
    }

    // ---- stress half-step at every cell of the region (in place)
#pragma unroll
    for (int q = 0; q < NC; ++q) {
      if (!in[q]) continue;
      const int r = rr[q], i = ii[q];
      const bool xp_ok = i >= 1 && i < nx - 2;
      if constexpr (SH) {
        float qc[2];
        const float vy0 = L[L_VY][r + 2][cc + 2];
        const float xc[2] = {zp_ok ? nplus(at(L_VY, r, cm1), vy0, at(L_VY, r, cp1), at(L_VY, r, cp2)) : 0.0f,
                             xp_ok ? nplus(at(L_VY, r - 1, cc), vy0, at(L_VY, r + 1, cc), at(L_VY, r + 2, cc)) : 0.0f};
        div_group(qc, xc, {Dz, Dx}, fast_c);
        const float mu_dt = mu[q] * dt;
        L[L_SXY][r + 2][cc + 2] = L[L_SXY][r + 2][cc + 2] + mu_dt * qc[1];
        L[L_SZY][r + 2][cc + 2] = L[L_SZY][r + 2][cc + 2] + mu_dt * qc[0];
      } else {
        constexpr int NCD = SPIN ? 6 : 4;
        float qc[NCD], xc[NCD];
        const float vx0 = L[L_VX][r + 2][cc + 2], vz0 = L[L_VZ][r + 2][cc + 2];
        xc[0] = keep(zp_ok, nplus(at(L_VX, r, cm1), vx0, at(L_VX, r, cp1), at(L_VX, r, cp2)));
        xc[1] = keep(zp_ok, nplus(at(L_VZ, r, cm1), vz0, at(L_VZ, r, cp1), at(L_VZ, r, cp2)));
        xc[2] = keep(xp_ok, nplus(at(L_VX, r - 1, cc), vx0, at(L_VX, r + 1, cc), at(L_VX, r + 2, cc)));
        xc[3] = keep(xp_ok, nplus(at(L_VZ, r - 1, cc), vz0, at(L_VZ, r + 1, cc), at(L_VZ, r + 2, cc)));
        float vyc0 = 0.0f;
        if constexpr (SPIN) {
          vyc0 = L[L_VYC][r + 2][cc + 2];
          xc[4] = keep(zp_ok, nplus(at(L_VYC, r, cm1), vyc0, at(L_VYC, r, cp1), at(L_VYC, r, cp2)));
          xc[5] = keep(xp_ok, nplus(at(L_VYC, r - 1, cc), vyc0, at(L_VYC, r + 1, cc), at(L_VYC, r + 2, cc)));
          div_group(qc, xc, {Dz, Dz, Dx, Dx, Dz, Dx}, fast_c);
        } else {
          div_group(qc, xc, {Dz, Dz, Dx, Dx}, fast_c);
        }
        const float dvxdz = qc[0], dvzdz = qc[1], dvxdx = qc[2], dvzdx = qc[3];
        const float l2m = lam[q] + 2.0f * mu[q];
        L[L_SXX][r + 2][cc + 2] = L[L_SXX][r + 2][cc + 2] + dt * (l2m * dvxdx + lam[q] * dvzdz);
        L[L_SZZ][r + 2][cc + 2] = L[L_SZZ][r + 2][cc + 2] + dt * (l2m * dvzdz + lam[q] * dvxdx);
        L[L_SXZ][r + 2][cc + 2] = L[L_SXZ][r + 2][cc + 2] + dt * mu[q] * (dvxdz + dvzdx);
        if constexpr (SPIN) {
          const float dvycdz = qc[4], dvycdx = qc[5];
          L[L_SYY][r + 2][cc + 2] = L[L_SYY][r + 2][cc + 2] + 2.0f * dt * nu[q] * (-vyc0 - 0.5f * (dvzdx - dvxdz));
          const float c2 = dt * (muc[q] + nuc[q]);
          L[L_SYX][r + 2][cc + 2] = L[L_SYX][r + 2][cc + 2] + c2 * dvycdx;
          L[L_SYZ][r + 2][cc + 2] = L[L_SYZ][r + 2][cc + 2] + c2 * dvycdz;
        }
      }
    }
    __syncthreads();
  }

  // ---- the tile back to memory
#pragma unroll
  for (int q = 0; q < NC; ++q) {
    const int r = rr[q];
    if (!in[q] || r < TL_H || r >= TL_R - TL_H || cc < TL_H || cc >= TL_C - TL_H) continue;
    const int k = ii[q] * nz + j;
#pragma unroll
    for (int f = 0; f < NF; ++f) s_out[slot[f] + k] = L[f][r + 2][cc + 2];
  }
  if (a.advance) advance_clock(a.clock, it, K, a.nblocks);
}

// ------------------------------------------------------ adjoint time steps
// Exact transpose of one forward step (SH or P-SV), run in reverse time.
// With the forward step
//   v' = B (v + dt (D- sigma + f) / rho),   sigma' = sigma + dt C(lambda, mu) D+ v'
// and the traces d^n = u^{n+1} = u^n + dt v' at the receivers, the adjoint
// step maps (vbar', sigbar') to (vbar, sigbar):
//   vbar''  = vbar' + dt A^n (receivers) + (D+)^T t,     t = dt C sigbar'
//   g       = B vbar'',   vbar = g,   q = dt g / rho
//   sigbar  = sigbar' + (D-)^T q
//   K_rho  += -dt g . ds / rho^2             (ds = D- sigma + f of the forward step)
//   K_mu, K_lambda += dt sigbar' . dC/dm . strain(v')
// with (D+)^T w = -D-(w~) and (D-)^T w = -D+(w^), where w~ / w^ are w zeroed
// outside the rows/columns on which the forward operator is active (D+: 1 <=
// i < n-2, D-: 2 <= i < n-2).  Both terms come from the forward velocities
// v' = v^{n+1} the recompute stored: the strains are D+ v^{n+1} (the
// operators of the forward's stress half-step), and since v^{n+1} = B (v^n +
// dt ds / rho), dt g . ds / rho^2 = (vbar'' . v^{n+1} - g . v^n) / rho.
// Summed over the steps, g of step n+1 is the vbar' of step n, so the sum
// telescopes (g after the last adjoint step and v^0 are zero) into
//   K_rho += -v^{n+1} . (vbar'' - vbar') / rho = -v^{n+1} . (dt A^n + (D+)^T t) / rho
// per step: one time level of stored velocities, and the difference is the
// update itself, formed before it is added (no cancellation).  Same
// marching/tiling as the forward step.
// The adjoint is not a reference path (its parity is against the FP64
// adjoint, to a tolerance): its divisions by 8h and by rho are
// multiplications by the correctly rounded reciprocals (within 1.5 units in
// the last place of the quotient), so it reads 1/rho and not rho.
__device__ __forceinline__ float ndm(float fm2, float fm1, float f0, float fp1, Div d) {
  return -((9.0f * (f0 - fm1) - (fp1 - fm2)) * d.r);
}
__device__ __forceinline__ float ndp(float fm1, float f0, float fp1, float fp2, Div d) {
  return -((9.0f * (fp1 - f0) - (fp2 - fm1)) * d.r);
}
// D+ of a forward field (the strain operator)
__device__ __forceinline__ float dpf(float fm1, float f0, float fp1, float fp2, Div d) {
  return (9.0f * (fp1 - f0) - (fp2 - fm1)) * d.r;
}

// sum of column t's values of the S shot groups of a workgroup, in shot order
template <int S, int NT>
__device__ __forceinline__ float shot_sum(const float* v, int t) {
  float acc = v[t];
  for (int s = 1; s < S; ++s) acc = acc + v[s * NT + t];
  return acc;
}

// The stored forward velocities of a shot at time level n (slot n % period).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t fwd_vel(const StepArgs& a, int n, int shot, int nf) {
  return make_rsrc(a.aux + (int64_t(n % a.aux_period) * a.nshots + shot) * nf * int64_t(a.npt),
                   uint32_t(nf) * 4u * uint32_t(a.npt));
}

template <int NT, int S>
__device__ __forceinline__ void adj_sh(const StepArgs& a, int logical) {
  __shared__ float lds_t[2][S * NT];
  __shared__ float lds_q[2][S * NT];
  // forward vy^{n+1} of rows m-3..m (slot as the register ring's), for the
  // z-derivative two rows later
  __shared__ float lds_f[4][S * NT];
  // several shots per workgroup: each row's K_rho, K_mu contributions of
  // every shot, summed over the shots one iteration later
  __shared__ float lds_k[S > 1 ? 2 : 1][S > 1 ? 2 : 1][S > 1 ? S * NT : 1];
  const TileGeom g = tile_geom<NT, S>(a, logical);
  const int t = g.t, lane = g.lane, j = g.j, nx = a.nx, nz = a.nz, i0 = g.i0, i1 = g.i1;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (it < a.it_lo || it >= a.it_hi) return;
  const int64_t so = int64_t(g.shot) * a.shot_floats;
  const uint32_t sbytes = 4u * uint32_t(a.shot_floats);
  const auto rin = make_rsrc(a.st[(it + 1) & 1] + so, sbytes);
  const auto rout = make_rsrc(a.st[it & 1] + so, sbytes);
  const uint32_t npt4 = 4u * uint32_t(a.npt);
  const auto rmod = make_rsrc(a.model, uint32_t(MF_N) * npt4);
  const auto rf1 = fwd_vel(a, it + 1, g.shot, NAUX_SH);
  // one accumulator set per workgroup, as in adj_psv (field f summed by
  // shot group f % S)
  const auto rk = make_rsrc(a.kgrad + int64_t(g.shot / S) * NK * int64_t(a.npt), NK * npt4);
  const int sub = g.base / NT;
  const bool own_r = S == 1 || sub == 0, own_m = S == 1 || sub == 1 % S;
  constexpr int KR_LAG = S > 1 ? 1 : 0, KS_LAG = S > 1 ? 3 : 2;   // K rows behind iteration m
  const uint32_t o_vy = 4u * a.slot_off[S_VY], o_sxy = 4u * a.slot_off[S_SXY], o_szy = 4u * a.slot_off[S_SZY];
  const uint32_t m_irho = MF_IRHO * npt4, m_mu = MF_MU * npt4;
  const uint32_t k_rho = K_RHO * npt4, k_mu = K_MU * npt4;
  const uint32_t jc4 = g.jc4;
  const int jc = int(jc4 >> 2);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const float dt = a.dt;
  const Div Dx{a.h8x, a.r8x, true}, Dz{a.h8z, a.r8z, true};
  const int tm2 = g.base + max(t - 2, 0), tm1 = g.base + max(t - 1, 0), tp1 = g.base + min(t + 1, NT - 1),
            tp2 = g.base + min(t + 2, NT - 1);
  const bool zt = j >= 1 && j < nz - 2, zq = j >= 2 && j < nz - 2;
  auto xt = [&](int r) { return r >= 1 && r < nx - 2; };
  auto xq = [&](int r) { return r >= 2 && r < nx - 2; };
  const int tps = a.nstrips * a.nsegs;
  int ip = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile + 1]);
  int inj_row = next_row(a.inj, ip, ie);     // row of the next adjoint-source entry

  // rings (as in the forward): sigbar rows m-2..m+1 and their x-masked t;
  // q (x-masked) rows m-3..m; -D+z(q^) rows m-2..m; forward vy^{n+1} rows m-3..m
  float sbx[4], sbz[4], tx[4], q[4] = {0, 0, 0, 0}, dzq[4] = {0, 0, 0, 0}, f1[4] = {0, 0, 0, 0};
  float tz_cur = 0.0f;    // z-masked t_z of row m
  for (int k = 0; k < 3; ++k) {
    const int r = i0 - 3 + k;
    const uint32_t ro = g.row(r);
    sbx[k] = settle(bload(rin, jc4, o_sxy + ro));
    sbz[k] = settle(bload(rin, jc4, o_szy + ro));
    const float mu = bload(rmod, jc4, m_mu + ro);
    tx[k] = xt(r) ? dt * mu * sbx[k] : 0.0f;
    if (k == 2) tz_cur = zt ? dt * mu * sbz[k] : 0.0f;
  }
  sbx[3] = sbz[3] = tx[3] = 0.0f;
  // D-deep prefetch FIFO, slot (m - (i0-1)) % D: sigbar and mu rows m+1;
  // vbar, 1/rho, the forward vy^{n+1} and K_rho rows m; K_mu rows m-2
  // of iteration m.  D = 2, or 1 with several shots per workgroup.
  constexpr int D = S > 1 ? 1 : 2;
  static_assert(S == 1 || D == 1, "the shot sums' epilogue reads prefetch slot 0");
  float pf_sbx[D], pf_sbz[D], pf_mu[D], pf_vb[D], pf_irho[D], pf_f1[D], pf_kr[D], pf_km[D];
  float bx_next = taper_x(a.bx, i0 - 1, nx);
  for (int k = 0; k < D; ++k) {
    const uint32_t r1 = g.row(i0 + k), r0 = g.row(i0 - 1 + k);
    const uint32_t rkr = g.row(i0 - 1 + k - KR_LAG), rks = g.row(i0 - 1 + k - KS_LAG);
    pf_sbx[k] = bload(rin, jc4, o_sxy + r1);
    pf_sbz[k] = bload(rin, jc4, o_szy + r1);
    pf_mu[k] = bload(rmod, jc4, m_mu + r1);
    pf_vb[k] = bload(rin, jc4, o_vy + r0);
    pf_irho[k] = bload(rmod, jc4, m_irho + r0);
    pf_f1[k] = bload(rf1, jc4, A_VY * npt4 + r0);
    pf_kr[k] = own_r ? bload(rk, jc4, k_rho + rkr) : 0.0f;
    pf_km[k] = own_m ? bload(rk, jc4, k_mu + rks) : 0.0f;
  }

  // one barrier per row (as in the forward step): t of row i0-1 goes to
  // lds_t[0]; iteration m writes q and the forward vy of row m and t of row
  // m+1 before it
  lds_t[0][lane] = tz_cur;
  __syncthreads();

  auto body = [&](auto phase, int m) __attribute__((always_inline)) {
    constexpr int R = decltype(phase)::value;
    constexpr int s0 = R & 3, s1 = (R + 1) & 3, s2 = (R + 2) & 3, s3 = (R + 3) & 3, par = R & 1, f = R % D;
    // row m+1 enters the rings
    sbx[s3] = settle(pf_sbx[f]);
    sbz[s3] = settle(pf_sbz[f]);
    const float mu1 = pf_mu[f];
    tx[s3] = xt(m + 1) ? dt * mu1 * sbx[s3] : 0.0f;
    const float tz_next = zt ? dt * mu1 * sbz[s3] : 0.0f;
    const float vb_o = pf_vb[f], irho = pf_irho[f];
    f1[s3] = settle(pf_f1[f]);     // forward vy^{n+1}, row m
    const float kr = pf_kr[f], km = pf_km[f], bxm = bx_next;
    bx_next = taper_x(a.bx, m + 1, nx);
    {
      const uint32_t r3 = g.row(m + D + 1), r2 = g.row(m + D);
      const uint32_t rkr = g.row(m + D - KR_LAG), rks = g.row(m + D - KS_LAG);
      pf_sbx[f] = bload(rin, jc4, o_sxy + r3);
      pf_sbz[f] = bload(rin, jc4, o_szy + r3);
      pf_mu[f] = bload(rmod, jc4, m_mu + r3);
      pf_vb[f] = bload(rin, jc4, o_vy + r2);
      pf_irho[f] = bload(rmod, jc4, m_irho + r2);
      pf_f1[f] = bload(rf1, jc4, A_VY * npt4 + r2);
      if (own_r) pf_kr[f] = bload(rk, jc4, k_rho + rkr);
      if (own_m) pf_km[f] = bload(rk, jc4, k_mu + rks);
    }

    // adjoint "velocity" at row m: vbar'' = vbar' + inj + div
    float vb = vb_o, inj = 0.0f;
    if (m >= inj_row) {
      for (; ip < ie; ++ip) {
        const int4 e = list_entry(a.inj, ip);
        if (e.x > m) break;
        if (e.x == m && e.y == j) {
          const float v = uniform_value(a.val_y, int64_t(e.z) * a.nt_val + it);
          vb = vb + v;
          inj = inj + v;
        }
      }
      inj_row = next_row(a.inj, ip, ie);
    }
    const float div = ndm(tx[s0], tx[s1], tx[s2], tx[s3], Dx) +
                      ndm(lds_t[par][tm2], lds_t[par][tm1], tz_cur, lds_t[par][tp1], Dz);
    vb = vb + div;
    const float b = (bxm * bzb) * bzt;
    const float gv = b * vb;
    const float qv = (dt * gv) * irho;     // dt g / rho
    const float cr = (-(f1[s3] * (inj + div))) * irho;   // K_rho += -v^{n+1} (vbar'' - vbar') / rho
    if (g.jst && m >= i0 && m < i1) {
      const uint32_t ro = uint32_t(m) * g.rowb;
      bstore_out(gv, rout, jc4, o_vy + ro, a.stream_out);
      if constexpr (S == 1) bstore(kr + cr, rk, jc4, k_rho + ro);
    }
    q[s3] = xq(m) ? qv : 0.0f;
    const float qz = zq ? qv : 0.0f;
    lds_q[par][lane] = qz;
    lds_t[par ^ 1][lane] = tz_next;
    lds_f[s3][lane] = f1[s3];
    __syncthreads();
    dzq[s2] = ndp(lds_q[par][tm1], qz, lds_q[par][tp1], lds_q[par][tp2], Dz);
    tz_cur = tz_next;

    // adjoint "stress" at row i = m-2, and the strains of vy^{n+1} there
    const int i = m - 2;
    if (i >= i0 && g.jst) {
      const uint32_t ro = uint32_t(i) * g.rowb;
      const float nsx = sbx[s0] + ndp(q[s0], q[s1], q[s2], q[s3], Dx);
      const float nsz = sbz[s0] + dzq[s0];
      bstore_out(nsx, rout, jc4, o_sxy + ro, a.stream_out);
      bstore_out(nsz, rout, jc4, o_szy + ro, a.stream_out);
    }
    const float exy = xt(i) ? dpf(f1[s0], f1[s1], f1[s2], f1[s3], Dx) : 0.0f;
    const float ezy = zt ? dpf(lds_f[s1][tm1], f1[s1], lds_f[s1][tp1], lds_f[s1][tp2], Dz) : 0.0f;
    const float cm = dt * (sbx[s0] * exy + sbz[s0] * ezy);
    if constexpr (S == 1) {
      if (i >= i0 && g.jst) bstore(km + cm, rk, jc4, k_mu + uint32_t(i) * g.rowb);
    } else {
      // the previous iteration's contributions (K_rho row m-1, K_mu row m-3)
      if (g.jst) {
        if (own_r && m - 1 >= i0 && m - 1 < i1)
          bstore(kr + shot_sum<S, NT>(lds_k[par ^ 1][0], t), rk, jc4, k_rho + uint32_t(m - 1) * g.rowb);
        if (own_m && m - 3 >= i0)
          bstore(km + shot_sum<S, NT>(lds_k[par ^ 1][1], t), rk, jc4, k_mu + uint32_t(m - 3) * g.rowb);
      }
      lds_k[par][0][lane] = cr;
      lds_k[par][1][lane] = cm;
    }
  };

  const int mlast = i1 + 1;
  for (int m = i0 - 1;;) {
    body(Phase<0>{}, m);
    if (++m > mlast) break;
    body(Phase<1>{}, m);
    if (++m > mlast) break;
    body(Phase<2>{}, m);
    if (++m > mlast) break;
    body(Phase<3>{}, m);
    if (++m > mlast) break;
  }
  if constexpr (S > 1) {
    // the last iteration's K_mu contributions (row i1 - 1)
    __syncthreads();
    const int par = (mlast - (i0 - 1)) & 1, r = mlast - 2;
    if (g.jst && r >= i0 && own_m)
      bstore(pf_km[0] + shot_sum<S, NT>(lds_k[par][1], t), rk, jc4, k_mu + uint32_t(r) * g.rowb);
  }
  if (a.advance) advance_clock(a.clock, it, -1, a.nblocks);
}

template <int NT, int S>
__device__ __forceinline__ void adj_psv(const StepArgs& a, int logical) {
  __shared__ float lds_t[2][2][S * NT];   // z-masked t_xz, t_zz of row m
  __shared__ float lds_q[2][2][S * NT];   // z-masked qx, qz of row m
  __shared__ float lds_f[4][2][S * NT];   // forward vx, vz^{n+1} of rows m-3..m (register-ring slots)
  // several shots per workgroup: each row's K_rho, K_lambda, K_mu
  // contributions of every shot, summed over the shots one iteration later
  __shared__ float lds_k[S > 1 ? 2 : 1][S > 1 ? 3 : 1][S > 1 ? S * NT : 1];
  const TileGeom g = tile_geom<NT, S>(a, logical);
  const int t = g.t, lane = g.lane, j = g.j, nx = a.nx, nz = a.nz, i0 = g.i0, i1 = g.i1;
  const int it = __hip_atomic_load(&a.clock[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (it < a.it_lo || it >= a.it_hi) return;
  const int64_t so = int64_t(g.shot) * a.shot_floats;
  const uint32_t sbytes = 4u * uint32_t(a.shot_floats);
  const auto rin = make_rsrc(a.st[(it + 1) & 1] + so, sbytes);
  const auto rout = make_rsrc(a.st[it & 1] + so, sbytes);
  const uint32_t npt4 = 4u * uint32_t(a.npt);
  const auto rmod = make_rsrc(a.model, uint32_t(MF_N) * npt4);
  const auto rf1 = fwd_vel(a, it + 1, g.shot, NAUX_PSV);
  // kernel accumulators: one set per workgroup ([shot / S][NK][npt]); with
  // S > 1 the S shots' contributions are added in shot order by the shot
  // group that owns the field (field f: group f % S), one read-modify-write
  // per field and cell instead of S
  const auto rk = make_rsrc(a.kgrad + int64_t(g.shot / S) * NK * int64_t(a.npt), NK * npt4);
  const int sub = g.base / NT;
  const bool own_r = S == 1 || sub == 0, own_l = S == 1 || sub == 1 % S, own_m = S == 1 || sub == 2 % S;
  constexpr int KR_LAG = S > 1 ? 1 : 0, KS_LAG = S > 1 ? 3 : 2;   // K rows behind iteration m
  const uint32_t o_vx = 4u * a.slot_off[S_VX], o_vz = 4u * a.slot_off[S_VZ];
  const uint32_t o_sxx = 4u * a.slot_off[S_SXX], o_szz = 4u * a.slot_off[S_SZZ], o_sxz = 4u * a.slot_off[S_SXZ];
  const uint32_t m_irho = MF_IRHO * npt4, m_lam = MF_LAM * npt4, m_mu = MF_MU * npt4;
  const uint32_t k_rho = K_RHO * npt4, k_lam = K_LAM * npt4, k_mu = K_MU * npt4;
  const uint32_t jc4 = g.jc4;
  const int jc = int(jc4 >> 2);
  const float bzb = a.bzb[jc], bzt = a.bzt[jc];
  const float dt = a.dt;
  const Div Dx{a.h8x, a.r8x, true}, Dz{a.h8z, a.r8z, true};
  const int tm2 = g.base + max(t - 2, 0), tm1 = g.base + max(t - 1, 0), tp1 = g.base + min(t + 1, NT - 1),
            tp2 = g.base + min(t + 2, NT - 1);
  const bool zt = j >= 1 && j < nz - 2, zq = j >= 2 && j < nz - 2;
  auto xt = [&](int r) { return r >= 1 && r < nx - 2; };
  auto xq = [&](int r) { return r >= 2 && r < nx - 2; };
  const int tps = a.nstrips * a.nsegs;
  int ip = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile]);
  const int ie = __builtin_amdgcn_readfirstlane(a.inj_off[g.shot * tps + g.tile + 1]);
  int inj_row = next_row(a.inj, ip, ie);     // row of the next adjoint-source entry

  float sbxx[4], sbzz[4], sbxz[4], txx[4], txz[4];
  float qx[4] = {0, 0, 0, 0}, qz[4] = {0, 0, 0, 0}, dzqx[4] = {0, 0, 0, 0}, dzqz[4] = {0, 0, 0, 0};
  float fx[4] = {0, 0, 0, 0}, fz[4] = {0, 0, 0, 0};   // forward vx, vz^{n+1}, rows m-3..m
  float tzxz_cur = 0.0f, tzzz_cur = 0.0f;   // z-masked t_xz, t_zz of row m
  auto tvals = [&](float lam, float mu, float sxx, float szz, float sxz, float& oxx, float& ozz, float& oxz) {
    const float l2m = lam + 2.0f * mu;
    oxx = dt * (l2m * sxx + lam * szz);
    ozz = dt * (lam * sxx + l2m * szz);
    oxz = dt * mu * sxz;
  };
  for (int k = 0; k < 3; ++k) {
    const int r = i0 - 3 + k;
    const uint32_t ro = g.row(r);
    sbxx[k] = settle(bload(rin, jc4, o_sxx + ro));
    sbzz[k] = settle(bload(rin, jc4, o_szz + ro));
    sbxz[k] = settle(bload(rin, jc4, o_sxz + ro));
    float a_xx, a_zz, a_xz;
    tvals(bload(rmod, jc4, m_lam + ro), bload(rmod, jc4, m_mu + ro), sbxx[k], sbzz[k], sbxz[k], a_xx, a_zz, a_xz);
    txx[k] = xt(r) ? a_xx : 0.0f;
    txz[k] = xt(r) ? a_xz : 0.0f;
    if (k == 2) { tzxz_cur = zt ? a_xz : 0.0f; tzzz_cur = zt ? a_zz : 0.0f; }
  }
  sbxx[3] = sbzz[3] = sbxz[3] = txx[3] = txz[3] = 0.0f;
  // D-deep prefetch FIFO, slot (m - (i0-1)) % D: sigbar, lambda, mu rows
  // m+1; vbar, 1/rho, the forward velocities v^{n+1} and K_rho rows m;
  // K_lambda and K_mu rows m-2 of iteration m.  D = 1: with D = 2 the 13
  // fields in flight per row exceed the 128 VGPRs of four waves per SIMD.
  constexpr int D = 1;
  static_assert(S == 1 || D == 1, "the shot sums' epilogue reads prefetch slot 0");
  float pf_sxx[D], pf_szz[D], pf_sxz[D], pf_lam[D], pf_mu[D], pf_vbx[D], pf_vbz[D], pf_irho[D];
  float pf_fx1[D], pf_fz1[D], pf_kr[D], pf_kl[D], pf_km[D];
  float bx_next = taper_x(a.bx, i0 - 1, nx);
  for (int k = 0; k < D; ++k) {
    const uint32_t r1 = g.row(i0 + k), r0 = g.row(i0 - 1 + k);
    const uint32_t rkr = g.row(i0 - 1 + k - KR_LAG), rks = g.row(i0 - 1 + k - KS_LAG);
    pf_sxx[k] = bload(rin, jc4, o_sxx + r1);
    pf_szz[k] = bload(rin, jc4, o_szz + r1);
    pf_sxz[k] = bload(rin, jc4, o_sxz + r1);
    pf_lam[k] = bload(rmod, jc4, m_lam + r1);
    pf_mu[k] = bload(rmod, jc4, m_mu + r1);
    pf_vbx[k] = bload(rin, jc4, o_vx + r0);
    pf_vbz[k] = bload(rin, jc4, o_vz + r0);
    pf_irho[k] = bload(rmod, jc4, m_irho + r0);
    pf_fx1[k] = bload(rf1, jc4, A_VX * npt4 + r0);
    pf_fz1[k] = bload(rf1, jc4, A_VZ * npt4 + r0);
    pf_kr[k] = own_r ? bload(rk, jc4, k_rho + rkr) : 0.0f;
    pf_kl[k] = own_l ? bload(rk, jc4, k_lam + rks) : 0.0f;
    pf_km[k] = own_m ? bload(rk, jc4, k_mu + rks) : 0.0f;
  }

  // one barrier per row: t of row i0-1 goes to lds_t[0]
  lds_t[0][0][lane] = tzxz_cur;
  lds_t[0][1][lane] = tzzz_cur;
  __syncthreads();

  auto body = [&](auto phase, int m) __attribute__((always_inline)) {
    constexpr int R = decltype(phase)::value;
    constexpr int s0 = R & 3, s1 = (R + 1) & 3, s2 = (R + 2) & 3, s3 = (R + 3) & 3, par = R & 1, f = R % D;
    sbxx[s3] = settle(pf_sxx[f]); sbzz[s3] = settle(pf_szz[f]); sbxz[s3] = settle(pf_sxz[f]);     // row m+1
    float n_xx, n_zz, n_xz;
    tvals(pf_lam[f], pf_mu[f], sbxx[s3], sbzz[s3], sbxz[s3], n_xx, n_zz, n_xz);
    txx[s3] = xt(m + 1) ? n_xx : 0.0f;
    txz[s3] = xt(m + 1) ? n_xz : 0.0f;
    const float tzxz_next = zt ? n_xz : 0.0f, tzzz_next = zt ? n_zz : 0.0f;
    const float vbx_o = pf_vbx[f], vbz_o = pf_vbz[f], irho = pf_irho[f];
    fx[s3] = settle(pf_fx1[f]);     // forward v^{n+1}, row m
    fz[s3] = settle(pf_fz1[f]);
    const float kr = pf_kr[f], kl = pf_kl[f], km = pf_km[f];
    const float bxm = bx_next;
    bx_next = taper_x(a.bx, m + 1, nx);
    {
      const uint32_t r3 = g.row(m + D + 1), r2 = g.row(m + D);
      const uint32_t rkr = g.row(m + D - KR_LAG), rks = g.row(m + D - KS_LAG);
      pf_sxx[f] = bload(rin, jc4, o_sxx + r3);
      pf_szz[f] = bload(rin, jc4, o_szz + r3);
      pf_sxz[f] = bload(rin, jc4, o_sxz + r3);
      pf_lam[f] = bload(rmod, jc4, m_lam + r3);
      pf_mu[f] = bload(rmod, jc4, m_mu + r3);
      pf_vbx[f] = bload(rin, jc4, o_vx + r2);
      pf_vbz[f] = bload(rin, jc4, o_vz + r2);
      pf_irho[f] = bload(rmod, jc4, m_irho + r2);
      pf_fx1[f] = bload(rf1, jc4, A_VX * npt4 + r2);
      pf_fz1[f] = bload(rf1, jc4, A_VZ * npt4 + r2);
      if (own_r) pf_kr[f] = bload(rk, jc4, k_rho + rkr);
      if (own_l) pf_kl[f] = bload(rk, jc4, k_lam + rks);
      if (own_m) pf_km[f] = bload(rk, jc4, k_mu + rks);
    }

    // adjoint "velocity" at row m: vbar'' = vbar' + inj + div
    float vbx = vbx_o, vbz = vbz_o, injx = 0.0f, injz = 0.0f;
    if (m >= inj_row) {
      for (; ip < ie; ++ip) {
        const int4 e = list_entry(a.inj, ip);
        if (e.x > m) break;
        if (e.x == m && e.y == j) {
          const float vx = uniform_value(a.val_x, int64_t(e.z) * a.nt_val + it), vz = uniform_value(a.val_z, int64_t(e.z) * a.nt_val + it);
          vbx = vbx + vx;
          vbz = vbz + vz;
          injx = injx + vx;
          injz = injz + vz;
        }
      }
      inj_row = next_row(a.inj, ip, ie);
    }
    const float divx = ndm(txx[s0], txx[s1], txx[s2], txx[s3], Dx) +
                       ndm(lds_t[par][0][tm2], lds_t[par][0][tm1], tzxz_cur, lds_t[par][0][tp1], Dz);
    const float divz = ndm(lds_t[par][1][tm2], lds_t[par][1][tm1], tzzz_cur, lds_t[par][1][tp1], Dz) +
                       ndm(txz[s0], txz[s1], txz[s2], txz[s3], Dx);
    vbx = vbx + divx;
    vbz = vbz + divz;
    const float b = (bxm * bzb) * bzt;
    const float gx = b * vbx, gz = b * vbz;
    const float qxv = (dt * gx) * irho, qzv = (dt * gz) * irho;   // dt g / rho
    // K_rho += -v^{n+1}.(vbar'' - vbar') / rho (the telescoped -dt g.ds / rho^2)
    const float cr = (-(fx[s3] * (injx + divx) + fz[s3] * (injz + divz))) * irho;
    if (g.jst && m >= i0 && m < i1) {
      const uint32_t ro = uint32_t(m) * g.rowb;
      bstore_out(gx, rout, jc4, o_vx + ro, a.stream_out);
      bstore_out(gz, rout, jc4, o_vz + ro, a.stream_out);
      if constexpr (S == 1) bstore(kr + cr, rk, jc4, k_rho + ro);
    }
    qx[s3] = xq(m) ? qxv : 0.0f;
    qz[s3] = xq(m) ? qzv : 0.0f;
    const float qxz = zq ? qxv : 0.0f, qzz = zq ? qzv : 0.0f;
    lds_q[par][0][lane] = qxz;
    lds_q[par][1][lane] = qzz;
    lds_t[par ^ 1][0][lane] = tzxz_next;
    lds_t[par ^ 1][1][lane] = tzzz_next;
    lds_f[s3][0][lane] = fx[s3];
    lds_f[s3][1][lane] = fz[s3];
    __syncthreads();
    dzqx[s2] = ndp(lds_q[par][0][tm1], qxz, lds_q[par][0][tp1], lds_q[par][0][tp2], Dz);
    dzqz[s2] = ndp(lds_q[par][1][tm1], qzz, lds_q[par][1][tp1], lds_q[par][1][tp2], Dz);
    tzxz_cur = tzxz_next;
    tzzz_cur = tzzz_next;

    // adjoint "stress" at row i = m-2, and the strains of v^{n+1} there
    const int i = m - 2;
    if (i >= i0 && g.jst) {
      const uint32_t ro = uint32_t(i) * g.rowb;
      const float nxx = sbxx[s0] + ndp(qx[s0], qx[s1], qx[s2], qx[s3], Dx);
      const float nzz = sbzz[s0] + dzqz[s0];
      const float nxz = sbxz[s0] + (dzqx[s0] + ndp(qz[s0], qz[s1], qz[s2], qz[s3], Dx));
      bstore_out(nxx, rout, jc4, o_sxx + ro, a.stream_out);
      bstore_out(nzz, rout, jc4, o_szz + ro, a.stream_out);
      bstore_out(nxz, rout, jc4, o_sxz + ro, a.stream_out);
    }
    const bool xs = xt(i);
    const float exx = xs ? dpf(fx[s0], fx[s1], fx[s2], fx[s3], Dx) : 0.0f;
    const float dvzdx = xs ? dpf(fz[s0], fz[s1], fz[s2], fz[s3], Dx) : 0.0f;
    const float ezz = zt ? dpf(lds_f[s1][1][tm1], fz[s1], lds_f[s1][1][tp1], lds_f[s1][1][tp2], Dz) : 0.0f;
    const float dvxdz = zt ? dpf(lds_f[s1][0][tm1], fx[s1], lds_f[s1][0][tp1], lds_f[s1][0][tp2], Dz) : 0.0f;
    const float exz = dvxdz + dvzdx;
    const float cl = dt * (sbxx[s0] + sbzz[s0]) * (exx + ezz);
    const float cm = dt * (2.0f * sbxx[s0] * exx + 2.0f * sbzz[s0] * ezz + sbxz[s0] * exz);
    if constexpr (S == 1) {
      if (i >= i0 && g.jst) {
        const uint32_t ro = uint32_t(i) * g.rowb;
        bstore(kl + cl, rk, jc4, k_lam + ro);
        bstore(km + cm, rk, jc4, k_mu + ro);
      }
    } else {
      // the previous iteration's contributions (K_rho row m-1, K_lambda and
      // K_mu row m-3), complete in LDS since this iteration's barrier
      if (g.jst) {
        if (own_r && m - 1 >= i0 && m - 1 < i1)
          bstore(kr + shot_sum<S, NT>(lds_k[par ^ 1][0], t), rk, jc4, k_rho + uint32_t(m - 1) * g.rowb);
        if (m - 3 >= i0) {
          if (own_l) bstore(kl + shot_sum<S, NT>(lds_k[par ^ 1][1], t), rk, jc4, k_lam + uint32_t(m - 3) * g.rowb);
          if (own_m) bstore(km + shot_sum<S, NT>(lds_k[par ^ 1][2], t), rk, jc4, k_mu + uint32_t(m - 3) * g.rowb);
        }
      }
      lds_k[par][0][lane] = cr;
      lds_k[par][1][lane] = cl;
      lds_k[par][2][lane] = cm;
    }
  };

  const int mlast = i1 + 1;
  for (int m = i0 - 1;;) {
    body(Phase<0>{}, m);
    if (++m > mlast) break;
    body(Phase<1>{}, m);
    if (++m > mlast) break;
    body(Phase<2>{}, m);
    if (++m > mlast) break;
    body(Phase<3>{}, m);
    if (++m > mlast) break;
  }
  if constexpr (S > 1) {
    // the last iteration's K_lambda / K_mu contributions (row i1 - 1)
    __syncthreads();
    const int par = (mlast - (i0 - 1)) & 1, r = mlast - 2;
    if (g.jst && r >= i0) {
      if (own_l) bstore(pf_kl[0] + shot_sum<S, NT>(lds_k[par][1], t), rk, jc4, k_lam + uint32_t(r) * g.rowb);
      if (own_m) bstore(pf_km[0] + shot_sum<S, NT>(lds_k[par][2], t), rk, jc4, k_mu + uint32_t(r) * g.rowb);
    }
  }
  if (a.advance) advance_clock(a.clock, it, -1, a.nblocks);
}

// ------------------------------------------- receivers of two-step launches
// Traces of a two-step marching launch (steps it and it+1; it = clock - 2,
// the launch has advanced the clock), one thread per (shot, receiver entry of
// the launch's tile lists, whose fourth field is the owning tile).  The
// launch keeps the velocities of step it+1 on chip, so their receiver-cell
// values are recomputed here from the launch's input state st[0] with the
// operations of its first stage (same operands, same order: the same bits);
// those of step it+2 are read from its output st[1].  Then, as one-step
// launches do: u += v dt at step it, trace(it) = u, again with the step it+2
// velocities, trace(it+1) = u.
template <bool SH>
__global__ __launch_bounds__(64) void k_record2(const StepArgs a, int nent) {
  const int64_t idx = int64_t(blockIdx.x) * 64 + threadIdx.x;
  if (idx >= int64_t(nent) * a.nshots) return;
  const int shot = int(idx / nent);
  const int4 e = a.rec[idx - int64_t(shot) * nent];
  const int i = e.x, j = e.y, r = e.z, tile = e.w;
  const int it = a.clock[0] - 2;
  const int nx = a.nx, nz = a.nz;
  const int64_t npt = a.npt, k = int64_t(i) * nz + j;
  const float* s_in = a.st[0] + int64_t(shot) * a.shot_floats;
  const float* s_out = a.st[1] + int64_t(shot) * a.shot_floats;
  // a state value at a row / column clamped to the grid, as the march loads it
  auto at = [&](const float* f, int ii, int jj) {
    return f[int64_t(min(max(ii, 0), nx - 1)) * nz + min(max(jj, 0), nz - 1)];
  };
  const CDiv Dx{a.h8x, a.r8x, a.z8x, a.l8x, a.c_lo, (a.fast_div & FD_CONST) != 0},
             Dz{a.h8z, a.r8z, a.z8z, a.l8z, a.c_lo, (a.fast_div & FD_CONST) != 0};
  const int fast_c = div_mode(a.fast_div, FD_CONST);
  const int fast_rho = div_mode(a.fast_div, FD_RHO);
  const bool xm_ok = i >= 2 && i < nx - 2, zm_ok = j >= 2 && j < nz - 2;
  const float dt = a.dt;
  const float b = (a.bx[min(max(i, 0), nx - 1)] * a.bzb[j]) * a.bzt[j];
  const float rho = a.model[MF_RHO * npt + k], irho = a.model[MF_IRHO * npt + k];
  const int tps = a.nstrips * a.nsegs;
  const int q0 = a.inj_off[shot * tps + tile], q1 = a.inj_off[shot * tps + tile + 1];
  float* u = a.uacc + (int64_t(shot) * a.nrec + r) * N_UACC;
  float* obs = a.obs + int64_t(shot) * 5 * a.nrec * a.nt;
  if constexpr (SH) {
    // slots vy, sxy, szy (as the march reads them)
    const float *vy = s_in, *sxy = s_in + npt, *szy = s_in + 2 * npt;
    // every load before the first store (the trace and displacement stores
    // may alias the state for the compiler, which would then order the later
    // loads behind them)
    const float v2 = s_out[k], u0 = u[U_UY];
    float qa[2];
    {
      const float n0 = nminus(at(sxy, i - 2, j), at(sxy, i - 1, j), at(sxy, i, j), at(sxy, i + 1, j));
      const float n1 = nminus(at(szy, i, j - 2), at(szy, i, j - 1), at(szy, i, j), at(szy, i, j + 1));
      const float xa[2] = {keep(xm_ok, n0), keep(zm_ok, n1)};
      div_group(qa, xa, {Dx, Dz}, fast_c);
    }
    float dsy = qa[0] + qa[1];
    for (int q = q0; q < q1; ++q) {
      const int4 f = a.inj[q];
      if (f.x > i) break;
      if (f.x == i && f.y == j) dsy = dsy + a.val_y[int64_t(f.z) * a.nt_val + it];
    }
    float qb[1];
    {
      const float xb[1] = {dt * dsy};
      div_group(qb, xb, {Div{rho, irho, true}}, fast_rho);
    }
    const float v1 = b * (vy[k] + qb[0]);
    const float uy = u0 + v1 * dt;
    obs[(1 * a.nrec + r) * a.nt + it] = uy;
    const float uy2 = uy + v2 * dt;
    obs[(1 * a.nrec + r) * a.nt + it + 1] = uy2;
    u[U_UY] = uy2;
  } else {
    // slots vx, vz, sxx, szz, sxz
    const float *vx = s_in, *vz = s_in + npt, *sxx = s_in + 2 * npt, *szz = s_in + 3 * npt, *sxz = s_in + 4 * npt;
    const float v2x = s_out[k], v2z = s_out[npt + k], u0x = u[U_UX], u0z = u[U_UZ];   // loads before stores
    float qa[4];
    {
      const float n0 = nminus(at(sxx, i - 2, j), at(sxx, i - 1, j), at(sxx, i, j), at(sxx, i + 1, j));
      const float n1 = nminus(at(sxz, i, j - 2), at(sxz, i, j - 1), at(sxz, i, j), at(sxz, i, j + 1));
      const float n2 = nminus(at(sxz, i - 2, j), at(sxz, i - 1, j), at(sxz, i, j), at(sxz, i + 1, j));
      const float n3 = nminus(at(szz, i, j - 2), at(szz, i, j - 1), at(szz, i, j), at(szz, i, j + 1));
      const float xa[4] = {keep(xm_ok, n0), keep(zm_ok, n1), keep(xm_ok, n2), keep(zm_ok, n3)};
      div_group(qa, xa, {Dx, Dz, Dx, Dz}, fast_c);
    }
    float dsx = qa[0] + qa[1];
    float dsz = qa[2] + qa[3];
    for (int q = q0; q < q1; ++q) {
      const int4 f = a.inj[q];
      if (f.x > i) break;
      if (f.x == i && f.y == j) {
        dsx = dsx + a.val_x[int64_t(f.z) * a.nt_val + it];
        dsz = dsz + a.val_z[int64_t(f.z) * a.nt_val + it];
      }
    }
    float qb[2];
    {
      const float xb[2] = {dt * dsx, dt * dsz};
      div_group(qb, xb, {Div{rho, irho, true}, Div{rho, irho, true}}, fast_rho);
    }
    const float v1x = b * (vx[k] + qb[0]), v1z = b * (vz[k] + qb[1]);
    const int nobs = a.nrec * a.nt;
    const float ux = u0x + v1x * dt, uz = u0z + v1z * dt;
    obs[0 * nobs + r * a.nt + it] = ux;
    obs[2 * nobs + r * a.nt + it] = uz;
    const float ux2 = ux + v2x * dt, uz2 = uz + v2z * dt;
    obs[0 * nobs + r * a.nt + it + 1] = ux2;
    obs[2 * nobs + r * a.nt + it + 1] = uz2;
    u[U_UX] = ux2;
    u[U_UZ] = uz2;
  }
}

// --------------------------------------------------------- entry points
template <int NT, int S, bool AUX>
__global__ __launch_bounds__(NT * S) void k_step_sh(const StepArgs a) {
  step_sh<NT, S, AUX>(a, xcd_remap(blockIdx.x, gridDim.x));
}
// P-SV without spin fits 80 VGPRs with the 2-deep prefetch: 6 waves per
// SIMD; four shots per workgroup (16 waves, one workgroup per CU) trade
// occupancy for the 4-deep prefetch
template <int NT, int S, bool AUX>
__global__ __launch_bounds__(NT * S) __attribute__((amdgpu_waves_per_eu(S >= 4 ? 4 : S == 2 ? 5 : 6))) void k_step_psv(
    const StepArgs a) {
  step_psv<NT, S, false, AUX>(a, xcd_remap(blockIdx.x, gridDim.x));
}
template <int NT, int S, bool AUX, bool REC = false>
__global__ __launch_bounds__(NT * S) __attribute__((amdgpu_waves_per_eu(4))) void k_step_psv2(
    const StepArgs a) {
  step_psv2<NT, S, AUX, REC>(a, xcd_remap(blockIdx.x, gridDim.x));
}
template <int NT, int S>
__global__ __launch_bounds__(NT * S) __attribute__((amdgpu_waves_per_eu(3))) void k_step_psv3(
    const StepArgs a) {
  step_psv3<NT, S>(a, xcd_remap(blockIdx.x, gridDim.x));
}
template <int NT, int S, bool AUX, bool REC = false>
__global__ __launch_bounds__(NT * S) __attribute__((amdgpu_waves_per_eu(4))) void k_step_sh2(const StepArgs a) {
  step_sh2<NT, S, AUX, REC>(a, xcd_remap(blockIdx.x, gridDim.x));
}
template <int PH, int K>
__global__ __launch_bounds__(1024) void k_tile2(const StepArgs a) {
  tile2<PH, K>(a, xcd_remap(blockIdx.x, gridDim.x));
}
template <int NT, int S>
__global__ __launch_bounds__(NT * S) void k_step_spin(const StepArgs a) {
  step_psv<NT, S, true, false>(a, xcd_remap(blockIdx.x, gridDim.x));
}
template <int NT, int S>
__global__ __launch_bounds__(NT * S) void k_adj_sh(const StepArgs a) {
  adj_sh<NT, S>(a, xcd_remap(blockIdx.x, gridDim.x));
}
// the P-SV adjoint keeps 17 fields in flight per row: 4 waves per SIMD
template <int NT, int S>
__global__ __launch_bounds__(NT * S) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_psv(const StepArgs a) {
  adj_psv<NT, S>(a, xcd_remap(blockIdx.x, gridDim.x));
}
// multi-problem entry point: one launch advances several problems
template <int NT, int KIND>
__global__ __launch_bounds__(NT) void k_multi(const StepArgs* __restrict__ args, const int2* __restrict__ work) {
  const int2 w = work[xcd_remap(blockIdx.x, gridDim.x)];
  const StepArgs& a = args[w.x];
  if constexpr (KIND == MK_SH) step_sh<NT, 1, false>(a, w.y);
  else if constexpr (KIND == MK_PSV) step_psv<NT, 1, false, false>(a, w.y);
  else if constexpr (KIND == MK_SPIN) step_psv<NT, 1, true, false>(a, w.y);
  else if constexpr (KIND == MK_SH_AUX) step_sh<NT, 1, true>(a, w.y);
  else if constexpr (KIND == MK_PSV_AUX) step_psv<NT, 1, false, true>(a, w.y);
  else if constexpr (KIND == MK_ADJ_SH) adj_sh<NT, 1>(a, w.y);
  else adj_psv<NT, 1>(a, w.y);
}

template <int NT>
hipError_t launch_multi_nt(int kind, const StepArgs* args, const int2* work, int nb, hipStream_t s) {
  switch (kind) {
    case MK_SH: hipLaunchKernelGGL((k_multi<NT, MK_SH>), dim3(nb), dim3(NT), 0, s, args, work); break;
    case MK_PSV: hipLaunchKernelGGL((k_multi<NT, MK_PSV>), dim3(nb), dim3(NT), 0, s, args, work); break;
    case MK_SPIN: hipLaunchKernelGGL((k_multi<NT, MK_SPIN>), dim3(nb), dim3(NT), 0, s, args, work); break;
    case MK_SH_AUX: hipLaunchKernelGGL((k_multi<NT, MK_SH_AUX>), dim3(nb), dim3(NT), 0, s, args, work); break;
    case MK_PSV_AUX: hipLaunchKernelGGL((k_multi<NT, MK_PSV_AUX>), dim3(nb), dim3(NT), 0, s, args, work); break;
    case MK_ADJ_SH: hipLaunchKernelGGL((k_multi<NT, MK_ADJ_SH>), dim3(nb), dim3(NT), 0, s, args, work); break;
    case MK_ADJ_PSV: hipLaunchKernelGGL((k_multi<NT, MK_ADJ_PSV>), dim3(nb), dim3(NT), 0, s, args, work); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// adjoint sources: one thread per (shot, component, receiver) scans time
// backwards, A^n = sum_{k>=n} 2 w_k^2 (d^k - obs^k) (double accumulation),
// writes dt * A^n into adj[comp][shot*nrec + r][n]; per-shot misfit partials
// are reduced in a fixed order by a second kernel.
// Wavefront-level sums (64 lanes, fixed order: lane pairs at distance 32, 16,
// ..., 1, so results do not depend on timing).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;   // lane 0 holds the sum
}
// Inclusive suffix sum over the lanes: lane l gets v_l + v_{l+1} + ... + v_63.
__device__ __forceinline__ double wave_suffix_sum(double v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double y = __shfl_down(v, off, 64);
    if (lane + off < 64) v += y;
  }
  return v;
}

// Misfit and adjoint source of the gradient (objective J = sum (w (syn -
// obs))^2, w the cosine taper of compute_waveform_misfit, solver.rs:441-455):
// one wavefront per trace (shot, component, receiver).  The adjoint source
// at step n is dt * sum_{m >= n} 2 w_m^2 (syn_m - obs_m) (the adjoint of the
// displacement accumulation u^{n+1} = u^n + dt v): the trace is walked in
// 64-sample tiles from its end, each tile's suffix sums formed by a
// wavefront scan and carried into the tile before it, so loads and stores
// are coalesced; the squared misfit is summed per lane and reduced across
// the wavefront at the end (part[trace]).
__global__ __launch_bounds__(256) void k_adjsrc(const float* __restrict__ obs_syn, const float* __restrict__ obs_ref,
                                                const float* __restrict__ taper, float* __restrict__ adj,
                                                double* __restrict__ part, int nshots, int ncomp,
                                                const int* __restrict__ comps, int nrec, int nt, float dt) {
  const int lane = int(threadIdx.x) & 63;
  const int g = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * 4 + int(threadIdx.x >> 6));
  const int total = nshots * ncomp * nrec;
  if (g >= total) return;
  const int r = g % nrec, c = (g / nrec) % ncomp, sh = g / (nrec * ncomp);
  const float* syn = obs_syn + ((int64_t(sh) * 5 + comps[c]) * nrec + r) * nt;
  const float* ref = obs_ref + ((int64_t(sh) * ncomp + c) * nrec + r) * nt;
  float* out = adj + (int64_t(c) * nshots * nrec + int64_t(sh) * nrec + r) * nt;
  double carry = 0.0, mis = 0.0;
  for (int t0 = ((nt - 1) / 64) * 64; t0 >= 0; t0 -= 64) {
    const int n = t0 + lane;
    double v = 0.0;
    if (n < nt) {
      const double w = taper[n];
      const double d = double(syn[n]) - double(ref[n]);
      v = 2.0 * w * w * d;
      mis += (w * d) * (w * d);
    }
    const double suf = wave_suffix_sum(v, lane);
    if (n < nt) out[n] = float(double(dt) * (carry + suf));
    carry += __shfl(suf, 0, 64);
  }
  mis = wave_sum(mis);
  if (lane == 0) part[g] = mis;
}
// per-shot misfit: one wavefront per shot sums its traces' partials
__global__ __launch_bounds__(64) void k_misfit_shots(const double* __restrict__ part, double* __restrict__ mis,
                                                     int nshots, int per) {
  const int sh = blockIdx.x, lane = threadIdx.x;
  double acc = 0.0;
  for (int q = lane; q < per; q += 64) acc += part[int64_t(sh) * per + q];
  acc = wave_sum(acc);
  if (lane == 0) mis[sh] = acc;
}
__global__ void k_reduce_shots(const float* __restrict__ kshot, float* __restrict__ kout, int nshots, int64_t n,
                               int64_t shot_stride) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float acc = kout[k];
  for (int s = 0; s < nshots; ++s) acc = acc + kshot[int64_t(s) * shot_stride + k];
  kout[k] = acc;
}

// ------------------------------------------------ plain SH imaging condition
// div_uy, div_fw, interaction_muy (fd2d.wgsl:501-509, 571-588)
__global__ void k_imaging_sh(const float* __restrict__ va, const float* __restrict__ vf, float* __restrict__ kmu,
                             int nx, int nz, float h8x, float h8z, float r8x, float r8z, float ndt,
                             int fast_div) {
  const Div Dx{h8x, r8x, (fast_div & FD_CONST) != 0}, Dz{h8z, r8z, (fast_div & FD_CONST) != 0};
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  const int k = i * nz + j;
  const bool xp = i >= 1 && i < nx - 2, zp = j >= 1 && j < nz - 2;
  float ax = 0.0f, az = 0.0f, fx = 0.0f, fz = 0.0f;
  if (xp) {
    ax = dplus(va[k - nz], va[k], va[k + nz], va[k + 2 * nz], Dx);
    fx = dplus(vf[k - nz], vf[k], vf[k + nz], vf[k + 2 * nz], Dx);
  }
  if (zp) {
    az = dplus(va[k - 1], va[k], va[k + 1], va[k + 2], Dz);
    fz = dplus(vf[k - 1], vf[k], vf[k + 1], vf[k + 2], Dz);
  }
  kmu[k] = kmu[k] + (ax * fx + az * fz) * ndt;
}

// ----------------------------------------------------- Gaussian smoothing
// init_gsum, gaussian_x, gaussian_z (fd2d.wgsl:517-554, 592-609).  w[d] is the
// 1-D weight exp(-d*d/(2*sigma*sigma)) tabulated by the host, and w[d] == 0
// for every d >= nw.  The reference sums over the full axis; the terms with
// |i - n| >= nw are 0 * k = +-0, and adding +-0 to a sum that starts at +0
// (which can never become -0 in round-to-nearest) leaves it unchanged, so
// summing over |i - n| < nw only, in the same order, gives the same bits
// for finite fields at O(npt * nw) instead of O(npt * (nx + nz)).
__global__ void k_gauss_x(const float* __restrict__ w, int nw, const float* __restrict__ kin,
                          float* __restrict__ gtmp, int nx, int nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  float s = 0.0f;
  const int lo = max(i - nw + 1, 0), hi = min(i + nw, nx);
  for (int n = lo; n < hi; ++n) s = s + w[abs(i - n)] * kin[int64_t(n) * nz + j];
  gtmp[int64_t(i) * nz + j] = s;
}
// GSUM (init_gsum) is the product of the 1-D weight sums, formed per cell
__global__ void k_gauss_z(const float* __restrict__ w, int nw, const float* __restrict__ gtmp,
                          const float* __restrict__ sumx, const float* __restrict__ sumz, float* __restrict__ kout,
                          int nx, int nz) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= nz) return;
  float s = 0.0f;
  const float* row = gtmp + int64_t(i) * nz;
  const int lo = max(j - nw + 1, 0), hi = min(j + nw, nz);
  for (int n = lo; n < hi; ++n) s = s + w[abs(j - n)] * row[n];
  const float g = sumx[i] * sumz[j];
  kout[int64_t(i) * nz + j] = (g > 1e-30f) ? s / g : 0.0f;
}

// ------------------------------------------ waveform misfit + adjoint source
// compute_waveform_misfit (solver.rs:432-467): adstf time-reversed; the
// squared-misfit partial sums are reduced per block in a fixed order.
__global__ __launch_bounds__(256) void k_misfit(const float* __restrict__ syn, const float* __restrict__ ref,
                                                const float* __restrict__ taper, float* __restrict__ adstf,
                                                double* __restrict__ partial, int nrec, int nt) {
  __shared__ double red[4];
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
  // wavefront sums, then the block's four in order
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}
__global__ __launch_bounds__(256) void k_reduce_sum(const double* __restrict__ partial, int n, double* out) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) acc += partial[q];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = ((red[0] + red[1]) + red[2]) + red[3];
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

__global__ void k_recip(const float* __restrict__ x, float* __restrict__ out, int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) out[k] = 1.0f / x[k];
}

__global__ void k_fill(float* __restrict__ p, float v, int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}
__global__ void k_fill_f64(double* __restrict__ p, double v, int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// Blow-up check: *bad = min(*bad, tag) when any of the n values is NaN or
// infinite.  Each wavefront votes once, so at most one atomic per wavefront.
__global__ __launch_bounds__(256) void k_nonfinite(const float* __restrict__ p, int64_t n, int tag, int* bad) {
  bool nf = false;
  const int64_t stride = int64_t(gridDim.x) * 256;
  for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < n; k += stride)
    nf |= (__builtin_bit_cast(uint32_t, p[k]) & 0x7f800000u) == 0x7f800000u;
  if (__builtin_amdgcn_ballot_w64(nf) != 0 && __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)) == 0)
    atomicMin(bad, tag);
}

// plain streaming copy (float4), the calibration kernel for HBM counters
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n4) {
  for (int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x; k < n4; k += int64_t(gridDim.x) * 256) dst[k] = src[k];
}

// forward step launch for a tile width NT and S shots per workgroup; kind is
// MK_SH, MK_SH_AUX, MK_PSV, MK_PSV_AUX or MK_SPIN
// The grid and the tile geometry the kernels derive from the arguments must
// describe the same tiles: a mismatch would map workgroups to shots past the
// batch.
bool args_match(const StepArgs& a, const Tiling& t, int S) {
  return a.nstrips == t.nstrips && a.nsegs == t.nsegs && a.seglen == t.seglen && S >= 1 && a.nshots % S == 0 &&
         a.nblocks == a.nshots / S * t.nstrips * t.nsegs;
}

template <int NT, int S>
hipError_t launch_fwd_ns(int kind, const StepArgs& a, const Tiling& t, hipStream_t s) {
  if (!args_match(a, t, S)) return hipErrorInvalidValue;
  const int nblk = a.nshots / S * t.nstrips * t.nsegs;
  const dim3 grid(nblk), block(NT * S);
  switch (kind) {
    case MK_SH: hipLaunchKernelGGL((k_step_sh<NT, S, false>), grid, block, 0, s, a); break;
    case MK_SH_AUX: hipLaunchKernelGGL((k_step_sh<NT, S, true>), grid, block, 0, s, a); break;
    case MK_PSV: hipLaunchKernelGGL((k_step_psv<NT, S, false>), grid, block, 0, s, a); break;
    case MK_PSV_AUX: hipLaunchKernelGGL((k_step_psv<NT, S, true>), grid, block, 0, s, a); break;
    case MK_SPIN:
      // the spin kernel needs more than 64 VGPRs: at most 512 threads per workgroup
      if constexpr (NT * S <= 512) hipLaunchKernelGGL((k_step_spin<NT, S>), grid, block, 0, s, a);
      else return hipErrorInvalidValue;
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_fwd(int kind, const StepArgs& a, const Tiling& t, hipStream_t s) {
  if (t.group < 1 || a.nshots % t.group != 0) return hipErrorInvalidValue;
  switch (t.nt_threads * 8 + t.group) {
    case 64 * 8 + 1: return launch_fwd_ns<64, 1>(kind, a, t, s);
    case 128 * 8 + 1: return launch_fwd_ns<128, 1>(kind, a, t, s);
    case 128 * 8 + 2: return launch_fwd_ns<128, 2>(kind, a, t, s);
    case 128 * 8 + 4: return launch_fwd_ns<128, 4>(kind, a, t, s);
    case 256 * 8 + 1: return launch_fwd_ns<256, 1>(kind, a, t, s);
    case 256 * 8 + 2: return launch_fwd_ns<256, 2>(kind, a, t, s);
    case 256 * 8 + 4: return launch_fwd_ns<256, 4>(kind, a, t, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool tiling_supported(int nt) { return nt == 64 || nt == 128 || nt == 256; }
bool group_supported(int nt, int group) { return group == 1 || (nt >= 128 && (group == 2 || group == 4)); }

namespace {
template <typename K>
int occupancy(K k, int nt) {
  int n = 0;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, nt, 0) == hipSuccess ? n : 0;
}
template <int NT, int S>
int blocks_per_cu_ns(int kind) {
  switch (kind) {
    case MK_SH: return occupancy(k_step_sh<NT, S, false>, NT * S);
    case MK_PSV: return occupancy(k_step_psv<NT, S, false>, NT * S);
    case MK_SPIN:
      if constexpr (NT * S <= 512) return occupancy(k_step_spin<NT, S>, NT * S);
      return 0;
    case MK_SH_AUX: return occupancy(k_step_sh<NT, S, true>, NT * S);
    case MK_PSV_AUX: return occupancy(k_step_psv<NT, S, true>, NT * S);
  }
  if (kind == MK_ADJ_SH) return occupancy(k_adj_sh<NT, S>, NT * S);
  if (kind == MK_ADJ_PSV) return occupancy(k_adj_psv<NT, S>, NT * S);
  return 0;
}
}  // namespace

int step_blocks_per_cu(int kind, int nt_threads, int group) {
  switch (nt_threads * 8 + group) {
    case 64 * 8 + 1: return blocks_per_cu_ns<64, 1>(kind);
    case 128 * 8 + 1: return blocks_per_cu_ns<128, 1>(kind);
    case 128 * 8 + 2: return blocks_per_cu_ns<128, 2>(kind);
    case 128 * 8 + 4: return blocks_per_cu_ns<128, 4>(kind);
    case 256 * 8 + 1: return blocks_per_cu_ns<256, 1>(kind);
    case 256 * 8 + 2: return blocks_per_cu_ns<256, 2>(kind);
    case 256 * 8 + 4: return blocks_per_cu_ns<256, 4>(kind);
  }
  return 0;
}

hipError_t launch_step_sh(const StepArgs& a, const Tiling& t, hipStream_t s, bool aux) {
  return launch_fwd(aux ? MK_SH_AUX : MK_SH, a, t, s);
}

// the state layout the two-step kernels take as fixed: one physics, its
// slots first to last
bool two_step_slots(bool sh, const StepArgs& a) {
  const int n = a.npt;
  if (sh) return a.slot_off[S_VY] == 0 && a.slot_off[S_SXY] == n && a.slot_off[S_SZY] == 2 * n;
  return a.slot_off[S_VX] == 0 && a.slot_off[S_VZ] == n && a.slot_off[S_SXX] == 2 * n &&
         a.slot_off[S_SZZ] == 3 * n && a.slot_off[S_SXZ] == 4 * n;
}

template <int NT, int S>
hipError_t launch_two_ns(bool sh, const StepArgs& a, const Tiling& t, hipStream_t s, bool aux, bool rec) {
  if (!args_match(a, t, S) || !two_step_slots(sh, a)) return hipErrorInvalidValue;
  const int nblk = a.nshots / S * t.nstrips * t.nsegs;
  const dim3 grid(nblk), block(NT * S);
  if (sh && aux) hipLaunchKernelGGL((k_step_sh2<NT, S, true>), grid, block, 0, s, a);
  else if (sh && rec) hipLaunchKernelGGL((k_step_sh2<NT, S, false, true>), grid, block, 0, s, a);
  else if (sh) hipLaunchKernelGGL((k_step_sh2<NT, S, false>), grid, block, 0, s, a);
  else if (aux) hipLaunchKernelGGL((k_step_psv2<NT, S, true>), grid, block, 0, s, a);
  else if (rec) hipLaunchKernelGGL((k_step_psv2<NT, S, false, true>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((k_step_psv2<NT, S, false>), grid, block, 0, s, a);
  return hipGetLastError();
}
hipError_t launch_two(bool sh, const StepArgs& a, const Tiling& t, hipStream_t s, bool aux, bool rec) {
  if (t.group < 1 || a.nshots % t.group != 0) return hipErrorInvalidValue;
  switch (t.nt_threads * 8 + t.group) {
    case 64 * 8 + 1: return launch_two_ns<64, 1>(sh, a, t, s, aux, rec);
    case 128 * 8 + 1: return launch_two_ns<128, 1>(sh, a, t, s, aux, rec);
    case 128 * 8 + 2: return launch_two_ns<128, 2>(sh, a, t, s, aux, rec);
    case 128 * 8 + 4: return launch_two_ns<128, 4>(sh, a, t, s, aux, rec);
    case 256 * 8 + 1: return launch_two_ns<256, 1>(sh, a, t, s, aux, rec);
    case 256 * 8 + 2: return launch_two_ns<256, 2>(sh, a, t, s, aux, rec);
    case 256 * 8 + 4: return launch_two_ns<256, 4>(sh, a, t, s, aux, rec);
    case 320 * 8 + 1: return launch_two_ns<320, 1>(sh, a, t, s, aux, rec);
    case 320 * 8 + 2: return launch_two_ns<320, 2>(sh, a, t, s, aux, rec);
    case 512 * 8 + 1: return launch_two_ns<512, 1>(sh, a, t, s, aux, rec);
    case 512 * 8 + 2: return launch_two_ns<512, 2>(sh, a, t, s, aux, rec);
  }
  return hipErrorInvalidValue;
}
// K steps on LDS tiles (small grids): the tile geometry is fixed by the
// kernel (lds_tile_rows(K)-row segments, lds_tile_cols(K)-column strips), one
// shot per workgroup
hipError_t launch_tile2(int phys, const StepArgs& a, const Tiling& t, hipStream_t s) {
  const int K = t.lds_steps;
  if (t.nt_threads != kLdsTileThreads || K < 2 || K > 4 || t.seglen != lds_tile_rows(K) || t.group != 1 ||
      !args_match(a, t, 1) || a.nstrips != (a.nz + lds_tile_cols(K) - 1) / lds_tile_cols(K) ||
      a.nsegs != (a.nx + lds_tile_rows(K) - 1) / lds_tile_rows(K))
    return hipErrorInvalidValue;
  const dim3 grid(a.nshots * t.nstrips * t.nsegs), block(kLdsTileThreads);
  switch (phys * 8 + K) {
    case TP_SH * 8 + 2: hipLaunchKernelGGL((k_tile2<TP_SH, 2>), grid, block, 0, s, a); break;
    case TP_SH * 8 + 3: hipLaunchKernelGGL((k_tile2<TP_SH, 3>), grid, block, 0, s, a); break;
    case TP_PSV * 8 + 2: hipLaunchKernelGGL((k_tile2<TP_PSV, 2>), grid, block, 0, s, a); break;
    case TP_PSV * 8 + 3: hipLaunchKernelGGL((k_tile2<TP_PSV, 3>), grid, block, 0, s, a); break;
    case TP_SH * 8 + 4: hipLaunchKernelGGL((k_tile2<TP_SH, 4>), grid, block, 0, s, a); break;
    case TP_PSV * 8 + 4: hipLaunchKernelGGL((k_tile2<TP_PSV, 4>), grid, block, 0, s, a); break;
    case TP_SPIN * 8 + 2: hipLaunchKernelGGL((k_tile2<TP_SPIN, 2>), grid, block, 0, s, a); break;
    case TP_SPIN * 8 + 3: hipLaunchKernelGGL((k_tile2<TP_SPIN, 3>), grid, block, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
// three steps per launch (marching tiles, P-SV alone, one shot per workgroup)
hipError_t launch_three(const StepArgs& a, const Tiling& t, hipStream_t s) {
  if (t.group != 1 || !args_match(a, t, 1) || !two_step_slots(false, a)) return hipErrorInvalidValue;
  const dim3 grid(a.nshots * t.nstrips * t.nsegs);
  switch (t.nt_threads) {
    case 256: hipLaunchKernelGGL((k_step_psv3<256, 1>), grid, dim3(256), 0, s, a); break;
    case 384: hipLaunchKernelGGL((k_step_psv3<384, 1>), grid, dim3(384), 0, s, a); break;
    case 512: hipLaunchKernelGGL((k_step_psv3<512, 1>), grid, dim3(512), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
bool tiling3_supported(int nt) { return nt == 256 || nt == 384 || nt == 512; }
int step3_blocks_per_cu(int nt) {
  switch (nt) {
    case 256: return occupancy(k_step_psv3<256, 1>, 256);
    case 384: return occupancy(k_step_psv3<384, 1>, 384);
    case 512: return occupancy(k_step_psv3<512, 1>, 512);
  }
  return 0;
}

hipError_t launch_step_psv2(const StepArgs& a, const Tiling& t, hipStream_t s, bool aux, bool spin, bool rec) {
  if (t.nt_threads == kLdsTileThreads) return aux ? hipErrorInvalidValue : launch_tile2(spin ? TP_SPIN : TP_PSV, a, t, s);
  if (spin) return hipErrorInvalidValue;
  if (t.msteps == 3) return aux ? hipErrorInvalidValue : launch_three(a, t, s);
  return launch_two(false, a, t, s, aux, rec);
}
hipError_t launch_record2(bool sh, const StepArgs& a, int nent, hipStream_t s) {
  const int64_t n = int64_t(nent) * a.nshots;
  if (n == 0) return hipSuccess;
  // one wavefront per workgroup: the receivers spread over every CU (the
  // kernel waits on scattered loads, not on arithmetic)
  const dim3 grid(unsigned((n + 63) / 64));
  if (sh) hipLaunchKernelGGL(k_record2<true>, grid, dim3(64), 0, s, a, nent);
  else hipLaunchKernelGGL(k_record2<false>, grid, dim3(64), 0, s, a, nent);
  return hipGetLastError();
}
hipError_t launch_step_sh2(const StepArgs& a, const Tiling& t, hipStream_t s, bool aux, bool rec) {
  if (t.nt_threads == kLdsTileThreads) return aux ? hipErrorInvalidValue : launch_tile2(TP_SH, a, t, s);
  return launch_two(true, a, t, s, aux, rec);
}
int lds_tile_rows(int k) { return TL_R - 6 * k; }
int lds_tile_cols(int k) { return TL_C - 6 * k; }
// 320 (five waves, 308-column strips) fits grids of ~600 columns in two strips
bool tiling2_supported(int nt) {
  return nt == 64 || nt == 128 || nt == 256 || nt == 320 || nt == 512 || nt == kLdsTileThreads;
}
bool group2_supported(int nt, int group) {
  if (nt == kLdsTileThreads) return group == 1;
  return tiling2_supported(nt) && (group == 1 || (nt >= 128 && (group == 2 || group == 4) && nt * group <= 1024));
}
int lds_tile_blocks_per_cu(bool sh, bool spin, int k) {
  int n = 0;
  switch ((sh ? TP_SH : spin ? TP_SPIN : TP_PSV) * 8 + k) {
    case TP_SH * 8 + 2: n = occupancy(k_tile2<TP_SH, 2>, kLdsTileThreads); break;
    case TP_SH * 8 + 3: n = occupancy(k_tile2<TP_SH, 3>, kLdsTileThreads); break;
    case TP_SH * 8 + 4: n = occupancy(k_tile2<TP_SH, 4>, kLdsTileThreads); break;
    case TP_PSV * 8 + 2: n = occupancy(k_tile2<TP_PSV, 2>, kLdsTileThreads); break;
    case TP_PSV * 8 + 3: n = occupancy(k_tile2<TP_PSV, 3>, kLdsTileThreads); break;
    case TP_PSV * 8 + 4: n = occupancy(k_tile2<TP_PSV, 4>, kLdsTileThreads); break;
    case TP_SPIN * 8 + 2: n = occupancy(k_tile2<TP_SPIN, 2>, kLdsTileThreads); break;
    case TP_SPIN * 8 + 3: n = occupancy(k_tile2<TP_SPIN, 3>, kLdsTileThreads); break;
    default: break;
  }
  return std::max(n, 1);
}
int step2_blocks_per_cu(int nt, int group) {
  if (nt == kLdsTileThreads) return occupancy(k_tile2<TP_PSV, 2>, kLdsTileThreads);
  switch (nt * 8 + group) {
    case 64 * 8 + 1: return occupancy(k_step_psv2<64, 1, false>, 64);
    case 128 * 8 + 1: return occupancy(k_step_psv2<128, 1, false>, 128);
    case 128 * 8 + 2: return occupancy(k_step_psv2<128, 2, false>, 256);
    case 128 * 8 + 4: return occupancy(k_step_psv2<128, 4, false>, 512);
    case 256 * 8 + 1: return occupancy(k_step_psv2<256, 1, false>, 256);
    case 256 * 8 + 2: return occupancy(k_step_psv2<256, 2, false>, 512);
    case 256 * 8 + 4: return occupancy(k_step_psv2<256, 4, false>, 1024);
    case 320 * 8 + 1: return occupancy(k_step_psv2<320, 1, false>, 320);
    case 320 * 8 + 2: return occupancy(k_step_psv2<320, 2, false>, 640);
    case 512 * 8 + 1: return occupancy(k_step_psv2<512, 1, false>, 512);
    case 512 * 8 + 2: return occupancy(k_step_psv2<512, 2, false>, 1024);
  }
  return 0;
}

hipError_t launch_step_psv(const StepArgs& a, const Tiling& t, bool spin, hipStream_t s, bool aux) {
  if (spin && aux) return hipErrorInvalidValue;
  return launch_fwd(spin ? MK_SPIN : aux ? MK_PSV_AUX : MK_PSV, a, t, s);
}

hipError_t launch_multi(int kind, int nt_threads, const StepArgs* args, const int2* work, int nb, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  switch (nt_threads) {
    case 64: return launch_multi_nt<64>(kind, args, work, nb, s);
    case 128: return launch_multi_nt<128>(kind, args, work, nb, s);
    case 256: return launch_multi_nt<256>(kind, args, work, nb, s);
  }
  return hipErrorInvalidValue;
}

template <int NT, int S>
hipError_t launch_adj_ns(bool sh, const StepArgs& a, const Tiling& t, hipStream_t s) {
  if (!args_match(a, t, S)) return hipErrorInvalidValue;
  const int nblk = a.nshots / S * t.nstrips * t.nsegs;
  if (sh) hipLaunchKernelGGL((k_adj_sh<NT, S>), dim3(nblk), dim3(NT * S), 0, s, a);
  else hipLaunchKernelGGL((k_adj_psv<NT, S>), dim3(nblk), dim3(NT * S), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_adj(bool sh, const StepArgs& a, const Tiling& t, hipStream_t s) {
  if (t.group < 1 || a.nshots % t.group != 0) return hipErrorInvalidValue;
  switch (t.nt_threads * 8 + t.group) {
    case 64 * 8 + 1: return launch_adj_ns<64, 1>(sh, a, t, s);
    case 128 * 8 + 1: return launch_adj_ns<128, 1>(sh, a, t, s);
    case 128 * 8 + 2: return launch_adj_ns<128, 2>(sh, a, t, s);
    case 128 * 8 + 4: return launch_adj_ns<128, 4>(sh, a, t, s);
    case 256 * 8 + 1: return launch_adj_ns<256, 1>(sh, a, t, s);
    case 256 * 8 + 2: return launch_adj_ns<256, 2>(sh, a, t, s);
    case 256 * 8 + 4: return launch_adj_ns<256, 4>(sh, a, t, s);
  }
  return hipErrorInvalidValue;
}
hipError_t launch_adj_sh(const StepArgs& a, const Tiling& t, hipStream_t s) { return launch_adj(true, a, t, s); }
hipError_t launch_adj_psv(const StepArgs& a, const Tiling& t, hipStream_t s) { return launch_adj(false, a, t, s); }
hipError_t launch_adjsrc(const float* obs_syn, const float* obs_ref, const float* taper, float* adj, double* misfit,
                         int nshots, int ncomp, const int* comps, int nrec, int nt, float dt, hipStream_t s) {
  const int total = nshots * ncomp * nrec;
  if (total > 0) {
    // partials live after the per-shot misfits in the same buffer
    double* part = misfit + nshots;
    // one wavefront per trace, four per workgroup; one wavefront per shot
    hipLaunchKernelGGL(k_adjsrc, dim3((total + 3) / 4), dim3(256), 0, s, obs_syn, obs_ref, taper, adj, part,
                       nshots, ncomp, comps, nrec, nt, dt);
    hipLaunchKernelGGL(k_misfit_shots, dim3(nshots), dim3(64), 0, s, part, misfit, nshots, ncomp * nrec);
  } else if (nshots > 0) {
    // no receivers: no traces, a zero misfit per shot (and no adjoint source)
    hipLaunchKernelGGL(k_fill_f64, dim3((nshots + 63) / 64), dim3(64), 0, s, misfit, 0.0, int64_t(nshots));
  }
  return hipGetLastError();
}
hipError_t launch_reduce_shots(const float* kshot, float* kout, int nshots, int64_t n, int64_t shot_stride,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_reduce_shots, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, kshot, kout, nshots, n,
                     shot_stride);
  return hipGetLastError();
}

hipError_t launch_imaging_sh(const float* va, const float* vf, float* kmu, int nx, int nz, float h8x, float h8z,
                             float ndt, hipStream_t s) {
  const float r8x = 1.0f / h8x, r8z = 1.0f / h8z;
  const int fd = (fast_divisor(h8x) && fast_divisor(h8z)) ? FD_CONST : 0;
  hipLaunchKernelGGL(k_imaging_sh, dim3((nz + 255) / 256, nx), dim3(256), 0, s, va, vf, kmu, nx, nz, h8x, h8z, r8x,
                     r8z, ndt, fd);
  return hipGetLastError();
}
hipError_t launch_gauss_x(const float* w, int nw, const float* k, float* gtmp, int nx, int nz, hipStream_t s) {
  hipLaunchKernelGGL(k_gauss_x, dim3((nz + 255) / 256, nx), dim3(256), 0, s, w, nw, k, gtmp, nx, nz);
  return hipGetLastError();
}
hipError_t launch_gauss_z(const float* w, int nw, const float* gtmp, const float* sumx, const float* sumz, float* k,
                          int nx, int nz, hipStream_t s) {
  hipLaunchKernelGGL(k_gauss_z, dim3((nz + 255) / 256, nx), dim3(256), 0, s, w, nw, gtmp, sumx, sumz, k, nx, nz);
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
hipError_t launch_recip(const float* x, float* out, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_recip, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, x, out, n);
  return hipGetLastError();
}
hipError_t launch_fill(float* p, float v, int64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_nonfinite(const float* p, int64_t n, int tag, int* bad, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_nonfinite, dim3(unsigned(nb)), dim3(256), 0, s, p, n, tag, bad);
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
