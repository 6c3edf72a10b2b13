// Device-side argument blocks and launch wrappers shared by the C-ABI layer
// (capi.cpp) and the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sw {

// State-field slots of the ping-pong state buffers.  A context stores only
// the slots its physics needs.
enum StateSlot : int {
  S_VY = 0, S_SXY, S_SZY,                       // SH
  S_VX, S_VZ, S_SXX, S_SZZ, S_SXZ,              // P-SV
  S_VYC, S_SYXC, S_SYYC, S_SYZC,                // micropolar spin
  N_SLOTS
};

// Per-receiver displacement accumulators (u = sum v*dt) kept only where the
// traces need them.
enum UaccSlot : int {
  U_UY = 0, U_UX, U_UZ, U_UYC,
  U_UZ_M1, U_UZ_P1, U_UZ_P2,   // uz at i-1, i+1, i+2 (for d+uz/dx)
  U_UX_M1, U_UX_P1, U_UX_P2,   // ux at j-1, j+1, j+2 (for d+ux/dz)
  N_UACC
};

struct Tiling {
  int nt_threads;   // threads per shot = columns per strip incl. 6 halo
  int seglen;       // rows (x) marched by one workgroup
  int nstrips, nsegs;
  int group = 1;    // shots per workgroup of the forward steps (1, 2 or 4)
  int lds_steps = 0;   // multi-step launches: 0 marching tiles; 2..4 LDS tiles, steps per launch
  int msteps = 2;      // marching multi-step tiles: steps per launch (2, or 3 for P-SV alone)
};

// Rows of the model buffer [MF_N][npt].
// MF_IRHO / MF_IJ hold the correctly rounded reciprocals 1/rho, 1/J.
enum ModelField : int { MF_RHO = 0, MF_LAM, MF_MU, MF_NU, MF_J, MF_LAM_C, MF_MU_C, MF_NU_C, MF_IRHO, MF_IJ, MF_N };

// Exact divisions by rho (FD_RHO) and J (FD_J) may use the
// reciprocal-and-correction form when every such divisor lies in
// [2^-20, 2^20]; divisions by 8h (FD_CONST) take the two-operation form
// fma(x, zh, x * zl) with 1/(8h) split as zh + zl, once the host has shown
// it exact for that divisor (bmr_exact).  Otherwise the kernels divide with
// the IEEE sequence.  FD_RCP: the forward step kernels multiply by the
// correctly rounded reciprocal instead (x * RN(1/y), within 1.5 units of x / y).
enum FastDiv : int { FD_CONST = 1, FD_RHO = 2, FD_J = 4, FD_RCP = 8 };
inline bool fast_divisor(float y) { return y >= 0x1p-20f && y <= 0x1p20f; }

struct StepArgs {
  // ping-pong state buffers, each laid out [shot][slot][npt].  A forward step
  // `it` reads st[it & 1] and writes st[(it+1) & 1]; an adjoint step `it`
  // reads st[(it+1) & 1] and writes st[it & 1].
  float* st[2];
  int slot_off[N_SLOTS];      // float offset of each slot inside one shot (-1: not stored)
  int shot_floats;            // floats per shot
  const float* model;         // [MF_N][npt], shared by all shots
  int npt;
  const float* bx; const float* bzb; const float* bzt;   // 1-D taper factors
  int nx, nz, nshots;
  int nstrips, nsegs, seglen;
  float dt, h8x, h8z;         // h8 = 8*h (fd2d.wgsl:156)
  float r8x, r8z;             // correctly rounded 1/h8
  float z8x, z8z, l8x, l8z;   // 1/h8 split as zh + zl for FD_CONST (zl >= 0)
  unsigned c_lo;              // FD_CONST: numerators x of magnitude below T take the
                              // IEEE division; c_lo = (bits(T) << 1) - 1
  int fast_div;               // FastDiv bits
  // body-force injections: CSR over (shot, tile), entries (i, j, q) sorted by i
  const int* inj_off; const int4* inj;
  const float* val_x; const float* val_y; const float* val_z;   // [q][nt_val]
  int nt_val;
  // receivers: CSR over tiles (shared by shots), entries (i, j, r) sorted by i
  const int* rec_off; const int4* rec;
  float* uacc;                // [nshots][nrec][N_UACC]
  float* obs;                 // [nshots][5][nrec][nt]
  int nrec, nt;
  int record;                 // 1: accumulate u and write traces
  int* clock;                 // [0] = current step it, [1] = finished-block counter
  int advance;                // forward: last block sets clock[0] = it+1; adjoint: it-1
  int it_lo, it_hi;           // the launch is a no-op unless it_lo <= it < it_hi
  int nblocks;                // workgroups of this problem in the launch
  // gradient support: the forward velocities of the recomputed segment
  // (forward recompute writes v^{n+1} after step n, the adjoint of step n
  // reads it) laid out [n % aux_period][shot][naux][npt], with aux_period
  // the checkpoint interval, and the kernel accumulators (rho, mu, lambda)
  float* aux;
  int aux_period;
  float* kgrad;               // kernel accumulators [shot / S][3][npt]
  int stream_out;             // 1: state and aux stores non-temporal (working set far beyond the caches)
  int tile_major;             // 1: consecutive workgroups take the same tile of consecutive shot groups
};

// velocity fields kept per recomputed forward step
enum AuxSh : int { A_VY = 0, NAUX_SH };
enum AuxPsv : int { A_VX = 0, A_VZ, NAUX_PSV };
enum KField : int { K_RHO = 0, K_MU, K_LAM, NK };

// launches (kernels.hip)
hipError_t launch_step_sh(const StepArgs& a, const Tiling& t, hipStream_t s, bool aux = false);
hipError_t launch_step_psv(const StepArgs& a, const Tiling& t, bool spin, hipStream_t s, bool aux = false);
// two P-SV steps per launch: reads st[0], writes st[1] (a buffer distinct
// from st[0]); marching tiles of nt_threads-12 columns (no spin); aux: also
// the velocities of both steps (forward recompute of the gradient).  With
// nt_threads == kLdsTileThreads: t.lds_steps (K = 2..4) steps on LDS tiles
// of lds_tile_rows(K) x lds_tile_cols(K) cells (small grids; spin too, K = 2
// or 3; no aux)
constexpr int kLdsTileThreads = 1024;
// rec: a marching launch records the receivers itself (else the caller runs
// launch_record2 after it when traces are recorded)
hipError_t launch_step_psv2(const StepArgs& a, const Tiling& t, hipStream_t s, bool aux = false, bool spin = false,
                            bool rec = false);
int lds_tile_rows(int k);
int lds_tile_cols(int k);
// the same for SH
hipError_t launch_step_sh2(const StepArgs& a, const Tiling& t, hipStream_t s, bool aux = false, bool rec = false);
// traces of a two-step marching launch (SH or P-SV alone), after it: nent
// receiver entries in a.rec (fourth field: the owning two-step tile), every shot
hipError_t launch_record2(bool sh, const StepArgs& a, int nent, hipStream_t s);
hipError_t launch_adj_sh(const StepArgs& a, const Tiling& t, hipStream_t s);
hipError_t launch_adj_psv(const StepArgs& a, const Tiling& t, hipStream_t s);

// Several independent problems advanced by one launch: args[p] in device
// memory, work[b] = (problem, workgroup of that problem) for every workgroup
// b of the launch; all problems share the tile width nt_threads.
enum MultiKind : int { MK_SH = 0, MK_PSV, MK_SPIN, MK_SH_AUX, MK_PSV_AUX, MK_ADJ_SH, MK_ADJ_PSV };
hipError_t launch_multi(int kind, int nt_threads, const StepArgs* args, const int2* work, int nblocks,
                        hipStream_t s);
// adjoint sources dt * sum_{k>=n} 2 w_k^2 (d^k - obs^k) and per-shot misfit
hipError_t launch_adjsrc(const float* obs_syn, const float* obs_ref, const float* taper, float* adj, double* misfit,
                         int nshots, int ncomp, const int* comps, int nrec, int nt, float dt, hipStream_t s);
hipError_t launch_reduce_shots(const float* kshot, float* kout, int nshots, int64_t n, int64_t shot_stride,
                               hipStream_t s);
bool tiling_supported(int nt_threads);
bool group_supported(int nt_threads, int group);
// two-step P-SV launches: tile widths 64..512, 2 or 4 shots per workgroup up
// to 1024 threads; resident workgroups per CU of a shape
bool tiling2_supported(int nt_threads);
bool group2_supported(int nt_threads, int group);
int step2_blocks_per_cu(int nt_threads, int group);
bool tiling3_supported(int nt_threads);
int step3_blocks_per_cu(int nt_threads);
// resident LDS-tile workgroups per CU (K steps per launch)
int lds_tile_blocks_per_cu(bool sh, bool spin, int k);
// resident workgroups per CU of the step kernel of a MultiKind at a tile width
// and shots per workgroup (adjoint kinds: group 1 only)
int step_blocks_per_cu(int kind, int nt_threads, int group);

hipError_t launch_imaging_sh(const float* vadj, const float* vfw, float* kmu, int nx, int nz,
                             float h8x, float h8z, float ndt, hipStream_t s);
// separable Gaussian sums; w[d] == 0 for d >= nw
hipError_t launch_gauss_x(const float* w, int nw, const float* k, float* gtmp, int nx, int nz, hipStream_t s);
hipError_t launch_gauss_z(const float* w, int nw, const float* gtmp, const float* sumx, const float* sumz, float* k,
                          int nx, int nz, hipStream_t s);
hipError_t launch_misfit(const float* syn, const float* ref, const float* taper, float* adstf,
                         double* partial, int nrec, int nt, hipStream_t s, int* nblocks_out);
hipError_t launch_reduce_sum(const double* partial, int n, double* out, hipStream_t s);
hipError_t launch_vps2lm(const float* rho, float* lam_vp, float* mu_vs, int64_t npt, hipStream_t s);
hipError_t launch_fill(float* p, float v, int64_t n, hipStream_t s);
hipError_t launch_recip(const float* x, float* out, int64_t n, hipStream_t s);
hipError_t launch_copy(const float* src, float* dst, int64_t n, hipStream_t s);
// *bad = min(*bad, tag) if any of p[0..n) is NaN or infinite
hipError_t launch_nonfinite(const float* p, int64_t n, int tag, int* bad, hipStream_t s);

}  // namespace sw
