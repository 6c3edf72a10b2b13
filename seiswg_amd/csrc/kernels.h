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
  int nt_threads;   // threads per workgroup = columns per strip incl. 6 halo
  int seglen;       // rows (x) marched by one workgroup
  int nstrips, nsegs;
};

struct StepArgs {
  const float* in[N_SLOTS];   // shot-0 base of each state slot (read)
  float* out[N_SLOTS];        // shot-0 base of each state slot (write)
  int64_t shot_stride;        // floats between consecutive shots of one slot
  // model (shared by all shots)
  const float* rho; const float* lam; const float* mu;
  const float* nu; const float* jj; const float* muc; const float* nuc;
  const float* bx; const float* bzb; const float* bzt;   // 1-D taper factors
  int nx, nz, nshots;
  int nstrips, nsegs, seglen;
  float dt, h8x, h8z;         // h8 = 8*h (fd2d.wgsl:156)
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
  int advance;                // 1: the last block advances clock[0]
};

// launches (kernels.hip)
hipError_t launch_step_sh(const StepArgs& a, const Tiling& t, hipStream_t s);
hipError_t launch_step_psv(const StepArgs& a, const Tiling& t, bool spin, hipStream_t s);
bool tiling_supported(int nt_threads);

hipError_t launch_imaging_sh(const float* vadj, const float* vfw, float* kmu, int nx, int nz,
                             float h8x, float h8z, float ndt, hipStream_t s);
hipError_t launch_gsum(const float* sumx, const float* sumz, float* gsum, int nx, int nz, hipStream_t s);
hipError_t launch_gauss_x(const float* w, const float* k, float* gtmp, int nx, int nz, hipStream_t s);
hipError_t launch_gauss_z(const float* w, const float* gtmp, const float* gsum, float* k, int nx, int nz,
                          hipStream_t s);
hipError_t launch_misfit(const float* syn, const float* ref, const float* taper, float* adstf,
                         double* partial, int nrec, int nt, hipStream_t s, int* nblocks_out);
hipError_t launch_reduce_sum(const double* partial, int n, double* out, hipStream_t s);
hipError_t launch_vps2lm(const float* rho, float* lam_vp, float* mu_vs, int64_t npt, hipStream_t s);
hipError_t launch_fill(float* p, float v, int64_t n, hipStream_t s);
hipError_t launch_copy(const float* src, float* dst, int64_t n, hipStream_t s);

}  // namespace sw
