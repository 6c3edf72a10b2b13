// CPU oracle for the seiswg staggered-grid time loop.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (seiswg_amd/) links,
// loads or calls this file; only tests/, the smoke check and the CPU-baseline
// leg of the benchmark use it, and only as the checker.
//
// It is a literal, unfused restatement of the reference algorithm:
//   * one loop per WGSL compute entry point of shaders/fd2d.wgsl, with the
//     same arithmetic order (compile with -ffp-contract=off so that no FMA is
//     formed), the same edge conditions and the same 44-slice field layout
//     (fd2d.wgsl:53-105, solver.rs:34-85);
//   * the host sequencing of src/solver.rs: dispatch order of one forward step
//     (solver.rs:1000-1030), the forward driver (941-1119), the adjoint driver
//     (1298-1529), the Ricker table (350-356, 721-739), grid inference
//     (640-658), grid-index rounding (709-719, 745-751) and the waveform misfit
//     (432-467).
// Duplicate source/receiver cells are accumulated in index order (the
// reference races there, fd2d.wgsl:252-260; in index order it equals the
// reference whenever positions are unique).
//
// It is templated on the real type: T=float is the FP32 restatement of the
// reference scheme, T=double is the FP64 oracle of the same scheme.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#include <memory>
#endif

namespace {

// field slice ids (fd2d.wgsl:53-105)
enum : uint32_t {
  F_VY = 0, F_UY, F_SXY, F_SZY, F_DSY, F_DVYDX, F_DVYDZ,
  F_VX = 7, F_VZ, F_UX, F_UZ, F_SXX, F_SZZ, F_SXZ, F_DSX, F_DSZ,
  F_DVXDX = 16, F_DVXDZ, F_DVZDX, F_DVZDZ,
  F_VY_C = 20, F_UY_C, F_SYX_C, F_SYY_C, F_SYZ_C, F_DSY_C, F_DVYDX_C,
  F_DVYDZ_C, F_DUZDX, F_DUXDZ,
  F_LAM = 30, F_MU, F_NU, F_J, F_LAM_C, F_MU_C, F_NU_C, F_RHO, F_BOUND,
  F_K_LAM = 39, F_K_MU, F_K_RHO, F_GSUM, F_GTMP,
  N_FIELDS = 44, N_DYN_FIELDS = 30
};
enum : uint32_t { OBS_X = 0, OBS_Y, OBS_Z, OBS_YI, OBS_YC };
enum : uint32_t { STF_X = 0, STF_Y, STF_Z };

}  // namespace

extern "C" {
// Plain-old-data configuration shared with the Python wrapper (oracle.py).
typedef struct {
  int32_t nx, nz, nt;
  float dx, dz, dt;
  int32_t sh, psv, spin;
  int32_t abs_left, abs_right, abs_bottom, abs_top, abs_width;
  float abs_alpha;
  int32_t nsrc, nrec;
} or_cfg;
}

namespace {

template <typename T>
struct Oracle {
  // Params uniform (fd2d.wgsl:18-49)
  uint32_t nx, nz, nt, npt, nsrc, nrec;
  T dx, dz, dt;
  bool sh, psv, spin;
  bool abs_left, abs_right, abs_bottom, abs_top;
  uint32_t abs_width;
  T abs_alpha;
  uint32_t it = 0;
  int32_t isrc = -1;

  std::vector<T> fields;      // [44][npt]
  std::vector<int32_t> idata; // [nsrc | nrec]
  std::vector<T> obs;         // [5][nrec][nt]
  std::vector<T> stf;         // [3][nsrc][nt]

  explicit Oracle(const or_cfg& c)
      : nx(c.nx), nz(c.nz), nt(c.nt), npt(uint32_t(c.nx) * uint32_t(c.nz)),
        nsrc(c.nsrc), nrec(c.nrec), dx(T(c.dx)), dz(T(c.dz)), dt(T(c.dt)),
        sh(c.sh), psv(c.psv), spin(c.spin), abs_left(c.abs_left),
        abs_right(c.abs_right), abs_bottom(c.abs_bottom), abs_top(c.abs_top),
        abs_width(uint32_t(c.abs_width)), abs_alpha(T(c.abs_alpha)) {
    fields.assign(size_t(N_FIELDS) * npt, T(0));
    idata.assign(size_t(nsrc) + nrec, 0);
    obs.assign(size_t(5) * nrec * nt, T(0));
    stf.assign(size_t(3) * nsrc * nt, T(0));
  }

  // fd2d.wgsl:132-137
  inline T gf(uint32_t id, uint32_t k) const { return fields[size_t(id) * npt + k]; }
  inline void sf(uint32_t id, uint32_t k, T v) { fields[size_t(id) * npt + k] = v; }
  inline void af(uint32_t id, uint32_t k, T v) { fields[size_t(id) * npt + k] += v; }
  inline uint32_t src_id(uint32_t s) const { return uint32_t(idata[s]); }
  inline uint32_t rec_id(uint32_t r) const { return uint32_t(idata[nsrc + r]); }
  // fd2d.wgsl:140-145
  inline void ij_from(uint32_t k, uint32_t& i, uint32_t& j) const {
    j = k % nz;
    i = (k - j) / nz;
  }

  // fd2d.wgsl:151-180 (coefficient 1/8 on the 3-cell difference, as written)
  inline T diff_x(uint32_t fid, uint32_t i, uint32_t k) const {
    if (i < 2u || i >= nx - 2u) return T(0);
    return (T(9) * (gf(fid, k) - gf(fid, k - nz)) - (gf(fid, k + nz) - gf(fid, k - 2u * nz))) /
           (T(8) * dx);
  }
  inline T diff_z(uint32_t fid, uint32_t j, uint32_t k) const {
    if (j < 2u || j >= nz - 2u) return T(0);
    return (T(9) * (gf(fid, k) - gf(fid, k - 1u)) - (gf(fid, k + 1u) - gf(fid, k - 2u))) /
           (T(8) * dz);
  }
  inline T diff_x1(uint32_t fid, uint32_t i, uint32_t k) const {
    if (i < 1u || i >= nx - 2u) return T(0);
    return (T(9) * (gf(fid, k + nz) - gf(fid, k)) - (gf(fid, k + 2u * nz) - gf(fid, k - nz))) /
           (T(8) * dx);
  }
  inline T diff_z1(uint32_t fid, uint32_t j, uint32_t k) const {
    if (j < 1u || j >= nz - 2u) return T(0);
    return (T(9) * (gf(fid, k + 1u) - gf(fid, k)) - (gf(fid, k + 2u) - gf(fid, k - 1u))) /
           (T(8) * dz);
  }

#define GRID_LOOP _Pragma("omp parallel for schedule(static)") for (int64_t kk = 0; kk < int64_t(npt); ++kk)

  // fd2d.wgsl:188-201
  void vps2lm() {
    GRID_LOOP {
      uint32_t k = uint32_t(kk);
      T vp = gf(F_LAM, k), vs = gf(F_MU, k), rho = gf(F_RHO, k);
      sf(F_MU, k, rho * vs * vs);
      if (vp > vs) sf(F_LAM, k, rho * (vp * vp - T(2) * vs * vs));
      else sf(F_LAM, k, T(0));
    }
  }
  // fd2d.wgsl:205-235 (u32 edge arithmetic kept)
  void set_bound() {
    GRID_LOOP {
      uint32_t k = uint32_t(kk), i, j;
      ij_from(k, i, j);
      uint32_t w = abs_width;
      T alpha = abs_alpha, b = T(1);
      if (abs_left && (i + 1u) < w) { T aw = alpha * T(w - i - 1u); b *= std::exp(-aw * aw); }
      if (abs_right && i > (nx - w)) { T aw = alpha * T(w + i - nx); b *= std::exp(-aw * aw); }
      if (abs_bottom && j > (nz - w)) { T aw = alpha * T(w + j - nz); b *= std::exp(-aw * aw); }
      if (abs_top && (j + 1u) < w) { T aw = alpha * T(w - j - 1u); b *= std::exp(-aw * aw); }
      sf(F_BOUND, k, b);
    }
  }

  // ---- SH (fd2d.wgsl:242-301)
  void div_sy() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DSY, k, diff_x(F_SXY, i, k) + diff_z(F_SZY, j, k)); }
  }
  void stf_dsy() {
    for (uint32_t ib = 0; ib < nsrc; ++ib) {
      if (isrc >= 0 && int32_t(ib) != isrc) continue;
      uint32_t ks = ib * nt + it, km = src_id(ib);
      af(F_DSY, km, stf[size_t(STF_Y) * nsrc * nt + ks]);
    }
  }
  void add_vy() {
    GRID_LOOP { uint32_t k = uint32_t(kk);
      T vy_new = gf(F_BOUND, k) * (gf(F_VY, k) + dt * gf(F_DSY, k) / gf(F_RHO, k));
      sf(F_VY, k, vy_new);
      af(F_UY, k, vy_new * dt); }
  }
  void div_vy() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DVYDX, k, diff_x1(F_VY, i, k));
      sf(F_DVYDZ, k, diff_z1(F_VY, j, k)); }
  }
  void add_sy_sh() {
    GRID_LOOP { uint32_t k = uint32_t(kk);
      T mu_dt = gf(F_MU, k) * dt;
      af(F_SXY, k, mu_dt * gf(F_DVYDX, k));
      af(F_SZY, k, mu_dt * gf(F_DVYDZ, k)); }
  }
  void save_obs_y_sh() {
    for (uint32_t ir = 0; ir < nrec; ++ir)
      obs[size_t(OBS_Y) * nrec * nt + ir * nt + it] = gf(F_UY, rec_id(ir));
  }

  // ---- PSV (fd2d.wgsl:308-399)
  void div_sxz() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DSX, k, diff_x(F_SXX, i, k) + diff_z(F_SXZ, j, k));
      sf(F_DSZ, k, diff_x(F_SXZ, i, k) + diff_z(F_SZZ, j, k)); }
  }
  void div_sxyz_c() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      T syy = gf(F_SYY_C, k);
      af(F_DSX, k, diff_z(F_SYY_C, j, k));
      af(F_DSZ, k, -diff_x(F_SYY_C, i, k));
      af(F_DSY_C, k, T(2) * syy); }
  }
  void div_sy_c() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DSY_C, k, diff_x(F_SYX_C, i, k) + diff_z(F_SYZ_C, j, k)); }
  }
  void stf_dsxz() {
    size_t nst = size_t(nsrc) * nt;
    for (uint32_t ib = 0; ib < nsrc; ++ib) {
      if (isrc >= 0 && int32_t(ib) != isrc) continue;
      uint32_t ks = ib * nt + it, km = src_id(ib);
      af(F_DSX, km, stf[STF_X * nst + ks]);
      af(F_DSZ, km, stf[STF_Z * nst + ks]);
    }
  }
  void add_vxz() {
    GRID_LOOP { uint32_t k = uint32_t(kk);
      T rho = gf(F_RHO, k), bound = gf(F_BOUND, k);
      T vx = bound * (gf(F_VX, k) + dt * gf(F_DSX, k) / rho);
      T vz = bound * (gf(F_VZ, k) + dt * gf(F_DSZ, k) / rho);
      sf(F_VX, k, vx); sf(F_VZ, k, vz);
      af(F_UX, k, vx * dt); af(F_UZ, k, vz * dt); }
  }
  void div_vxz() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DVXDX, k, diff_x1(F_VX, i, k));
      sf(F_DVXDZ, k, diff_z1(F_VX, j, k));
      sf(F_DVZDX, k, diff_x1(F_VZ, i, k));
      sf(F_DVZDZ, k, diff_z1(F_VZ, j, k)); }
  }
  void add_sxz() {
    GRID_LOOP { uint32_t k = uint32_t(kk);
      T lam = gf(F_LAM, k), mu = gf(F_MU, k);
      T dvxdx = gf(F_DVXDX, k), dvxdz = gf(F_DVXDZ, k), dvzdx = gf(F_DVZDX, k), dvzdz = gf(F_DVZDZ, k);
      af(F_SXX, k, dt * ((lam + T(2) * mu) * dvxdx + lam * dvzdz));
      af(F_SZZ, k, dt * ((lam + T(2) * mu) * dvzdz + lam * dvxdx));
      af(F_SXZ, k, dt * mu * (dvxdz + dvzdx)); }
  }
  // ---- spin (fd2d.wgsl:406-492)
  void add_vy_c() {
    GRID_LOOP { uint32_t k = uint32_t(kk);
      T bound = gf(F_BOUND, k);
      T vy_c = bound * (gf(F_VY_C, k) + dt * gf(F_DSY_C, k) / gf(F_J, k));
      sf(F_VY_C, k, vy_c);
      af(F_UY_C, k, vy_c * dt); }
  }
  void div_vy_c() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DVYDX_C, k, diff_x1(F_VY_C, i, k));
      sf(F_DVYDZ_C, k, diff_z1(F_VY_C, j, k)); }
  }
  void add_sy_c() {
    GRID_LOOP { uint32_t k = uint32_t(kk);
      T nu = gf(F_NU, k), mu_c = gf(F_MU_C, k), nu_c = gf(F_NU_C, k), vy_c = gf(F_VY_C, k);
      T dvydx_c = gf(F_DVYDX_C, k), dvydz_c = gf(F_DVYDZ_C, k);
      T dvxdz = gf(F_DVXDZ, k), dvzdx = gf(F_DVZDX, k);
      af(F_SYY_C, k, T(2) * dt * nu * (-vy_c - T(0.5) * (dvzdx - dvxdz)));
      af(F_SYX_C, k, dt * (mu_c + nu_c) * dvydx_c);
      af(F_SYZ_C, k, dt * (mu_c + nu_c) * dvydz_c); }
  }
  void compute_du_grad() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DUZDX, k, diff_x1(F_UZ, i, k));
      sf(F_DUXDZ, k, diff_z1(F_UX, j, k)); }
  }
  void save_obs_x() {
    for (uint32_t ir = 0; ir < nrec; ++ir)
      obs[size_t(OBS_X) * nrec * nt + ir * nt + it] = gf(F_UX, rec_id(ir));
  }
  void save_obs_z() {
    for (uint32_t ir = 0; ir < nrec; ++ir)
      obs[size_t(OBS_Z) * nrec * nt + ir * nt + it] = gf(F_UZ, rec_id(ir));
  }
  void save_obs_ry() {
    size_t nobs = size_t(nrec) * nt;
    for (uint32_t ir = 0; ir < nrec; ++ir) {
      uint32_t km = rec_id(ir);
      size_t kt = size_t(ir) * nt + it;
      T uy_c = gf(F_UY_C, km);
      T curl = -(gf(F_DUZDX, km) - gf(F_DUXDZ, km));
      T iso = uy_c - T(0.5) * curl;
      obs[OBS_Y * nobs + kt] = uy_c;
      obs[OBS_YI * nobs + kt] = iso;
      obs[OBS_YC * nobs + kt] = curl;
    }
  }
  // ---- adjoint helpers (fd2d.wgsl:501-609)
  void interaction_muy() {
    GRID_LOOP { uint32_t k = uint32_t(kk);
      af(F_K_MU, k, (gf(F_DVYDX, k) * gf(F_DVYDX_C, k) + gf(F_DVYDZ, k) * gf(F_DVYDZ_C, k)) * abs_alpha); }
  }
  void adj_dsy() {
    for (uint32_t ir = 0; ir < nrec; ++ir) {
      uint32_t km = rec_id(ir), kr = ir * nt + it;
      af(F_DSY, km, obs[size_t(OBS_Y) * nrec * nt + kr]);
    }
  }
  void div_uy() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DVYDX, k, diff_x1(F_VY, i, k));
      sf(F_DVYDZ, k, diff_z1(F_VY, j, k)); }
  }
  void div_fw() {
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      sf(F_DVYDX_C, k, diff_x1(F_DSY, i, k));
      sf(F_DVYDZ_C, k, diff_z1(F_DSY, j, k)); }
  }
  void init_gsum() {
    T sigma = abs_alpha;
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      T sumx = T(0);
      for (uint32_t n = 0; n < nx; ++n) { T d = T(int32_t(i) - int32_t(n)); sumx += std::exp(-d * d / (T(2) * sigma * sigma)); }
      T sumz = T(0);
      for (uint32_t n = 0; n < nz; ++n) { T d = T(int32_t(j) - int32_t(n)); sumz += std::exp(-d * d / (T(2) * sigma * sigma)); }
      sf(F_GSUM, k, sumx * sumz); }
  }
  void gaussian_x() {
    T sigma = abs_alpha;
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      T sumx = T(0);
      for (uint32_t n = 0; n < nx; ++n) {
        T d = T(int32_t(i) - int32_t(n));
        T w = std::exp(-d * d / (T(2) * sigma * sigma));
        sumx += w * gf(F_K_MU, n * nz + j);
      }
      sf(F_GTMP, k, sumx); }
  }
  void gaussian_z() {
    T sigma = abs_alpha;
    GRID_LOOP { uint32_t k = uint32_t(kk), i, j; ij_from(k, i, j);
      T sumz = T(0);
      for (uint32_t n = 0; n < nz; ++n) {
        T d = T(int32_t(j) - int32_t(n));
        T w = std::exp(-d * d / (T(2) * sigma * sigma));
        sumz += w * gf(F_GTMP, i * nz + n);
      }
      T gsum = gf(F_GSUM, k);
      sf(F_K_MU, k, gsum > T(1e-30) ? sumz / gsum : T(0)); }
  }
#undef GRID_LOOP

  // ---- host sequencing (solver.rs)
  void init_model() {           // solver.rs:898-915
    if (!spin) vps2lm();
    set_bound();
  }
  void clear_dyn() {            // solver.rs:978, 1126-1132
    std::fill(fields.begin(), fields.begin() + size_t(N_DYN_FIELDS) * npt, T(0));
  }
  void clear_obs() { std::fill(obs.begin(), obs.end(), T(0)); }

  // one forward time step, dispatch order of solver.rs:999-1030
  void step(uint32_t it_, int32_t isrc_) {
    it = it_;
    isrc = isrc_;
    if (sh) { div_sy(); stf_dsy(); add_vy(); div_vy(); add_sy_sh(); save_obs_y_sh(); }
    if (psv) {
      div_sxz();
      if (spin) { div_sy_c(); div_sxyz_c(); }
      stf_dsxz(); add_vxz(); div_vxz(); add_sxz();
      if (spin) { add_vy_c(); div_vy_c(); add_sy_c(); }
      save_obs_x(); save_obs_z();
      if (spin) { compute_du_grad(); save_obs_ry(); }
    }
  }
  // SH adjoint step (solver.rs:1458-1464)
  void adjoint_step_sh(uint32_t it_) {
    it = it_;
    isrc = -1;
    if (sh) { div_sy(); adj_dsy(); add_vy(); div_vy(); add_sy_sh(); }
  }
  // smooth_gradient (solver.rs:1249-1274)
  void smooth(T sigma) {
    T saved = abs_alpha;
    abs_alpha = sigma;
    init_gsum(); gaussian_x(); gaussian_z();
    abs_alpha = saved;
  }
};

// compute_waveform_misfit (solver.rs:432-467); the taper is evaluated in T.
template <typename T>
double waveform_misfit(const T* syn, const T* obs, uint32_t nrec, uint32_t nt, T dt, T* adstf) {
  const T PI_T = T(3.14159265358979323846);
  double misfit = 0.0;
  T t_end = T(nt - 1) * dt;
  T taper_width = t_end / T(10);
  T t_min = taper_width, t_max = t_end - taper_width;
  for (uint32_t ir = 0; ir < nrec; ++ir)
    for (uint32_t it = 0; it < nt; ++it) {
      size_t kt = size_t(ir) * nt + it, akt = size_t(ir) * nt + (nt - 1 - it);
      T t = T(it) * dt;
      T taper;
      if (t <= t_min) taper = T(0.5) + T(0.5) * std::cos(PI_T * (t_min - t) / taper_width);
      else if (t >= t_max) taper = T(0.5) + T(0.5) * std::cos(PI_T * (t_max - t) / taper_width);
      else taper = T(1);
      T diff = (syn[kt] - obs[kt]) * taper;
      misfit += double(diff) * double(diff);
      adstf[akt] = diff * taper * T(2);
    }
  return std::sqrt(misfit);
}

// run_adjoint (solver.rs:1298-1529), SH only as in the reference.
// ref_traces: [nruns][nrec][nt]; runs use isrc_list[r].
template <typename T>
int adjoint_plain(Oracle<T>& o, int32_t nruns, const int32_t* isrc_list, const T* ref_traces,
                  uint32_t adj_interval, T smooth_sigma, double* misfits) {
  uint32_t nt = o.nt, nrec = o.nrec, npt = o.npt;
  if (adj_interval == 0) return -2;
  uint32_t nsa = nt / adj_interval;
  if (nsa == 0) return -1;
  std::fill(o.fields.begin() + size_t(F_K_MU) * npt, o.fields.begin() + size_t(F_K_MU + 1) * npt, T(0));
  T boundary_alpha = o.abs_alpha;
  T ndt = T(adj_interval) * o.dt;
  std::vector<T> uy_fwd(size_t(nsa) * npt);
  std::vector<T> adstf(size_t(nrec) * nt);
  for (int32_t r = 0; r < nruns; ++r) {
    int32_t isrc = isrc_list[r];
    o.clear_dyn();
    o.clear_obs();
    std::fill(uy_fwd.begin(), uy_fwd.end(), T(0));
    for (uint32_t it = 0; it < nt; ++it) {
      o.abs_alpha = boundary_alpha;
      o.it = it; o.isrc = isrc;
      if (o.sh) { o.div_sy(); o.stf_dsy(); o.add_vy(); o.div_vy(); o.add_sy_sh(); o.save_obs_y_sh(); }
      if ((it + 1) % adj_interval == 0) {
        size_t isa = nsa - (it + 1) / adj_interval;
        std::copy(o.fields.begin() + size_t(F_VY) * npt, o.fields.begin() + size_t(F_VY + 1) * npt,
                  uy_fwd.begin() + isa * npt);
      }
    }
    double misfit = 0.0;
    if (o.sh) {
      const T* syn = o.obs.data() + size_t(OBS_Y) * nrec * nt;
      misfit = waveform_misfit<T>(syn, ref_traces + size_t(r) * nrec * nt, nrec, nt, o.dt, adstf.data());
    } else {
      std::fill(adstf.begin(), adstf.end(), T(0));
    }
    if (misfits) misfits[r] = misfit;
    std::copy(adstf.begin(), adstf.end(), o.obs.begin() + size_t(OBS_Y) * nrec * nt);
    o.clear_dyn();
    for (uint32_t it = 0; it < nt; ++it) {
      o.abs_alpha = boundary_alpha;
      o.adjoint_step_sh(it);
      if (it % adj_interval == 0) {
        uint32_t isae = it / adj_interval;
        if (isae < nsa) {
          std::copy(uy_fwd.begin() + size_t(isae) * npt, uy_fwd.begin() + size_t(isae + 1) * npt,
                    o.fields.begin() + size_t(F_DSY) * npt);
          o.abs_alpha = ndt;
          o.div_uy(); o.div_fw(); o.interaction_muy();
        }
      }
    }
  }
  o.abs_alpha = boundary_alpha;
  o.smooth(smooth_sigma);
  return 0;
}

// ---------------------------------------------------------------------------
// Exact discrete adjoint of the forward scheme (no counterpart in the
// reference, whose only gradient is the SH cross-correlation imaging kernel of
// run_adjoint).  Objective, per run r and traced component c:
//   J = sum_{r,c,rec,n} ( w_n * (d_{r,c,rec}^n - obs_{r,c,rec}^n) )^2
// with w_n the cosine time taper of compute_waveform_misfit (solver.rs:441-455)
// and d^n = u^{n+1} at the receiver (u = sum v*dt, recorded after step n).
// SH traces uy; P-SV traces ux and uz.  Gradients are with respect to the
// stored rho, lambda, mu (after vps2lm).  All forward states are kept (tiny
// grids only); the sweep runs the transposed step operators in reverse order.
template <typename T>
struct Adjoint {
  Oracle<T>& o;
  explicit Adjoint(Oracle<T>& x) : o(x) {}
  uint32_t nx() const { return o.nx; }
  uint32_t nz() const { return o.nz; }

  // Transposes of the D operators, scatter form over the active output cells.
  // An x-operator only moves values along a column and a z-operator along a
  // row, so the threads take blocks of columns (x) or rows (z): every output
  // cell still receives its contributions in increasing i (or j) order, the
  // order of a serial loop over i then j.
  // [j0, j1): the columns of this thread (x-operators)
  void my_columns(uint32_t& j0, uint32_t& j1) const {
    const uint32_t nth = uint32_t(omp_get_num_threads()), t = uint32_t(omp_get_thread_num());
    j0 = uint32_t(uint64_t(o.nz) * t / nth);
    j1 = uint32_t(uint64_t(o.nz) * (t + 1) / nth);
  }
  void dxm_t(const std::vector<T>& w, std::vector<T>& out) const {   // (Dx-)^T
    const uint32_t nz = o.nz, nx = o.nx;
#pragma omp parallel
    {
      uint32_t j0, j1;
      my_columns(j0, j1);
      for (uint32_t i = 2; i + 2 < nx; ++i)
        for (uint32_t j = j0; j < j1; ++j) {
          const uint32_t k = i * nz + j;
          const T c = w[k] / (T(8) * o.dx);
          out[k] += T(9) * c; out[k - nz] -= T(9) * c; out[k + nz] -= c; out[k - 2 * nz] += c;
        }
    }
  }
  void dzm_t(const std::vector<T>& w, std::vector<T>& out) const {   // (Dz-)^T
    const uint32_t nz = o.nz;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < int64_t(o.nx); ++ii)
      for (uint32_t j = 2; j + 2 < nz; ++j) {
        const uint32_t k = uint32_t(ii) * nz + j;
        const T c = w[k] / (T(8) * o.dz);
        out[k] += T(9) * c; out[k - 1] -= T(9) * c; out[k + 1] -= c; out[k - 2] += c;
      }
  }
  void dxp_t(const std::vector<T>& w, std::vector<T>& out) const {   // (Dx+)^T
    const uint32_t nz = o.nz, nx = o.nx;
#pragma omp parallel
    {
      uint32_t j0, j1;
      my_columns(j0, j1);
      for (uint32_t i = 1; i + 2 < nx; ++i)
        for (uint32_t j = j0; j < j1; ++j) {
          const uint32_t k = i * nz + j;
          const T c = w[k] / (T(8) * o.dx);
          out[k + nz] += T(9) * c; out[k] -= T(9) * c; out[k + 2 * nz] -= c; out[k - nz] += c;
        }
    }
  }
  void dzp_t(const std::vector<T>& w, std::vector<T>& out) const {   // (Dz+)^T
    const uint32_t nz = o.nz;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < int64_t(o.nx); ++ii)
      for (uint32_t j = 1; j + 2 < nz; ++j) {
        const uint32_t k = uint32_t(ii) * nz + j;
        const T c = w[k] / (T(8) * o.dz);
        out[k + 1] += T(9) * c; out[k] -= T(9) * c; out[k + 2] -= c; out[k - 1] += c;
      }
  }
  // sources firing in run isrc, as per-cell lists in source order: head[k] is
  // the first, next[s] the following one at the same cell (-1: none)
  void source_lists(int32_t isrc, std::vector<int32_t>& head, std::vector<int32_t>& next) const {
    head.assign(o.npt, -1);
    next.assign(o.nsrc, -1);
    for (int32_t s = int32_t(o.nsrc) - 1; s >= 0; --s) {
      if (isrc >= 0 && s != isrc) continue;
      const uint32_t k = o.src_id(uint32_t(s));
      if (k >= o.npt) continue;
      next[s] = head[k];
      head[k] = s;
    }
  }
  // forward operators on a standalone field (same edge rules as fd2d.wgsl)
  T dxm(const T* f, uint32_t k) const {
    uint32_t i = k / o.nz; const uint32_t nz = o.nz;
    if (i < 2u || i >= o.nx - 2u) return T(0);
    return (T(9) * (f[k] - f[k - nz]) - (f[k + nz] - f[k - 2 * nz])) / (T(8) * o.dx);
  }
  T dzm(const T* f, uint32_t k) const {
    uint32_t j = k % o.nz;
    if (j < 2u || j >= o.nz - 2u) return T(0);
    return (T(9) * (f[k] - f[k - 1]) - (f[k + 1] - f[k - 2])) / (T(8) * o.dz);
  }
  T dxp(const T* f, uint32_t k) const {
    uint32_t i = k / o.nz; const uint32_t nz = o.nz;
    if (i < 1u || i >= o.nx - 2u) return T(0);
    return (T(9) * (f[k + nz] - f[k]) - (f[k + 2 * nz] - f[k - nz])) / (T(8) * o.dx);
  }
  T dzp(const T* f, uint32_t k) const {
    uint32_t j = k % o.nz;
    if (j < 1u || j >= o.nz - 2u) return T(0);
    return (T(9) * (f[k + 1] - f[k]) - (f[k + 2] - f[k - 1])) / (T(8) * o.dz);
  }

  std::vector<T> taper() const {
    const T PI_T = T(3.14159265358979323846);
    std::vector<T> w(o.nt);
    T t_end = T(o.nt - 1) * o.dt, tw = t_end / T(10), t_min = tw, t_max = t_end - tw;
    for (uint32_t n = 0; n < o.nt; ++n) {
      T t = T(n) * o.dt;
      if (t <= t_min) w[n] = T(0.5) + T(0.5) * std::cos(PI_T * (t_min - t) / tw);
      else if (t >= t_max) w[n] = T(0.5) + T(0.5) * std::cos(PI_T * (t_max - t) / tw);
      else w[n] = T(1);
    }
    return w;
  }

  // state field ids per physics
  std::vector<uint32_t> state_ids() const {
    if (o.sh) return {F_VY, F_SXY, F_SZY};
    return {F_VX, F_VZ, F_SXX, F_SZZ, F_SXZ};
  }
  std::vector<uint32_t> comps() const {
    if (o.sh) return {OBS_Y};
    return {OBS_X, OBS_Z};
  }

  // the state fields (state_ids order) before step 0 and after every step;
  // entry n + 1 is filled by the forward run's step n
  struct States {
    std::vector<std::unique_ptr<T[]>> s;
    const T* operator[](size_t n) const { return s[n].get(); }
  };

  // J for run isrc against obs [ncomp][nrec][nt]; optionally keeps all states
  double forward(int32_t isrc, const T* obs, States* states) {
    o.clear_dyn(); o.clear_obs();
    const auto ids = state_ids();
    const size_t npt = o.npt, per = ids.size() * npt;
    if (states) {
      states->s.clear();
      states->s.resize(size_t(o.nt) + 1);
      states->s[0].reset(new T[per]());   // the cleared state
    }
    for (uint32_t n = 0; n < o.nt; ++n) {
      o.step(n, isrc);
      if (states) {
        states->s[n + 1].reset(new T[per]);
        for (size_t q = 0; q < ids.size(); ++q) {
          const T* src = o.fields.data() + size_t(ids[q]) * npt;
          T* dst = states->s[n + 1].get() + q * npt;
#pragma omp parallel for schedule(static)
          for (int64_t k = 0; k < int64_t(npt); ++k) dst[k] = src[k];
        }
      }
    }
    const auto w = taper();
    const auto cs = comps();
    double J = 0.0;
    for (size_t c = 0; c < cs.size(); ++c)
      for (uint32_t r = 0; r < o.nrec; ++r)
        for (uint32_t n = 0; n < o.nt; ++n) {
          const T d = o.obs[(size_t(cs[c]) * o.nrec + r) * o.nt + n] - obs[(c * o.nrec + r) * o.nt + n];
          const T e = w[n] * d;
          J += double(e) * double(e);
        }
    return J;
  }

  // adds this run's gradient into kr, km, kl; returns J
  double gradient(int32_t isrc, const T* obs, std::vector<T>& kr, std::vector<T>& km, std::vector<T>& kl) {
    States st;
    const double J = forward(isrc, obs, &st);
    const size_t npt = o.npt;
    const uint32_t nt = o.nt, nrec = o.nrec, nsrc = o.nsrc;
    const auto w = taper();
    const auto cs = comps();
    // reverse-accumulated adjoint sources A^n = sum_{k>=n} 2 w_k^2 (d^k - obs^k)
    std::vector<T> A(cs.size() * nrec * nt, T(0));
    for (size_t c = 0; c < cs.size(); ++c)
      for (uint32_t r = 0; r < nrec; ++r) {
        T acc = T(0);
        for (int64_t n = int64_t(nt) - 1; n >= 0; --n) {
          const size_t q = (c * nrec + r) * nt + size_t(n);
          const T d = o.obs[(size_t(cs[c]) * nrec + r) * nt + size_t(n)] - obs[q];
          acc += T(2) * w[n] * w[n] * d;
          A[q] = acc;
        }
      }
    const T dt = o.dt;
    auto P = [&](uint32_t f, uint32_t k) { return o.fields[size_t(f) * npt + k]; };
    std::vector<int32_t> src_head, src_next;
    source_lists(isrc, src_head, src_next);
    (void)nsrc;
    if (o.sh) {
      std::vector<T> vb(npt, 0), sbx(npt, 0), sbz(npt, 0), tx(npt), tz(npt), q(npt), ex(npt), ez(npt);
      for (int64_t n = int64_t(nt) - 1; n >= 0; --n) {
        const T* s1 = st[size_t(n) + 1];   // after step n: vy, sxy, szy
        const T* s0 = st[size_t(n)];       // before step n
        const T *vy1 = s1, *sxy0 = s0 + npt, *szy0 = s0 + 2 * npt;
        for (uint32_t r = 0; r < nrec; ++r) vb[o.rec_id(r)] += dt * A[size_t(r) * nt + size_t(n)];
#pragma omp parallel for schedule(static)
        for (size_t k = 0; k < npt; ++k) {
          const T mu = P(F_MU, uint32_t(k));
          tx[k] = dt * mu * sbx[k];
          tz[k] = dt * mu * sbz[k];
          km[k] += dt * (sbx[k] * dxp(vy1, uint32_t(k)) + sbz[k] * dzp(vy1, uint32_t(k)));
        }
        dxp_t(tx, vb); dzp_t(tz, vb);
#pragma omp parallel for schedule(static)
        for (size_t k = 0; k < npt; ++k) {
          const T b = P(F_BOUND, uint32_t(k)), rho = P(F_RHO, uint32_t(k));
          const T g = b * vb[k];
          T ds = dxm(sxy0, uint32_t(k)) + dzm(szy0, uint32_t(k));
          for (int32_t s = src_head[k]; s >= 0; s = src_next[s]) ds += o.stf[(size_t(STF_Y) * o.nsrc + s) * nt + size_t(n)];
          kr[k] += -dt * g * ds / (rho * rho);
          q[k] = dt * g / rho;
          vb[k] = g;
        }
        dxm_t(q, sbx); dzm_t(q, sbz);
      }
    } else {
      std::vector<T> vbx(npt, 0), vbz(npt, 0), sbxx(npt, 0), sbzz(npt, 0), sbxz(npt, 0);
      std::vector<T> txx(npt), tzz(npt), txz(npt), qx(npt), qz(npt);
      for (int64_t n = int64_t(nt) - 1; n >= 0; --n) {
        const T* s1 = st[size_t(n) + 1];   // vx, vz, sxx, szz, sxz after step n
        const T* s0 = st[size_t(n)];
        const T *vx1 = s1, *vz1 = s1 + npt;
        const T *sxx0 = s0 + 2 * npt, *szz0 = s0 + 3 * npt, *sxz0 = s0 + 4 * npt;
        for (uint32_t r = 0; r < nrec; ++r) {
          vbx[o.rec_id(r)] += dt * A[(0 * size_t(nrec) + r) * nt + size_t(n)];
          vbz[o.rec_id(r)] += dt * A[(1 * size_t(nrec) + r) * nt + size_t(n)];
        }
#pragma omp parallel for schedule(static)
        for (size_t k = 0; k < npt; ++k) {
          const uint32_t kk = uint32_t(k);
          const T lam = P(F_LAM, kk), mu = P(F_MU, kk);
          txx[k] = dt * ((lam + T(2) * mu) * sbxx[k] + lam * sbzz[k]);
          tzz[k] = dt * (lam * sbxx[k] + (lam + T(2) * mu) * sbzz[k]);
          txz[k] = dt * mu * sbxz[k];
          const T exx = dxp(vx1, kk), ezz = dzp(vz1, kk), exz = dzp(vx1, kk) + dxp(vz1, kk);
          kl[k] += dt * (sbxx[k] + sbzz[k]) * (exx + ezz);
          km[k] += dt * (T(2) * sbxx[k] * exx + T(2) * sbzz[k] * ezz + sbxz[k] * exz);
        }
        dxp_t(txx, vbx); dzp_t(txz, vbx);
        dzp_t(tzz, vbz); dxp_t(txz, vbz);
#pragma omp parallel for schedule(static)
        for (size_t k = 0; k < npt; ++k) {
          const uint32_t kk = uint32_t(k);
          const T b = P(F_BOUND, kk), rho = P(F_RHO, kk);
          const T gx = b * vbx[k], gz = b * vbz[k];
          T dsx = dxm(sxx0, kk) + dzm(sxz0, kk), dsz = dxm(sxz0, kk) + dzm(szz0, kk);
          for (int32_t s = src_head[k]; s >= 0; s = src_next[s]) {
            dsx += o.stf[(size_t(STF_X) * o.nsrc + s) * nt + size_t(n)];
            dsz += o.stf[(size_t(STF_Z) * o.nsrc + s) * nt + size_t(n)];
          }
          kr[k] += -dt * (gx * dsx + gz * dsz) / (rho * rho);
          qx[k] = dt * gx / rho;
          qz[k] = dt * gz / rho;
          vbx[k] = gx;
          vbz[k] = gz;
        }
        dxm_t(qx, sbxx); dzm_t(qx, sbxz);
        dxm_t(qz, sbxz); dzm_t(qz, sbzz);
      }
    }
    return J;
  }
};

}  // namespace

// ------------------------------------------------------------------------
// C interface used by oracle/oracle.py.  prec: 4 = float, 8 = double.
// Host arrays crossing this interface are always double (converted to T).
// ------------------------------------------------------------------------
extern "C" {

struct or_handle {
  int prec;
  void* o;
};

or_handle* or_new(int prec, const or_cfg* c) {
  or_handle* h = new or_handle;
  h->prec = prec;
  if (prec == 4) h->o = new Oracle<float>(*c);
  else h->o = new Oracle<double>(*c);
  return h;
}

void or_free(or_handle* h) {
  if (!h) return;
  if (h->prec == 4) delete static_cast<Oracle<float>*>(h->o);
  else delete static_cast<Oracle<double>*>(h->o);
  delete h;
}

#define DISPATCH(h, ...)                                                \
  do {                                                                  \
    if ((h)->prec == 4) { auto& o = *static_cast<Oracle<float>*>((h)->o); __VA_ARGS__; } \
    else { auto& o = *static_cast<Oracle<double>*>((h)->o); __VA_ARGS__; } \
  } while (0)

void or_set_field(or_handle* h, int32_t fid, const double* v) {
  DISPATCH(h, { for (uint32_t k = 0; k < o.npt; ++k) o.fields[size_t(fid) * o.npt + k] = decltype(o.dt)(v[k]); });
}
void or_get_field(or_handle* h, int32_t fid, double* v) {
  DISPATCH(h, { for (uint32_t k = 0; k < o.npt; ++k) v[k] = double(o.fields[size_t(fid) * o.npt + k]); });
}
void or_set_idata(or_handle* h, const int32_t* idx) {
  DISPATCH(h, { std::copy(idx, idx + o.nsrc + o.nrec, o.idata.begin()); });
}
void or_set_stf(or_handle* h, const double* v) {
  DISPATCH(h, { for (size_t n = 0; n < o.stf.size(); ++n) o.stf[n] = decltype(o.dt)(v[n]); });
}
void or_get_obs(or_handle* h, double* v) {
  DISPATCH(h, { for (size_t n = 0; n < o.obs.size(); ++n) v[n] = double(o.obs[n]); });
}
void or_set_obs(or_handle* h, const double* v) {
  DISPATCH(h, { for (size_t n = 0; n < o.obs.size(); ++n) o.obs[n] = decltype(o.dt)(v[n]); });
}
void or_init_model(or_handle* h) { DISPATCH(h, o.init_model()); }
void or_clear_dyn(or_handle* h) { DISPATCH(h, o.clear_dyn()); }
void or_clear_obs(or_handle* h) { DISPATCH(h, o.clear_obs()); }
void or_step(or_handle* h, int32_t it, int32_t isrc) { DISPATCH(h, o.step(uint32_t(it), isrc)); }
void or_adjoint_step_sh(or_handle* h, int32_t it) { DISPATCH(h, o.adjoint_step_sh(uint32_t(it))); }
void or_smooth(or_handle* h, double sigma) { DISPATCH(h, o.smooth(decltype(o.dt)(sigma))); }

// Forward run (solver.rs:976-1046): clears, runs nt steps; after step it with
// it>0 && it%snap==0, copies (vy, vx, vz, vy_c) into snaps[nsnap_taken][4][npt].
int32_t or_forward(or_handle* h, int32_t isrc, int32_t snap, double* snaps, int32_t max_snaps) {
  int32_t taken = 0;
  DISPATCH(h, {
    o.clear_dyn();
    o.clear_obs();
    const uint32_t ids[4] = {F_VY, F_VX, F_VZ, F_VY_C};
    for (uint32_t it = 0; it < o.nt; ++it) {
      o.step(it, isrc);
      if (snap > 0 && it > 0 && it % uint32_t(snap) == 0) {
        if (snaps && taken < max_snaps) {
          for (int q = 0; q < 4; ++q)
            for (uint32_t k = 0; k < o.npt; ++k)
              snaps[(size_t(taken) * 4 + q) * o.npt + k] = double(o.fields[size_t(ids[q]) * o.npt + k]);
        }
        ++taken;
      }
    }
  });
  return taken;
}

int32_t or_adjoint_plain(or_handle* h, int32_t nruns, const int32_t* isrc_list, const double* ref,
                         int32_t adj_interval, double smooth, double* misfits) {
  int32_t rc = 0;
  DISPATCH(h, {
    using T = decltype(o.dt);
    std::vector<T> r(size_t(nruns) * o.nrec * o.nt);
    for (size_t n = 0; n < r.size(); ++n) r[n] = T(ref[n]);
    rc = adjoint_plain<T>(o, nruns, isrc_list, r.data(), uint32_t(adj_interval), T(smooth), misfits);
  });
  return rc;
}

// Misfit J (sum over runs) of the forward scheme against obs
// [nruns][ncomp][nrec][nt] (ncomp = 1 for SH: uy; 2 for P-SV: ux, uz).
double or_objective(or_handle* h, int32_t nruns, const int32_t* isrc, const double* obs) {
  double J = 0.0;
  DISPATCH(h, {
    using T = decltype(o.dt);
    Adjoint<T> ad(o);
    const size_t per = ad.comps().size() * size_t(o.nrec) * o.nt;
    std::vector<T> ob(per);
    for (int32_t r = 0; r < nruns; ++r) {
      for (size_t q = 0; q < per; ++q) ob[q] = T(obs[size_t(r) * per + q]);
      J += ad.forward(isrc[r], ob.data(), nullptr);
    }
  });
  return J;
}

// Exact adjoint gradient of or_objective with respect to rho, mu, lambda.
double or_gradient(or_handle* h, int32_t nruns, const int32_t* isrc, const double* obs, double* k_rho, double* k_mu,
                   double* k_lam) {
  double J = 0.0;
  DISPATCH(h, {
    using T = decltype(o.dt);
    Adjoint<T> ad(o);
    const size_t per = ad.comps().size() * size_t(o.nrec) * o.nt;
    std::vector<T> ob(per), kr(o.npt, T(0)), km(o.npt, T(0)), kl(o.npt, T(0));
    for (int32_t r = 0; r < nruns; ++r) {
      for (size_t q = 0; q < per; ++q) ob[q] = T(obs[size_t(r) * per + q]);
      J += ad.gradient(isrc[r], ob.data(), kr, km, kl);
    }
    for (uint32_t k = 0; k < o.npt; ++k) { k_rho[k] = double(kr[k]); k_mu[k] = double(km[k]); k_lam[k] = double(kl[k]); }
  });
  return J;
}

// Ricker table (solver.rs:350-356, 721-739) in precision prec.
// src: [nsrc][4] = (f0, t0, ang_deg, amp) as f32 values; out: [3][nsrc][nt].
void or_ricker_table(int prec, int32_t nsrc, int32_t nt, float dt, const float* src, double* out) {
  size_t stride = size_t(nsrc) * nt;
  for (int32_t s = 0; s < nsrc; ++s) {
    float f0 = src[4 * s], t0 = src[4 * s + 1], ang = src[4 * s + 2], amp = src[4 * s + 3];
    for (int32_t it = 0; it < nt; ++it) {
      double x, y, z;
      if (prec == 4) {
        const float PI_F = 3.14159265358979323846f;
        float t = float(it) * dt;
        float pf = PI_F * f0;
        float a = pf * pf;                       // powi(2)
        float d = t - t0;
        float st = -amp * (1.0f - 2.0f * a * d * d) * std::exp(-a * d * d);
        const float RAD = PI_F / 180.0f;          // f32::to_radians
        float ar = ang * RAD;
        x = st * std::sin(ar); y = st; z = -st * std::cos(ar);
      } else {
        const double PI_D = 3.14159265358979323846;
        double t = double(it) * double(dt);
        double a = (PI_D * f0) * (PI_D * f0);
        double d = t - double(t0);
        double st = -double(amp) * (1.0 - 2.0 * a * d * d) * std::exp(-a * d * d);
        double ar = double(ang) * (PI_D / 180.0);
        x = st * std::sin(ar); y = st; z = -st * std::cos(ar);
      }
      out[0 * stride + size_t(s) * nt + it] = x;
      out[1 * stride + size_t(s) * nt + it] = y;
      out[2 * stride + size_t(s) * nt + it] = z;
    }
  }
}

// Grid inference (solver.rs:640-656).
void or_infer_grid(int32_t npt, const float* x, const float* z, int32_t* nx, int32_t* nz, float* dx, float* dz) {
  float xmax = -INFINITY, xmin = INFINITY, zmax = -INFINITY, zmin = INFINITY;
  for (int32_t k = 0; k < npt; ++k) {
    xmax = std::max(xmax, x[k]); xmin = std::min(xmin, x[k]);
    zmax = std::max(zmax, z[k]); zmin = std::min(zmin, z[k]);
  }
  float lx = xmax - xmin, lz = zmax - zmin;
  uint32_t nxv = uint32_t(std::round(std::sqrt(double(npt) * double(lx) / double(lz))));
  uint32_t nzv = uint32_t(std::round(std::sqrt(double(npt) * double(lz) / double(lx))));
  *nx = int32_t(nxv); *nz = int32_t(nzv);
  *dx = lx / float(nxv - 1);
  *dz = lz / float(nzv - 1);
}

int32_t or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// threads of the following parallel loops (timing on one core)
void or_set_num_threads(int32_t n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

// Misfit restated alone (solver.rs:432-467), for unit tests.
double or_misfit(int prec, const double* syn, const double* obs, int32_t nrec, int32_t nt, float dt, double* adstf) {
  size_t n = size_t(nrec) * nt;
  if (prec == 4) {
    std::vector<float> s(n), o(n), a(n);
    for (size_t q = 0; q < n; ++q) { s[q] = float(syn[q]); o[q] = float(obs[q]); }
    double m = waveform_misfit<float>(s.data(), o.data(), nrec, nt, dt, a.data());
    for (size_t q = 0; q < n; ++q) adstf[q] = a[q];
    return m;
  }
  std::vector<double> a(n);
  double m = waveform_misfit<double>(syn, obs, nrec, nt, double(dt), a.data());
  for (size_t q = 0; q < n; ++q) adstf[q] = a[q];
  return m;
}

}  // extern "C"
