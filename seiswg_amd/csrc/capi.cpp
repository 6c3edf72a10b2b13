// C ABI of the device module (include/seiswg_hip.h).
//
// Owns the HIP device state of one problem: model fields, 1-D taper tables,
// source time functions, receiver lists, ping-pong state buffers for a batch
// of shots, per-receiver displacement accumulators, traces and the device
// step clock.  Builds the per-tile injection/receiver lists the fused step
// kernels consume and drives the time loop (solver.rs:941-1119, 1298-1529).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/seiswg_hip.h"
#include "kernels.h"

using namespace sw;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(-3, "%s: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

// Device-memory cache.  Runs over many problems (gradients of an uneven
// batch, one context per problem) allocate and free buffers of tens of GB in
// turn; hipMalloc of memory that was just freed can then wait for the driver
// (measured: single gradient calls taking seconds instead of tens of ms).
// Freed blocks are kept per device and handed out again for requests of up
// to their size (at most twice the request); on an allocation failure the
// cache of that device is emptied and the allocation retried.  A free waits
// for the device to go idle first, as hipFree does, so a cached block is
// never still in use by a kernel.  SEISWG_NO_MEMORY_CACHE=1 turns it off.
struct MemCache {
  std::mutex mu;
  std::multimap<std::pair<int, size_t>, void*> blocks;   // (device, bytes) -> block
  bool on = getenv("SEISWG_NO_MEMORY_CACHE") == nullptr;
};
MemCache& mem_cache() {
  static MemCache* c = new MemCache;   // never destroyed: usable until exit
  return *c;
}
size_t cache_round(size_t bytes) {
  const size_t g = bytes >= (size_t(1) << 21) ? (size_t(1) << 21) : 4096;
  return (bytes + g - 1) / g * g;
}
size_t cache_trim(int dev) {
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  size_t freed = 0;
  for (auto it = c.blocks.begin(); it != c.blocks.end();) {
    if (dev >= 0 && it->first.first != dev) { ++it; continue; }
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(it->first.first);
    (void)hipFree(it->second);
    (void)hipSetDevice(cur);
    freed += it->first.second;
    it = c.blocks.erase(it);
  }
  return freed;
}
hipError_t cache_alloc(void** p, size_t& bytes, int dev) {
  MemCache& c = mem_cache();
  bytes = cache_round(bytes);
  if (c.on) {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.blocks.lower_bound({dev, bytes});
    if (it != c.blocks.end() && it->first.first == dev && it->first.second <= 2 * bytes) {
      *p = it->second;
      bytes = it->first.second;
      c.blocks.erase(it);
      return hipSuccess;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && c.on) {
    (void)hipGetLastError();
    cache_trim(dev);
    e = hipMalloc(p, bytes);
  }
  return e;
}
// idle: the caller has synchronised every stream that used the block, so
// the device-wide wait (what hipFree would do) can be skipped
void cache_free(void* p, size_t bytes, int dev, bool idle = false) {
  MemCache& c = mem_cache();
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != dev) (void)hipSetDevice(dev);
  if (c.on) {
    if (!idle) (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(c.mu);
    c.blocks.insert({{dev, bytes}, p});
  } else {
    (void)hipFree(p);
  }
  if (cur != dev) (void)hipSetDevice(cur);
}

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t bytes = 0;
  int dev = 0;
  ~DBuf() { release(); }
  void release(bool idle = false) {
    if (p) cache_free(p, bytes, dev, idle);
    p = nullptr;
    n = 0;
    bytes = 0;
  }
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    if (count == 0) return hipSuccess;
    (void)hipGetDevice(&dev);
    size_t b = count * sizeof(T);
    void* q = nullptr;
    hipError_t e = cache_alloc(&q, b, dev);
    if (e == hipSuccess) {
      p = static_cast<T*>(q);
      n = count;
      bytes = b;
    }
    return e;
  }
};

}  // namespace

namespace sw {
void set_last_error(const std::string& msg) { g_err = msg; }
}

// bookkeeping of a running gradient (one batch of shots)
struct GradPlan {
  int K = 1, nck = 0, nb = 0, ncomp = 1, naux = 0;
  int kgroup = 1;   // shots per kernel-accumulator set: the adjoint launches' shots per workgroup
  size_t state_floats = 0;
  std::vector<double> mis;
};

struct sw_ctx {
  sw_grid g{};
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t npt = 0;
  float h8x = 0, h8z = 0;
  // divisions by 8 dx, 8 dz in the step kernels (two-operation form, see const_div)
  struct ConstDiv { bool exact = false; float zh = 0, zl = 0; uint32_t lo_bits = 0; } cdx, cdz;
  bool model_set = false;
  // every rho / J in [2^-20, 2^20]: divisions by them may use the reciprocal form
  bool rho_fast = false, j_fast = false;
  bool recip_div = false;     // SW_DIV_RECIPROCAL: forward steps multiply by 1/y

  // model fields [MF_N][npt]: rho, lambda, mu, nu, j, lambda_c, mu_c, nu_c
  DBuf<float> model;
  float* mf(int f) { return model.p ? model.p + size_t(f) * size_t(npt) : nullptr; }
  DBuf<float> bx, bzb, bzt;
  std::vector<float> h_bx, h_bzb, h_bzt;

  // sources / receivers
  int nsrc = 0, nrec = 0;
  std::vector<int32_t> src_k, rec_k;
  DBuf<float> stf;   // [3][nsrc][nt]

  // tiling: chosen per batch by a heuristic unless fixed by sw_set_tiling or
  // sw_autotune
  Tiling tile{256, 64, 0, 0};
  bool tile_fixed = false;
  int group_req = 0;          // shots per forward workgroup requested (0: automatic)
  int ncu = 0;                // compute units of the device

  // batch
  int nshots = 0;
  std::vector<int32_t> shot_isrc;
  int slot_index[N_SLOTS];
  int nslots = 0;
  DBuf<float> state[2];       // [shot][slot][npt]
  int cur = 0;                // buffer holding the current state
  DBuf<float> uacc, obs;
  DBuf<int> clock;
  DBuf<int> inj_off, rec_off;
  DBuf<int4> inj, rec;
  // two-step launches (P-SV without spin): their tiling (6-column halos),
  // per-tile lists and the third state buffer they write between the two
  // ping-pong buffers (a pair of launches returns the state to state[cur])
  int steps_per_launch = 2;   // requested: 1 or 2
  bool two_ok = false;        // the prepared batch can use them
  Tiling tile2;
  bool tile2_fixed = false;   // set by sw_set_tiling2
  int group2_req = 0;
  DBuf<int> inj2_off, rec2_off;
  DBuf<int4> inj2, rec2;
  int nrec2_ent = 0;          // receiver entries of the multi-step tile lists
  DBuf<float> spare;
  int64_t steps_done = 0;
  // injection values of the current batch (forward: stf; adjoint: adjoint source)
  const float* val_x = nullptr;
  const float* val_y = nullptr;
  const float* val_z = nullptr;
  int nt_val = 0;
  bool record = true;
  bool sh_only = false;       // the adjoint workflow dispatches the SH kernels only
  int aux_period = 1;         // checkpoint interval of the running gradient

  // captured blocks of time steps (hipGraph), keyed by everything the
  // captured launches depend on (kernel choice, grid and the bytes of their
  // argument blocks: their hash and the bytes themselves), so a block is
  // replayed only with the arguments it was captured with; the graph is
  // kept with its executable until both go
  struct GraphEntry { uint64_t key; std::string bytes; hipGraph_t graph; hipGraphExec_t exec; uint64_t used; };
  std::vector<GraphEntry> graphs;
  uint64_t graph_clock = 0;
  int graph_steps = 16;       // steps per captured block (a multiple of 4)
  bool graphs_off = false;    // set after a failed capture (or by a caller)
  bool graphs_failed = false; // a capture failed: graph_error says why
  std::string graph_error;

  // plain adjoint
  DBuf<float> kmu, k_lam, k_rho, gtmp, wtab, sumx, sumz, fwd_store, adj_src, ref, taper;
  DBuf<double> partial, misfit_d;

  // differentiable gradient: adjoint state ([shot][slot][npt] ping-pong),
  // adjoint clock, receiver-injection lists, checkpoints, per-step aux
  // fields of the recomputed segment and per-shot kernels
  DBuf<float> adj[2];
  DBuf<int> aclock;
  DBuf<int> ainj_off;
  DBuf<int4> ainj;
  DBuf<float> ckpt, aux, kshot, adj_val, gref, gtaper;
  DBuf<double> gmis;
  DBuf<int> gcomps;
  GradPlan gp;
  std::vector<float> h_taper;
  int h_comps[2] = {0, 0};
  int h_clk[4] = {0, 0, 0, 0};

  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // snapshots: device staging and pinned host buffers (two of each), a copy
  // stream and its events, so a snapshot's transfer overlaps the next steps
  DBuf<float> snap_stage[2];
  float* snap_host[2] = {nullptr, nullptr};
  size_t snap_host_bytes = 0;
  hipStream_t copy_stream = nullptr;
  hipEvent_t snap_ready[2] = {nullptr, nullptr}, snap_copied[2] = {nullptr, nullptr};
  sw_timing timing{};
  int64_t launches = 0;

  // blow-up detection: forward runs scan the state of every shot at their
  // snapshot steps and after the last step; gradients and the plain adjoint
  // check their misfits and kernel fields
  bool finite_check = true;
  DBuf<int> dbad;              // smallest step found not finite (INT_MAX: none)
  int* bad_host = nullptr;     // pinned [2]: dbad as each snapshot slot's check left it
};

namespace {

bool uses_sh(const sw_ctx* c) { return c->g.sh != 0; }
bool uses_psv(const sw_ctx* c) { return c->g.psv != 0; }
bool uses_spin(const sw_ctx* c) { return c->g.psv != 0 && c->g.spin != 0; }

void compute_slots(sw_ctx* c) {
  for (int s = 0; s < N_SLOTS; ++s) c->slot_index[s] = -1;
  int n = 0;
  auto add = [&](int s) { c->slot_index[s] = n++; };
  if (uses_sh(c)) { add(S_VY); add(S_SXY); add(S_SZY); }
  if (uses_psv(c)) { add(S_VX); add(S_VZ); add(S_SXX); add(S_SZZ); add(S_SXZ); }
  if (uses_spin(c)) { add(S_VYC); add(S_SYXC); add(S_SYYC); add(S_SYZC); }
  c->nslots = n;
}

// set_bound (fd2d.wgsl:205-235) factored as bx(i) * bz_bottom(j) * bz_top(j)
// in the reference's multiplication order; edge tests keep u32 arithmetic.
void build_taper(sw_ctx* c) {
  const uint32_t nx = uint32_t(c->g.nx), nz = uint32_t(c->g.nz), w = uint32_t(c->g.abs_width);
  const float alpha = c->g.abs_alpha;
  c->h_bx.assign(nx, 1.0f);
  c->h_bzb.assign(nz, 1.0f);
  c->h_bzt.assign(nz, 1.0f);
  for (uint32_t i = 0; i < nx; ++i) {
    float b = 1.0f;
    if (c->g.abs_left && (i + 1u) < w) { float aw = alpha * float(w - i - 1u); b *= std::exp(-aw * aw); }
    if (c->g.abs_right && i > (nx - w)) { float aw = alpha * float(w + i - nx); b *= std::exp(-aw * aw); }
    c->h_bx[i] = b;
  }
  for (uint32_t j = 0; j < nz; ++j) {
    if (c->g.abs_bottom && j > (nz - w)) { float aw = alpha * float(w + j - nz); c->h_bzb[j] = std::exp(-aw * aw); }
    if (c->g.abs_top && (j + 1u) < w) { float aw = alpha * float(w - j - 1u); c->h_bzt[j] = std::exp(-aw * aw); }
  }
}

// Default tile shape: the widest strip and longest segment that still give
// the GPU's 256 CUs several workgroups each (>= 1024 workgroups).  Long
// segments amortise each workgroup's start-up (list lookups, first rows).
// Division by the constant y = 8h in the step kernels: fma(x, zh, x * zl)
// with 1/y split as zh + zl, zh = RN(1/y) and zl = RN(1/y - zh).  zl must
// not be negative: x = -0 would give (-0) + (+0) = +0 instead of -0, so
// where RN(1/y) lies above 1/y the split takes zh = RD(1/y) instead.  The
// kernels use the form only for a divisor it is exact for: checked here on
// every significand x in [1, 2) against the IEEE quotient (about 30 ms per
// divisor, cached; both splits fail for about 2 % of divisors, which then
// take the IEEE division), which carries over to x 2^k while x zl, x zh and
// x / y stay normal, so numerators of magnitude below the threshold T
// (>= 2^-100, one binade above where those products leave the normal range)
// take the IEEE division.
sw_ctx::ConstDiv const_div(float y) {
  static std::mutex mu;
  static std::map<uint32_t, sw_ctx::ConstDiv> cache;
  uint32_t key;
  std::memcpy(&key, &y, 4);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  sw_ctx::ConstDiv d;
  float zh = 1.0f / y;
  if (double(zh) > 1.0 / double(y)) zh = std::nextafter(zh, 0.0f);
  d.zh = zh;
  d.zl = float(1.0 / double(y) - double(zh));
  d.exact = fast_divisor(y) && d.zl >= 0.0f;
  for (uint32_t m = 0; d.exact && m < (1u << 23); ++m) {
    const uint32_t b = 0x3f800000u | m;
    float x;
    std::memcpy(&x, &b, 4);
    if (std::fma(x, zh, x * d.zl) != x / y) d.exact = false;
  }
  double t = 0x1p-100;
  if (d.zl != 0.0f) t = std::max(t, 0x1p-125 / std::fabs(double(d.zl)));
  t = std::max(t, 0x1p-125 * double(y));
  t = std::ldexp(1.0, int(std::ceil(std::log2(t))));
  const float tf = float(t);
  uint32_t tb;
  std::memcpy(&tb, &tf, 4);
  d.lo_bits = (tb << 1) - 1u;
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = d;
  return d;
}

bool all_fast_divisors(const float* v, size_t n) {
  for (size_t k = 0; k < n; ++k)
    if (!fast_divisor(v[k])) return false;
  return true;
}

// Shots per forward workgroup: the requested count (0: 2 where the tile is
// wide enough), reduced to one that the kernels support and that divides
// the number of shots.
int effective_group(int nt, int requested, int nshots, bool spin) {
  int g = requested > 0 ? requested : (nt >= 128 ? 2 : 1);
  while (g > 1 && (!group_supported(nt, g) || nshots % g != 0 || (spin && nt * g > 512))) g /= 2;
  return g;
}
int effective_group2(int nt, int requested, int nshots) {
  int g = requested > 0 ? requested : 1;
  while (g > 1 && (!group2_supported(nt, g) || nshots % g != 0)) g /= 2;
  return g;
}

// multi-step tiles: LDS tiles (K = lds_steps steps per launch, halo 3K) or
// two-step marching tiles (halo 6); columns of a strip and of its region
bool tile2_lds(const Tiling& t) { return t.nt_threads == kLdsTileThreads; }
int tile2_halo(const Tiling& t) { return tile2_lds(t) ? 3 * t.lds_steps : 3 * t.msteps; }
int tile2_strip_cols(const Tiling& t) {
  return tile2_lds(t) ? lds_tile_cols(t.lds_steps) : t.nt_threads - 2 * tile2_halo(t);
}
// three steps per marching launch: requested (sw_set_steps_per_launch) and P-SV alone
bool three_steps(const sw_ctx* c) {
  return c->steps_per_launch == 3 && uses_psv(c) && !uses_sh(c) && !uses_spin(c);
}
int tile2_region_cols(const Tiling& t) { return tile2_strip_cols(t) + 2 * tile2_halo(t); }

// Steps per LDS-tile launch: up to 4 (the region's 32 rows keep 8 after
// halos of 12), up to 3 with spin (its nine fields and eight moduli)
int lds_max_steps(const sw_ctx* c) { return uses_spin(c) ? 3 : 4; }
int lds_default_steps(const sw_ctx* c) { return uses_spin(c) ? 2 : 4; }

// Batches of at most this many shot-cells take LDS tiles for their two-step
// launches (latency-bound: a marching workgroup waits on its loads once per row)
constexpr int64_t kLdsTileMaxShotCells = 300000;

// Shape of the two-step launches (6-column halos, W = threads-12): the tile
// width that wastes the fewest lanes on the last strip (the wider on ties;
// for grids too small to fill the CUs, the one that gives more workgroups),
// shots per workgroup as requested (else 2 up to 256 threads, 1 above), and
// as many segments as fill the resident slots of every CU once, at least 8
// rows each, 4 on grids too small to fill the CUs (latency-bound: many short
// segments beat few long ones despite the 9 recomputed rows per segment).
void heuristic_tiling2(sw_ctx* c, int nshots) {
  Tiling& t = c->tile2;
  const int nz = c->g.nz;
  if (int64_t(nshots) * c->npt <= kLdsTileMaxShotCells) {
    // latency-bound batches: LDS tiles (each workgroup loads its region
    // once).  A launch of K steps takes about a + b K per round of resident
    // workgroups (a: launch, region load and store; b: one step in LDS;
    // measured a / b = 3.5 for SH, 6.7 with spin): the K with the fewest
    // rounds (4 + K) / K per step.
    t.nt_threads = kLdsTileThreads;
    const int64_t ncu = c->ncu > 0 ? c->ncu : 256;
    double best_cost = 1e300;
    for (int k = 2; k <= lds_max_steps(c); ++k) {
      const int64_t wg = int64_t(nshots) * ((nz + lds_tile_cols(k) - 1) / lds_tile_cols(k)) *
                         ((c->g.nx + lds_tile_rows(k) - 1) / lds_tile_rows(k));
      const int64_t slots = ncu * lds_tile_blocks_per_cu(uses_sh(c), uses_spin(c), k);
      const double cost = double((wg + slots - 1) / slots) * (4.0 + k) / k;
      if (cost < best_cost) {
        best_cost = cost;
        t.lds_steps = k;
      }
    }
    t.seglen = lds_tile_rows(t.lds_steps);
    t.group = 1;
    return;
  }
  t.lds_steps = 0;
  const int64_t ncu = c->ncu > 0 ? c->ncu : 256;
  if (three_steps(c)) {
    // three steps per launch (9-column halos, one shot per workgroup): the
    // width that wastes the fewest lanes on the last strip, segments that
    // fill every CU's resident slots once
    t.msteps = 3;
    int best3 = 384;
    double best_eff = -1.0;
    for (int nt : {512, 384, 256}) {
      const int64_t strips = (nz + nt - 18 - 1) / (nt - 18);
      const double eff = double(nz) / (double(strips) * nt);
      if (eff > best_eff + 1e-9) {
        best_eff = eff;
        best3 = nt;
      }
    }
    const int64_t resident = std::max(step3_blocks_per_cu(best3), 1);
    const int64_t base = int64_t(nshots) * ((nz + best3 - 18 - 1) / (best3 - 18));
    const int64_t nsegs = std::max<int64_t>(1, (ncu * resident) / base);
    t.nt_threads = best3;
    t.seglen = std::min(int(std::max<int64_t>((c->g.nx + nsegs - 1) / nsegs, 16)), c->g.nx);
    t.group = 1;
    return;
  }
  t.msteps = 2;
  int best = 64;
  double best_score = -1.0;
  for (int nt : {512, 256, 128, 64}) {
    const int64_t strips = (nz + nt - 12 - 1) / (nt - 12);
    const double eff = double(nz) / (double(strips) * nt);
    // a grid too small to give every CU a workgroup even with 8-row
    // segments prefers the narrower tile that gives more of them
    const int64_t most = int64_t(nshots) * strips * std::max(1, (c->g.nx + 7) / 8);
    const double score = eff * std::min(1.0, double(most) / double(ncu));
    if (score > best_score + 1e-9) {
      best_score = score;
      best = nt;
    }
  }
  // two shots per workgroup share their model rows; measured faster with
  // tiles up to 256 wide (1000x600, 16 shots: 106.6 against 102.5 shots/s
  // for the gradient), not with 512-wide tiles (2000x1000, 64 shots)
  const int group = effective_group2(best, c->group2_req > 0 ? c->group2_req : (best <= 256 ? 2 : 1), nshots);
  const int64_t resident = std::max(step2_blocks_per_cu(best, group), 1);
  const int64_t strips = (nz + best - 12 - 1) / (best - 12);
  const int64_t base = int64_t(nshots / group) * strips;
  int64_t nsegs = std::max<int64_t>(1, (ncu * resident) / base);
  // two rounds of resident workgroups where segments stay long (at least 128
  // rows): 2000x1000 x 64 shots, 250-row segments (two rounds) 1.382 ms per
  // launch against 1.433 with 500 rows (one round); 1.5 or 2.5 rounds (334 or
  // 200 rows) 1.480 and 1.444
  if ((c->g.nx + 2 * nsegs - 1) / (2 * nsegs) >= 128) nsegs *= 2;
  const int min_seg = base * ((c->g.nx + 7) / 8) < ncu ? 4 : 8;
  int seg = int(std::max<int64_t>((c->g.nx + nsegs - 1) / nsegs, min_seg));
  t.nt_threads = best;
  t.seglen = std::min(seg, c->g.nx);
  t.group = group;
}

// Tile width: 256 threads from 400 columns up, 128 from 100 (wide tiles
// measured faster even where they leave more lanes beyond nz idle, and on
// grids too small to fill the CUs: 500x500 P-SV+spin, one shot, 23.3 us per
// step with 256 threads against 26.9 with 128).
// Shots per workgroup: the largest of 4, 2, 1 that divides the batch and
// whose resident workgroups still give a CU at least 16 waves; more shots per
// workgroup read each model row fewer times.  Segments: as many as fill every
// CU's resident slots once (no second round), but at least 8 rows each (4 on
// grids too small to fill the CUs).
void heuristic_tiling(sw_ctx* c, int nshots) {
  Tiling& t = c->tile;
  const int64_t ncu = c->ncu > 0 ? c->ncu : 256;
  auto strips_of = [&](int w) { return int64_t((c->g.nz + (w - 6) - 1) / (w - 6)); };
  const int nt = c->g.nz >= 400 ? 256 : c->g.nz >= 100 ? 128 : 64;
  const bool spin = uses_spin(c);
  const int kind = uses_psv(c) ? (spin ? MK_SPIN : MK_PSV) : MK_SH;
  const int64_t nstrips = strips_of(nt);
  int group = 1;
  int64_t resident = std::max(step_blocks_per_cu(kind, nt, 1), 1);
  for (int g : {4, 2, 1}) {
    if (g != effective_group(nt, c->group_req > 0 ? c->group_req : g, nshots, spin)) continue;
    const int64_t r = step_blocks_per_cu(kind, nt, g);
    if (g > 1 && r * g * nt / 64 < 16) continue;
    group = g;
    resident = std::max<int64_t>(r, 1);
    break;
  }
  const int64_t base = int64_t(nshots / group) * nstrips;
  const int64_t nsegs = std::max<int64_t>(1, (ncu * resident) / base);
  // a grid that 8-row segments cannot spread over every CU is latency-bound:
  // 4-row segments (fewer rows marched per workgroup) measured faster there
  const int min_seg = base * ((c->g.nx + 7) / 8) < ncu ? 4 : 8;
  int seg = int(std::max<int64_t>((c->g.nx + nsegs - 1) / nsegs, min_seg));
  seg = std::min(seg, c->g.nx);
  t.nt_threads = nt;
  t.seglen = seg;
  t.group = group;
}

int check_tiling(sw_ctx* c, int nshots) {
  Tiling& t = c->tile;
  if (!tiling_supported(t.nt_threads)) return fail(-2, "unsupported tile width %d", t.nt_threads);
  if (t.seglen < 1) return fail(-2, "seglen must be >= 1");
  const int W = t.nt_threads - 6;
  t.nstrips = (c->g.nz + W - 1) / W;
  t.nsegs = (c->g.nx + t.seglen - 1) / t.seglen;
  t.group = effective_group(t.nt_threads, c->tile_fixed ? c->group_req : t.group, nshots, uses_spin(c));
  // two-step launches: their own shape if set or chosen by the heuristic,
  // else that of the one-step launches
  // two-step launches: marching tiles for SH or P-SV alone, LDS tiles also
  // for P-SV + spin
  c->two_ok = c->steps_per_launch >= 2 && uses_psv(c) != uses_sh(c);
  if (c->two_ok) {
    Tiling& t2 = c->tile2;
    if (!c->tile2_fixed && c->tile_fixed) {
      if (three_steps(c)) {
        heuristic_tiling2(c, nshots);   // the one-step widths are not three-step ones
      } else {
        t2 = t;
        c->group2_req = c->group_req;
      }
    }
    if (tile2_lds(t2)) {
      if (t2.lds_steps < 2 || t2.lds_steps > lds_max_steps(c)) t2.lds_steps = lds_default_steps(c);
      t2.seglen = lds_tile_rows(t2.lds_steps);
    } else {
      t2.lds_steps = 0;
      if (uses_spin(c)) c->two_ok = false;
      t2.msteps = three_steps(c) ? 3 : 2;
      if (t2.msteps == 3) {
        if (!tiling3_supported(t2.nt_threads))
          return fail(-2, "three steps per launch take tiles of 256, 384 or 512 threads (got %d)", t2.nt_threads);
        t2.group = 1;
        c->group2_req = 1;
      } else if (!tiling2_supported(t2.nt_threads)) {
        return fail(-2, "two steps per launch take tiles of 64, 128, 256, 320 or 512 threads (got %d)",
                    t2.nt_threads);
      }
    }
  }
  if (c->two_ok) {
    Tiling& t2 = c->tile2;
    const int w = tile2_strip_cols(t2);
    t2.nstrips = (c->g.nz + w - 1) / w;
    t2.nsegs = (c->g.nx + t2.seglen - 1) / t2.seglen;
    t2.group = effective_group2(t2.nt_threads, c->group2_req > 0 ? c->group2_req : t2.group, nshots);
  }
  return 0;
}

// Per-tile CSR lists.  Injection entries of a tile: every point whose new
// velocity the tile computes (rows [i0-1, i1+2), columns [c0, c0+NT)).
// Receiver entries: receivers whose cell the tile owns.  Both sorted by row,
// stable, so duplicates of one cell stay in index order.
struct Entry { int i, j, q; };

// tile `tile` of the one-step tiling (two2 false) or of the multi-step one
void tile_bounds(const sw_ctx* c, int tile, int& i0, int& i1, int& c0, bool two2 = false) {
  const Tiling& t = two2 ? c->tile2 : c->tile;
  const int strip = tile % t.nstrips, seg = tile / t.nstrips;
  i0 = seg * t.seglen;
  i1 = std::min(i0 + t.seglen, c->g.nx);
  c0 = two2 ? strip * tile2_strip_cols(t) - tile2_halo(t) : strip * (t.nt_threads - 6) - 3;
}

int upload_inj(sw_ctx* c, const std::vector<std::vector<Entry>>& shot_inj, DBuf<int>& off_d, DBuf<int4>& data_d) {
  const int ntile = c->tile.nstrips * c->tile.nsegs;
  const int NT = c->tile.nt_threads;
  std::vector<int> off(size_t(c->nshots) * ntile + 1, 0);
  std::vector<int4> data;
  for (int s = 0; s < c->nshots; ++s) {
    std::vector<Entry> pts = shot_inj[s];
    std::stable_sort(pts.begin(), pts.end(), [](const Entry& a, const Entry& b) { return a.i < b.i; });
    for (int t = 0; t < ntile; ++t) {
      int i0, i1, c0;
      tile_bounds(c, t, i0, i1, c0);
      off[size_t(s) * ntile + t] = int(data.size());
      for (const Entry& e : pts)
        if (e.i >= i0 - 1 && e.i < i1 + 2 && e.j >= c0 && e.j < c0 + NT) data.push_back(make_int4(e.i, e.j, e.q, 0));
    }
  }
  off.back() = int(data.size());
  HIPCHK(off_d.ensure(off.size()));
  HIPCHK(hipMemcpyAsync(off_d.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(data_d.ensure(std::max<size_t>(data.size(), 1)));
  if (!data.empty())
    HIPCHK(hipMemcpyAsync(data_d.p, data.data(), data.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// Lists of one tiling: the one-step tiles (halo 3) or the multi-step ones
// (two2: two-step marching tiles, halo 6, whose first stage computes
// velocities on rows [i0-4, i1+4]; LDS tiles, halo 3K, their whole region).
int build_tile_lists(sw_ctx* c, const std::vector<std::vector<Entry>>& shot_inj, bool two2, DBuf<int>& off_d,
                     DBuf<int4>& data_d, DBuf<int>& roff_d, DBuf<int4>& rdata_d) {
  const Tiling& tl = two2 ? c->tile2 : c->tile;
  const int ntile = tl.nstrips * tl.nsegs;
  const int halo = two2 ? tile2_halo(tl) : 3;
  const int NT = two2 ? tile2_region_cols(tl) : tl.nt_threads;
  // rows whose new velocities a tile computes: one-step marching rows
  // [i0-1, i1+2), two-step marching [i0-4, i1+5), LDS tiles their region
  const bool lds = two2 && tile2_lds(tl);
  const int before = lds ? halo : two2 ? 3 * tl.msteps - 2 : 1, after = lds ? halo : two2 ? 3 * tl.msteps - 1 : 2;
  std::vector<int> off(size_t(c->nshots) * ntile + 1, 0);
  std::vector<int4> data;
  for (int s = 0; s < c->nshots; ++s) {
    std::vector<Entry> pts = shot_inj[s];
    std::stable_sort(pts.begin(), pts.end(), [](const Entry& a, const Entry& b) { return a.i < b.i; });
    for (int t = 0; t < ntile; ++t) {
      int i0, i1, c0;
      tile_bounds(c, t, i0, i1, c0, two2);
      off[size_t(s) * ntile + t] = int(data.size());
      for (const Entry& e : pts)
        if (e.i >= i0 - before && e.i < i1 + after && e.j >= c0 && e.j < c0 + NT)
          data.push_back(make_int4(e.i, e.j, e.q, 0));
    }
  }
  off.back() = int(data.size());
  // receivers
  std::vector<Entry> recs;
  for (int r = 0; r < c->nrec; ++r) {
    const uint32_t k = uint32_t(c->rec_k[r]);
    if (int64_t(k) >= c->npt) continue;   // outside the grid: trace stays zero
    recs.push_back({int(k / uint32_t(c->g.nz)), int(k % uint32_t(c->g.nz)), r});
  }
  std::stable_sort(recs.begin(), recs.end(), [](const Entry& a, const Entry& b) { return a.i < b.i; });
  std::vector<int> roff(ntile + 1, 0);
  std::vector<int4> rdata;
  for (int t = 0; t < ntile; ++t) {
    int i0, i1, c0;
    tile_bounds(c, t, i0, i1, c0, two2);
    roff[t] = int(rdata.size());
    for (const Entry& e : recs)
      if (e.i >= i0 && e.i < i1 && e.j >= c0 + halo && e.j < c0 + NT - halo && e.j < c->g.nz)
        rdata.push_back(make_int4(e.i, e.j, e.q, t));   // t: the owning tile (k_record2)
  }
  roff[ntile] = int(rdata.size());
  if (two2) c->nrec2_ent = int(rdata.size());
  HIPCHK(off_d.ensure(off.size()));
  HIPCHK(hipMemcpyAsync(off_d.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(data_d.ensure(std::max<size_t>(data.size(), 1)));
  if (!data.empty())
    HIPCHK(hipMemcpyAsync(data_d.p, data.data(), data.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(roff_d.ensure(roff.size()));
  HIPCHK(hipMemcpyAsync(roff_d.p, roff.data(), roff.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(rdata_d.ensure(std::max<size_t>(rdata.size(), 1)));
  if (!rdata.empty())
    HIPCHK(hipMemcpyAsync(rdata_d.p, rdata.data(), rdata.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));   // host vectors go out of scope
  return 0;
}

int build_lists(sw_ctx* c, const std::vector<std::vector<Entry>>& shot_inj) {
  int rc = build_tile_lists(c, shot_inj, false, c->inj_off, c->inj, c->rec_off, c->rec);
  if (rc || !c->two_ok) return rc;
  return build_tile_lists(c, shot_inj, true, c->inj2_off, c->inj2, c->rec2_off, c->rec2);
}

std::vector<Entry> source_entries(const sw_ctx* c, int isrc) {
  std::vector<Entry> v;
  for (int q = 0; q < c->nsrc; ++q) {
    if (isrc >= 0 && q != isrc) continue;
    const uint32_t k = uint32_t(c->src_k[q]);
    if (int64_t(k) >= c->npt) continue;
    v.push_back({int(k / uint32_t(c->g.nz)), int(k % uint32_t(c->g.nz)), q});
  }
  return v;
}

// Allocate/zero the batch state for `nshots` shots with the given injection
// lists (solver.rs:978-981: clear dynamic fields and traces).
int prepare_batch(sw_ctx* c, int nshots, const std::vector<std::vector<Entry>>& inj, bool record) {
  if (!c->model_set) return fail(-2, "model not set");
  if (nshots < 1) return fail(-2, "nshots must be >= 1");
  if (!c->tile_fixed) heuristic_tiling(c, nshots);
  if (!c->tile2_fixed && !c->tile_fixed) heuristic_tiling2(c, nshots);
  int rc = check_tiling(c, nshots);
  if (rc) return rc;
  c->nshots = nshots;
  const size_t slot_elems = size_t(nshots) * size_t(c->npt);
  for (int p = 0; p < 2; ++p) {
    HIPCHK(c->state[p].ensure(std::max<size_t>(size_t(c->nslots) * slot_elems, 1)));
    HIPCHK(launch_fill(c->state[p].p, 0.0f, int64_t(c->nslots) * int64_t(slot_elems), c->stream));
  }
  const size_t nu = size_t(nshots) * std::max(c->nrec, 1) * N_UACC;
  HIPCHK(c->uacc.ensure(nu));
  HIPCHK(launch_fill(c->uacc.p, 0.0f, int64_t(nu), c->stream));
  const size_t no = size_t(nshots) * 5 * std::max(c->nrec, 1) * size_t(c->g.nt);
  HIPCHK(c->obs.ensure(no));
  HIPCHK(launch_fill(c->obs.p, 0.0f, int64_t(no), c->stream));
  HIPCHK(c->clock.ensure(2));
  const int zero2[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(c->clock.p, zero2, sizeof zero2, hipMemcpyHostToDevice, c->stream));
  rc = build_lists(c, inj);
  if (rc) return rc;
  c->cur = 0;
  c->steps_done = 0;
  c->record = record;
  return 0;
}

enum StepMode { M_FWD = 0, M_FWD_AUX, M_ADJ };

bool step_sh_on(const sw_ctx* c) { return uses_sh(c); }
bool step_psv_on(const sw_ctx* c) { return uses_psv(c) && !c->sh_only; }

// Launch arguments of one fused step of this context (the same for every
// step: buffers are selected by the parity of the device clock).
StepArgs make_args(sw_ctx* c, StepMode mode) {
  StepArgs a{};
  for (int q = 0; q < N_SLOTS; ++q) a.slot_off[q] = c->slot_index[q] < 0 ? -1 : int(c->slot_index[q] * c->npt);
  a.shot_floats = int(c->nslots * c->npt);
  a.model = c->model.p;
  a.npt = int(c->npt);
  a.bx = c->bx.p; a.bzb = c->bzb.p; a.bzt = c->bzt.p;
  a.nx = c->g.nx; a.nz = c->g.nz; a.nshots = c->nshots;
  a.nstrips = c->tile.nstrips; a.nsegs = c->tile.nsegs; a.seglen = c->tile.seglen;
  a.dt = c->g.dt; a.h8x = c->h8x; a.h8z = c->h8z;
  a.r8x = 1.0f / c->h8x; a.r8z = 1.0f / c->h8z;
  a.z8x = c->cdx.zh; a.z8z = c->cdz.zh;
  a.l8x = c->cdx.zl; a.l8z = c->cdz.zl;
  a.c_lo = std::max(c->cdx.lo_bits, c->cdz.lo_bits);
  a.fast_div = ((c->cdx.exact && c->cdz.exact) ? FD_CONST : 0) |
               (c->rho_fast ? FD_RHO : 0) | (c->j_fast ? FD_J : 0) | (c->recip_div ? FD_RCP : 0);
  a.inj_off = c->inj_off.p; a.inj = c->inj.p;
  a.val_x = c->val_x; a.val_y = c->val_y; a.val_z = c->val_z; a.nt_val = c->nt_val;
  a.rec_off = c->rec_off.p; a.rec = c->rec.p;
  a.uacc = c->uacc.p; a.obs = c->obs.p; a.nrec = c->nrec; a.nt = c->g.nt;
  a.record = (c->record && mode == M_FWD) ? 1 : 0;
  a.clock = c->clock.p;
  a.aux = c->aux.p;
  a.aux_period = c->aux_period;
  a.kgrad = c->kshot.p;
  // non-temporal state / aux stores once the state a launch reads and writes
  // is beyond the 256 MB Infinity Cache (below it, the next launch finds the
  // state there)
  a.stream_out = 2.0 * c->nshots * c->nslots * double(c->npt) * 4.0 > double(256 << 20) ? 1 : 0;
  // workgroup order: the shot groups of a tile side by side (they share its
  // model rows in one XCD's L2): 0.6-1.4 % faster on the 64-shot two-step
  // workload, neutral on the 16-shot gradient; SEISWG_TILE_MAJOR=0 restores
  // the shot-major order
  {
    const char* tm = getenv("SEISWG_TILE_MAJOR");
    a.tile_major = (tm && tm[0] == '0') ? 0 : 1;
  }
  a.st[0] = c->state[0].p;
  a.st[1] = c->state[1].p;
  a.it_lo = -0x7fffffff;
  a.it_hi = 0x7fffffff;
  // forward and adjoint steps run tile.group shots per workgroup
  a.nblocks = c->nshots / c->tile.group * c->tile.nstrips * c->tile.nsegs;
  a.advance = 1;
  if (mode == M_ADJ) {
    a.inj_off = c->ainj_off.p; a.inj = c->ainj.p;
    a.val_x = c->adj_val.p;
    a.val_y = c->adj_val.p;
    a.val_z = c->adj_val.p + size_t(c->nshots) * c->nrec * c->g.nt;
    a.nt_val = c->g.nt;
    a.clock = c->aclock.p;
    a.st[0] = c->adj[0].p;
    a.st[1] = c->adj[1].p;
  }
  return a;
}

// The bytes of the captured launches' arguments (the argument blocks are
// value-initialised, so their padding is zero) and their FNV-1a hash: a
// cached block is reused only when both match
struct GraphKey {
  uint64_t h = 1469598103934665603ull;
  std::string bytes;
  void add(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t q = 0; q < n; ++q) h = (h ^ b[q]) * 1099511628211ull;
    bytes.append(static_cast<const char*>(p), n);
  }
};
constexpr size_t kMaxGraphs = 16;   // captured blocks kept per context
// Two-step launches of at least this many shot-cells leave their traces to a
// receiver kernel after them: recording in the march costs about 4.5 % of a
// launch (64 shots of 2000x1000: 62 us of 1.38 ms), the extra launch a few
// microseconds, so the march records them itself below ~8 M shot-cells (one
// shot of 2000x1000: 23.4 against 25.5 us per step with the extra launch).
constexpr int64_t kRecordKernelShotCells = int64_t(8) << 20;

int launch_steps(sw_ctx* c, int nsteps, StepMode mode = M_FWD) {
  StepArgs a = make_args(c, mode);
  const bool sh = step_sh_on(c), psv = step_psv_on(c), spin = uses_spin(c);
  const bool aux = mode == M_FWD_AUX;
  StepArgs a_sh = a;
  a_sh.advance = psv ? 0 : 1;   // with both physics the P-SV kernel advances the clock
  // Multi-step launches (K steps each: two for marching tiles, lds_steps for
  // LDS tiles) keep the buffer of every step that of its parity.  Odd K:
  // launches alternate state[p] -> state[p ^ 1] and back, one launch per
  // unit; even K: launches go in pairs, state[cur] -> spare -> state[cur]
  // (2K steps).  LDS tiles have no variant that stores the gradient's
  // velocities.
  const int ldsk = tile2_lds(c->tile2) ? c->tile2.lds_steps : 0;
  const bool multi_ok = (mode == M_FWD || (mode == M_FWD_AUX && !ldsk && c->tile2.msteps == 2)) && c->two_ok &&
                        c->steps_per_launch >= 2;
  const int K = ldsk ? ldsk : c->tile2.msteps;   // steps per multi-step launch
  const bool odd = (K & 1) != 0;
  const int U = odd ? K : 2 * K;                 // steps per launch unit
  const bool multi = multi_ok && nsteps >= U;
  StepArgs a2a{}, a2b{};
  if (multi) {
    a2a = a;
    a2a.nstrips = c->tile2.nstrips;
    a2a.nsegs = c->tile2.nsegs;
    a2a.seglen = c->tile2.seglen;
    a2a.inj_off = c->inj2_off.p; a2a.inj = c->inj2.p;
    a2a.rec_off = c->rec2_off.p; a2a.rec = c->rec2.p;
    a2a.nblocks = c->nshots / c->tile2.group * c->tile2.nstrips * c->tile2.nsegs;
    a2b = a2a;
    if (odd) {
      // a2a from the current parity, a2b from the other one
      a2a.st[0] = c->state[c->cur].p; a2a.st[1] = c->state[c->cur ^ 1].p;
      a2b.st[0] = c->state[c->cur ^ 1].p; a2b.st[1] = c->state[c->cur].p;
    } else {
      HIPCHK(c->spare.ensure(size_t(c->nslots) * size_t(c->nshots) * size_t(c->npt)));
      a2a.st[0] = c->state[c->cur].p; a2a.st[1] = c->spare.p;
      a2b.st[0] = c->spare.p; a2b.st[1] = c->state[c->cur].p;
    }
  }
  const int cur0 = c->cur;
  // traces of two-step marching launches: from a receiver kernel after each
  // launch of at least kRecordKernelShotCells shot-cells, from the march
  // itself below that (LDS tiles and three-step launches record their own)
  const bool marching2 = multi && a2a.record && !ldsk && c->tile2.msteps == 2 && c->nrec2_ent > 0;
  // SEISWG_RECORD_KERNEL=1 / 0: the receiver kernel for every / no such launch
  const char* rk = getenv("SEISWG_RECORD_KERNEL");
  const bool rec_kernel = marching2 && (rk && (rk[0] == '0' || rk[0] == '1')
                                            ? rk[0] == '1'
                                            : int64_t(c->nshots) * c->npt >= kRecordKernelShotCells);
  const bool rec_loop = marching2 && !rec_kernel;
  auto launch2 = [&](const StepArgs& x) -> int {
    HIPCHK(sh ? launch_step_sh2(x, c->tile2, c->stream, aux, rec_loop)
              : launch_step_psv2(x, c->tile2, c->stream, aux, spin, rec_loop));
    if (rec_kernel) HIPCHK(launch_record2(sh, x, c->nrec2_ent, c->stream));
    c->launches += 1;
    return 0;
  };
  // one launch unit (U steps); odd K: from the buffer of parity cur, or with
  // flip from the other one
  auto enqueue_unit = [&](bool flip) -> int {
    if (odd) return launch2(flip ? a2b : a2a);
    const int rc = launch2(a2a);
    return rc ? rc : launch2(a2b);
  };
  // one time step: the fused kernel(s) of the enabled physics
  auto enqueue = [&]() -> int {
    if (mode == M_ADJ) {
      HIPCHK(sh ? launch_adj_sh(a, c->tile, c->stream) : launch_adj_psv(a, c->tile, c->stream));
      ++c->launches;
      return 0;
    }
    if (sh) {
      HIPCHK(launch_step_sh(a_sh, c->tile, c->stream, aux));
      ++c->launches;
    }
    if (psv) {
      HIPCHK(launch_step_psv(a, c->tile, spin, c->stream, aux));
      ++c->launches;
    }
    return 0;
  };
  auto count_step = [&](int n) {
    if (mode != M_ADJ) c->cur = (c->cur + n) & 1;
    if (mode == M_FWD) c->steps_done += n;
  };
  if (!sh && !psv) {
    // nothing to integrate; keep the clock in step with the host
    for (int s = 0; s < nsteps; ++s) {
      const int v[2] = {int(c->steps_done + 1), 0};
      HIPCHK(hipMemcpyAsync(c->clock.p, v, sizeof v, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      count_step(1);
    }
    return 0;
  }
  int done = 0;
  // steps per captured block: a whole number of launch units that returns
  // the state to the buffer it started in, since a block is replayed back to
  // back (multi-step launches: 2K steps; one step otherwise)
  const int G = multi ? c->graph_steps / (2 * K) * (2 * K) : c->graph_steps;   // graph_steps: a multiple of 4
  if (G >= 2 && nsteps >= G && !c->graphs_off && !getenv("SEISWG_NO_GRAPHS")) {
    // everything the captured launches depend on: the kernels (physics,
    // mode, tile shapes), the block structure and the argument blocks
    GraphKey k;
    const int32_t shape[13] = {int32_t(mode), int32_t(sh), int32_t(psv), int32_t(spin), int32_t(multi), G,
                               c->tile.nt_threads * 8 + c->tile.group, c->tile2.nt_threads * 8 + c->tile2.group,
                               int32_t(c->sh_only), K, int32_t(rec_kernel), int32_t(rec_loop), c->nrec2_ent};
    k.add(shape, sizeof shape);
    if (multi) {
      k.add(&a2a, sizeof a2a);
      k.add(&a2b, sizeof a2b);
    } else {
      k.add(&a, sizeof a);
      if (sh) k.add(&a_sh, sizeof a_sh);
    }
    hipGraphExec_t exec = nullptr;
    for (auto& e : c->graphs)
      if (e.key == k.h && e.bytes == k.bytes) {
        exec = e.exec;
        e.used = ++c->graph_clock;
      }
    if (!exec) {
      // Capture G steps.  Should the capture fail (HIP can invalidate a
      // capture because of work other host threads submit meanwhile), nothing
      // of it ran: this context stops using graphs, says so once on stderr
      // (sw_get_graph_state reports it) and launches directly.
      if (c->graphs.size() >= kMaxGraphs) {
        // evict the least recently used block; it may still be running
        HIPCHK(hipStreamSynchronize(c->stream));
        auto lru = std::min_element(c->graphs.begin(), c->graphs.end(),
                                    [](const sw_ctx::GraphEntry& x, const sw_ctx::GraphEntry& y) { return x.used < y.used; });
        (void)hipGraphExecDestroy(lru->exec);
        (void)hipGraphDestroy(lru->graph);
        c->graphs.erase(lru);
      }
      hipGraph_t graph = nullptr;
      const int64_t l0 = c->launches;
      hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
      int rc = 0;
      if (e == hipSuccess) {
        if (multi)
          for (int q = 0; q < G && !rc; q += U) rc = enqueue_unit(((q / U) & 1) != 0);
        else
          for (int q = 0; q < G && !rc; ++q) rc = enqueue();
        const hipError_t e_end = hipStreamEndCapture(c->stream, &graph);
        if (e_end != hipSuccess) e = e_end;
        else if (!graph) e = hipErrorUnknown;
        if (e == hipSuccess && rc == 0) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      }
      c->launches = l0;   // captured (or discarded), not launched
      if (e == hipSuccess && rc == 0) {
        c->graphs.push_back({k.h, std::move(k.bytes), graph, exec, ++c->graph_clock});
      } else {
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        c->graphs_off = true;
        c->graphs_failed = true;
        c->graph_error = rc ? g_err : std::string(hipGetErrorString(e));
        (void)hipGetLastError();
        fprintf(stderr, "seiswg: HIP graph capture failed (%s); this context launches its steps directly from now on\n",
                c->graph_error.c_str());
      }
    }
    const int64_t per_block = multi ? G / K : (mode != M_ADJ && sh && psv) ? 2 * G : G;
    for (; exec && done + G <= nsteps; done += G) {
      HIPCHK(hipGraphLaunch(exec, c->stream));
      c->launches += per_block;
      count_step(G);
    }
  }
  if (multi)
    for (; done + U <= nsteps; done += U) {
      // from the parity the (possibly replayed) blocks left
      int rc = enqueue_unit(c->cur != cur0);
      if (rc) return rc;
      count_step(U);
    }
  for (; done < nsteps; ++done) {
    int rc = enqueue();
    if (rc) return rc;
    count_step(1);
  }
  return 0;
}

float* slot_ptr(sw_ctx* c, int slot, int shot) {
  const int si = c->slot_index[slot];
  if (si < 0) return nullptr;
  return c->state[c->cur].p + (size_t(shot) * c->nslots + si) * size_t(c->npt);
}

int field_slot(int field) {
  switch (field) {
    case SW_F_VY: return S_VY;
    case SW_F_SXY: return S_SXY;
    case SW_F_SZY: return S_SZY;
    case SW_F_VX: return S_VX;
    case SW_F_VZ: return S_VZ;
    case SW_F_SXX: return S_SXX;
    case SW_F_SZZ: return S_SZZ;
    case SW_F_SXZ: return S_SXZ;
    case SW_F_VY_C: return S_VYC;
    case SW_F_SYX_C: return S_SYXC;
    case SW_F_SYY_C: return S_SYYC;
    case SW_F_SYZ_C: return S_SYZC;
  }
  return -1;
}

float* model_ptr(sw_ctx* c, int field) {
  switch (field) {
    case SW_F_LAM: return c->mf(MF_LAM);
    case SW_F_MU: return c->mf(MF_MU);
    case SW_F_NU: return c->mf(MF_NU);
    case SW_F_J: return c->mf(MF_J);
    case SW_F_LAM_C: return c->mf(MF_LAM_C);
    case SW_F_MU_C: return c->mf(MF_MU_C);
    case SW_F_NU_C: return c->mf(MF_NU_C);
    case SW_F_RHO: return c->mf(MF_RHO);
    case SW_F_K_MU: return c->kmu.p;
    case SW_F_K_LAM: return c->k_lam.p;
    case SW_F_K_RHO: return c->k_rho.p;
  }
  return nullptr;
}

// The forward time loop with snapshots (save_snapshot, solver.rs:1035,
// 1144-1176: after step it when it > 0 and it % snap == 0).  At a snapshot
// the fields are copied on the device into a staging buffer, which a copy
// stream moves to pinned host memory while the next steps run; the callbacks
// of a snapshot are made once the following block of steps has been queued
// (or at the end), in snapshot order.
int run_snapshots_body(sw_ctx* c, int snap, sw_snapshot_fn cb, void* user);
int run_with_snapshots(sw_ctx* c, int snap, sw_snapshot_fn cb, void* user) {
  const int rc = run_snapshots_body(c, snap, cb, user);
  // on an error, no transfer into the pinned buffers may still be running
  if (rc && c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  return rc;
}
// Tables of the separable Gaussian (smooth_gradient, solver.rs:1249-1274):
// w[d] = exp(-d*d / (2*sigma*sigma)) in f32 in the reference's expression
// order for d < max(nx, nz), and the 1-D weight sums sx[i], sz[j] whose
// product is init_gsum's GSUM.  Returns nw = 1 + the last d with w[d] != 0:
// the sums skip the zero terms, which leaves them unchanged.
int smooth_tables(int nx, int nz, float sigma, float* w, float* sx, float* sz) {
  const int nmax = std::max(nx, nz);
  int nw = 0;
  for (int d = 0; d < nmax; ++d) {
    const float fd = float(d);
    w[d] = std::exp(-fd * fd / (2.0f * sigma * sigma));
    if (w[d] != 0.0f) nw = d + 1;
  }
  for (int i = 0; i < nx; ++i) {
    float s = 0.0f;
    for (int n = std::max(i - nw + 1, 0); n < std::min(i + nw, nx); ++n) s = s + w[std::abs(i - n)];
    sx[i] = s;
  }
  for (int j = 0; j < nz; ++j) {
    float s = 0.0f;
    for (int n = std::max(j - nw + 1, 0); n < std::min(j + nw, nz); ++n) s = s + w[std::abs(j - n)];
    sz[j] = s;
  }
  return nw;
}

constexpr int kNoBlowup = 0x7fffffff;
const int kNoBlowupHost = kNoBlowup;   // static storage: a source that outlives any copy

int blowup_error(int step, int prev) {
  char earlier[96] = "";
  if (prev >= 0) snprintf(earlier, sizeof earlier, " (the check after step %d found none)", prev);
  return fail(-5,
              "wavefield not finite (NaN or infinite values) after step %d%s: the explicit scheme diverged (time "
              "step above the stability limit of this model) or an input is not finite",
              step, earlier);
}

// enqueue the scan of every shot's current state for the check tagged `step`
int enqueue_state_check(sw_ctx* c, int step) {
  HIPCHK(launch_nonfinite(c->state[c->cur].p, int64_t(c->nshots) * c->nslots * c->npt, step, c->dbad.p, c->stream));
  return 0;
}

// synchronous check of n device values (gradients, kernels)
int check_finite(sw_ctx* c, const float* p, int64_t n, const char* what) {
  if (!c->finite_check || n <= 0) return 0;
  HIPCHK(c->dbad.ensure(1));
  HIPCHK(hipMemcpyAsync(c->dbad.p, &kNoBlowupHost, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_nonfinite(p, n, 0, c->dbad.p, c->stream));
  int bad = kNoBlowup;
  HIPCHK(hipMemcpyAsync(&bad, c->dbad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (bad != kNoBlowup)
    return fail(-5, "%s not finite (NaN or infinite values): the forward wavefield diverged (time step above the "
                    "stability limit of this model) or an input is not finite", what);
  return 0;
}

int run_snapshots_body(sw_ctx* c, int snap, sw_snapshot_fn cb, void* user) {
  const int nt = c->g.nt;
  const bool check = c->finite_check && nt > 0;
  if (check) {
    HIPCHK(c->dbad.ensure(1));
    HIPCHK(hipMemcpyAsync(c->dbad.p, &kNoBlowupHost, sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (!c->bad_host) {
      HIPCHK(hipHostMalloc((void**)&c->bad_host, 2 * sizeof(int), hipHostMallocDefault));
      c->bad_host[0] = c->bad_host[1] = kNoBlowup;
    }
  }
  int last_clean = -1;   // step of the last check that found the state finite
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  const int64_t l0 = c->launches;
  const int fields[4] = {SW_F_VY, SW_F_VX, SW_F_VZ, SW_F_VY_C};
  const bool on[4] = {uses_sh(c), uses_psv(c), uses_psv(c), uses_spin(c)};
  int nf = 0;
  for (int q = 0; q < 4; ++q) nf += on[q] ? 1 : 0;
  const size_t per_field = size_t(c->npt);
  const size_t snap_floats = size_t(c->nshots) * nf * per_field;
  const bool snaps = snap > 0 && cb != nullptr && snap_floats > 0;
  if (snaps) {
    if (!c->copy_stream) {
      HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
      for (int q = 0; q < 2; ++q) {
        HIPCHK(hipEventCreateWithFlags(&c->snap_ready[q], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->snap_copied[q], hipEventDisableTiming));
      }
    }
    if (c->snap_host_bytes < snap_floats * 4) {
      for (int q = 0; q < 2; ++q) {
        if (c->snap_host[q]) (void)hipHostFree(c->snap_host[q]);
        c->snap_host[q] = nullptr;
      }
      c->snap_host_bytes = 0;
      for (int q = 0; q < 2; ++q) HIPCHK(hipHostMalloc((void**)&c->snap_host[q], snap_floats * 4, hipHostMallocDefault));
      c->snap_host_bytes = snap_floats * 4;
    }
    for (int q = 0; q < 2; ++q) {
      HIPCHK(c->snap_stage[q].ensure(snap_floats));
      HIPCHK(hipEventRecord(c->snap_copied[q], c->copy_stream));   // both slots free
    }
  }
  int pending = -1, pending_it = 0, slot = 0;
  auto deliver = [&](int q, int step) -> int {
    HIPCHK(hipEventSynchronize(c->snap_copied[q]));
    if (check) {
      if (c->bad_host[q] != kNoBlowup) return blowup_error(c->bad_host[q], last_clean);
      last_clean = step;
    }
    const float* h = c->snap_host[q];
    for (int s = 0; s < c->nshots; ++s)
      for (int f = 0; f < 4; ++f) {
        if (!on[f]) continue;
        if (cb(user, s, step, fields[f], h, c->npt) != 0) return fail(-4, "snapshot callback failed");
        h += per_field;
      }
    return 0;
  };
  int it = 0;
  while (it < nt) {
    int stop = nt - 1;
    if (snap > 0) {
      int next = ((it + snap - 1) / snap) * snap;   // first multiple of snap >= it
      if (next == 0) next = snap;
      if (next < nt) stop = next;
    }
    int rc = launch_steps(c, stop - it + 1);
    if (rc) return rc;
    if (snaps && stop > 0 && stop % snap == 0) {
      // staging slot: its previous transfer must be done (it was delivered)
      HIPCHK(hipStreamWaitEvent(c->stream, c->snap_copied[slot], 0));
      if (check) {
        // the scan's result travels with the snapshot (read when it is delivered)
        if ((rc = enqueue_state_check(c, stop))) return rc;
        HIPCHK(hipMemcpyAsync(&c->bad_host[slot], c->dbad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      }
      float* d = c->snap_stage[slot].p;
      for (int s = 0; s < c->nshots; ++s)
        for (int f = 0; f < 4; ++f) {
          if (!on[f]) continue;
          HIPCHK(hipMemcpyAsync(d, slot_ptr(c, field_slot(fields[f]), s), per_field * 4, hipMemcpyDeviceToDevice,
                                c->stream));
          d += per_field;
        }
      HIPCHK(hipEventRecord(c->snap_ready[slot], c->stream));
      HIPCHK(hipStreamWaitEvent(c->copy_stream, c->snap_ready[slot], 0));
      HIPCHK(hipMemcpyAsync(c->snap_host[slot], c->snap_stage[slot].p, snap_floats * 4, hipMemcpyDeviceToHost,
                            c->copy_stream));
      HIPCHK(hipEventRecord(c->snap_copied[slot], c->copy_stream));
      if (pending >= 0 && (rc = deliver(pending, pending_it))) return rc;
      pending = slot;
      pending_it = stop;
      slot ^= 1;
    }
    it = stop + 1;
  }
  if (pending >= 0) {
    int rc = deliver(pending, pending_it);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->timing.loop_ms = ms;
  c->timing.steps = nt;
  c->timing.step_launches = c->launches - l0;
  c->timing.cell_updates = int64_t(c->npt) * nt * c->nshots;
  if (check) {
    // the state after the last step (outside the timed loop)
    int rc = enqueue_state_check(c, nt - 1);
    if (rc) return rc;
    int bad = kNoBlowup;
    HIPCHK(hipMemcpyAsync(&bad, c->dbad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (bad != kNoBlowup) return blowup_error(bad, last_clean);
  }
  return 0;
}

}  // namespace

extern "C" {

SW_API const char* sw_last_error(void) { return g_err.c_str(); }

SW_API int sw_create(const sw_grid* g, int32_t device, sw_ctx** out) {
  if (!g || !out) return fail(-1, "null argument");
  if (g->nx < 4 || g->nz < 4) return fail(-2, "grid must be at least 4x4 (got %dx%d)", g->nx, g->nz);
  if (g->nt < 0) return fail(-2, "nt must be >= 0");
  // byte offsets inside one shot's state ([slot][npt]) and the model buffer are 32-bit
  if (int64_t(g->nx) * g->nz * int64_t(N_SLOTS) * 4 >= (int64_t(1) << 32))
    return fail(-2, "grid too large (%d x %d): at most %lld cells", g->nx, g->nz,
                (long long)(((int64_t(1) << 32) - 1) / (4 * N_SLOTS)));
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(-2, "no HIP device %d (found %d)", device, ndev);
  HIPCHK(hipSetDevice(device));
  sw_ctx* c = new sw_ctx;
  c->g = *g;
  c->device = device;
  c->npt = int64_t(g->nx) * g->nz;
  c->h8x = 8.0f * g->dx;
  c->h8z = 8.0f * g->dz;
  c->cdx = const_div(c->h8x);
  c->cdz = const_div(c->h8z);
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) c->ncu = 0;
  compute_slots(c);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    sw_destroy(c);
    return fail(-3, "stream/event creation failed: %s", hipGetErrorString(e));
  }
  *out = c;
  return 0;
}

SW_API void sw_destroy(sw_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& e : c->graphs) {
    (void)hipGraphExecDestroy(e.exec);
    (void)hipGraphDestroy(e.graph);
  }
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  for (int q = 0; q < 2; ++q) {
    if (c->snap_host[q]) (void)hipHostFree(c->snap_host[q]);
    if (c->snap_ready[q]) (void)hipEventDestroy(c->snap_ready[q]);
    if (c->snap_copied[q]) (void)hipEventDestroy(c->snap_copied[q]);
  }
  if (c->bad_host) (void)hipHostFree(c->bad_host);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

SW_API int sw_set_model(sw_ctx* c, int32_t kind, const float* const* f) {
  if (!c || !f) return fail(-1, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const size_t n = size_t(c->npt), bytes = n * 4;
  HIPCHK(c->model.ensure(size_t(MF_N) * n));
  HIPCHK(launch_fill(c->model.p, 0.0f, int64_t(MF_N) * int64_t(n), c->stream));
  auto up = [&](int mf, const float* h) -> int {
    HIPCHK(hipMemcpyAsync(c->mf(mf), h, bytes, hipMemcpyHostToDevice, c->stream));
    return 0;
  };
  int rc = 0;
  if (kind == SW_MODEL_VPVS) {
    if (!f[0] || !f[1] || !f[2]) return fail(-1, "rho, vp and vs are required");
    c->rho_fast = all_fast_divisors(f[0], n);
    c->j_fast = false;   // J = 0: the spin fields need the LAME model kind
    if ((rc = up(MF_RHO, f[0])) || (rc = up(MF_LAM, f[1])) || (rc = up(MF_MU, f[2]))) return rc;
    HIPCHK(launch_vps2lm(c->mf(MF_RHO), c->mf(MF_LAM), c->mf(MF_MU), c->npt, c->stream));
  } else if (kind == SW_MODEL_LAME) {
    const int order[8] = {MF_RHO, MF_LAM, MF_MU, MF_NU, MF_J, MF_LAM_C, MF_MU_C, MF_NU_C};
    for (int q = 0; q < 8; ++q) {
      if (!f[q]) return fail(-1, "model field %d missing", q);
      if ((rc = up(order[q], f[q]))) return rc;
    }
    c->rho_fast = all_fast_divisors(f[0], n);
    c->j_fast = all_fast_divisors(f[4], n);
  } else {
    return fail(-2, "unknown model kind %d", kind);
  }
  // correctly rounded reciprocals for the exact per-cell divisions
  HIPCHK(launch_recip(c->mf(MF_RHO), c->mf(MF_IRHO), int64_t(n), c->stream));
  if (kind == SW_MODEL_LAME) HIPCHK(launch_recip(c->mf(MF_J), c->mf(MF_IJ), int64_t(n), c->stream));
  else HIPCHK(launch_fill(c->mf(MF_IJ), 1.0f, int64_t(n), c->stream));
  build_taper(c);
  HIPCHK(c->bx.ensure(size_t(c->g.nx)));
  HIPCHK(c->bzb.ensure(size_t(c->g.nz)));
  HIPCHK(c->bzt.ensure(size_t(c->g.nz)));
  HIPCHK(hipMemcpyAsync(c->bx.p, c->h_bx.data(), c->h_bx.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->bzb.p, c->h_bzb.data(), c->h_bzb.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->bzt.p, c->h_bzt.data(), c->h_bzt.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->model_set = true;
  return 0;
}

SW_API int sw_set_sources(sw_ctx* c, int32_t nsrc, const int32_t* idx, const float* stf) {
  if (!c || nsrc < 0 || (nsrc > 0 && (!idx || !stf))) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(c->device));
  // a prepared batch points at the old source table and lists: it must be
  // prepared again (sw_prepare / the next run) before it can advance
  c->nshots = 0;
  c->nsrc = nsrc;
  c->src_k.assign(idx, idx + nsrc);
  const size_t n = size_t(3) * nsrc * size_t(c->g.nt);
  HIPCHK(c->stf.ensure(std::max<size_t>(n, 1)));
  if (n) {
    // on the context's stream: a copy on the legacy stream would have to wait
    // for (and is refused while) other threads capture graphs
    HIPCHK(hipMemcpyAsync(c->stf.p, stf, n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

SW_API int sw_set_receivers(sw_ctx* c, int32_t nrec, const int32_t* idx) {
  if (!c || nrec < 0 || (nrec > 0 && !idx)) return fail(-1, "bad arguments");
  // traces, displacement accumulators and per-tile receiver lists of a
  // prepared batch have the old receiver count: prepare again before advancing
  c->nshots = 0;
  c->nrec = nrec;
  c->rec_k.assign(idx, idx + nrec);
  return 0;
}

SW_API int sw_prepare(sw_ctx* c, int32_t nshots, const int32_t* isrc) {
  if (!c || nshots < 1 || !isrc) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(c->device));
  std::vector<std::vector<Entry>> inj(nshots);
  c->shot_isrc.assign(isrc, isrc + nshots);
  for (int s = 0; s < nshots; ++s) {
    if (isrc[s] >= c->nsrc) return fail(-2, "source %d out of range (nsrc=%d)", isrc[s], c->nsrc);
    inj[s] = source_entries(c, isrc[s]);
  }
  const size_t sstride = size_t(c->nsrc) * c->g.nt;
  c->val_x = c->stf.p;
  c->val_y = c->stf.p + sstride;
  c->val_z = c->stf.p + 2 * sstride;
  c->nt_val = c->g.nt;
  return prepare_batch(c, nshots, inj, true);
}

SW_API int sw_advance(sw_ctx* c, int32_t nsteps) {
  if (!c || nsteps < 0) return fail(-1, "bad arguments");
  if (c->nshots < 1) return fail(-2, "no prepared batch");
  if (c->steps_done + nsteps > c->g.nt)
    return fail(-2, "cannot advance %d steps: %lld of nt=%d done", nsteps, (long long)c->steps_done, c->g.nt);
  HIPCHK(hipSetDevice(c->device));
  const int64_t l0 = c->launches;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int rc = launch_steps(c, nsteps);
  if (rc) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->timing.loop_ms = ms;
  c->timing.steps = nsteps;
  c->timing.step_launches = c->launches - l0;
  c->timing.cell_updates = int64_t(c->npt) * nsteps * c->nshots;
  return 0;
}

SW_API int sw_forward_batch(sw_ctx* c, int32_t nshots, const int32_t* isrc, int32_t snap, sw_snapshot_fn cb,
                            void* user) {
  int rc = sw_prepare(c, nshots, isrc);
  if (rc) return rc;
  return run_with_snapshots(c, snap, cb, user);
}

SW_API int sw_forward(sw_ctx* c, int32_t isrc, int32_t snap, sw_snapshot_fn cb, void* user) {
  return sw_forward_batch(c, 1, &isrc, snap, cb, user);
}

SW_API int sw_read_obs(sw_ctx* c, int32_t shot, int32_t comp, float* out) {
  if (!c || !out || comp < 0 || comp > 4) return fail(-1, "bad arguments");
  if (shot < 0 || shot >= c->nshots) return fail(-2, "shot %d not in last run", shot);
  HIPCHK(hipSetDevice(c->device));
  const size_t n = size_t(c->nrec) * c->g.nt;
  if (n == 0) return 0;
  HIPCHK(hipMemcpyAsync(out, c->obs.p + (size_t(shot) * 5 + comp) * n, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

SW_API int sw_read_field(sw_ctx* c, int32_t shot, int32_t field, float* out) {
  if (!c || !out) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = size_t(c->npt) * 4;
  if (field == SW_F_BOUND) {
    if (!c->model_set) return fail(-2, "model not set");
    for (int i = 0; i < c->g.nx; ++i)
      for (int j = 0; j < c->g.nz; ++j)
        out[size_t(i) * c->g.nz + j] = (c->h_bx[i] * c->h_bzb[j]) * c->h_bzt[j];
    return 0;
  }
  const int slot = field_slot(field);
  const float* src = nullptr;
  if (slot >= 0) {
    if (shot < 0 || shot >= c->nshots) return fail(-2, "shot %d not prepared", shot);
    src = slot_ptr(c, slot, shot);
  } else {
    src = model_ptr(c, field);
  }
  if (!src) return fail(-2, "field %d is not materialised for this configuration", field);
  HIPCHK(hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

SW_API int sw_write_field(sw_ctx* c, int32_t shot, int32_t field, const float* in) {
  if (!c || !in) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(c->device));
  const int slot = field_slot(field);
  float* dst = nullptr;
  if (slot >= 0) {
    if (shot < 0 || shot >= c->nshots) return fail(-2, "shot %d not prepared", shot);
    dst = slot_ptr(c, slot, shot);
  } else {
    if (field == SW_F_K_MU) HIPCHK(c->kmu.ensure(size_t(c->npt)));
    if (field == SW_F_K_LAM) HIPCHK(c->k_lam.ensure(size_t(c->npt)));
    if (field == SW_F_K_RHO) HIPCHK(c->k_rho.ensure(size_t(c->npt)));
    dst = model_ptr(c, field);
  }
  if (!dst) return fail(-2, "field %d is not materialised for this configuration", field);
  HIPCHK(hipMemcpyAsync(dst, in, size_t(c->npt) * 4, hipMemcpyHostToDevice, c->stream));
  if (field == SW_F_RHO) {
    HIPCHK(launch_recip(c->mf(MF_RHO), c->mf(MF_IRHO), c->npt, c->stream));
    c->rho_fast = all_fast_divisors(in, size_t(c->npt));
  }
  if (field == SW_F_J) {
    HIPCHK(launch_recip(c->mf(MF_J), c->mf(MF_IJ), c->npt, c->stream));
    c->j_fast = all_fast_divisors(in, size_t(c->npt));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

SW_API int sw_set_steps_per_launch(sw_ctx* c, int32_t steps) {
  if (!c || steps < 1 || steps > 3) return fail(-1, "steps per launch must be 1, 2 or 3");
  c->steps_per_launch = steps;
  if (steps == 1) c->two_ok = false;
  return 0;
}

SW_API int sw_set_tiling(sw_ctx* c, int32_t threads, int32_t seglen, int32_t shots_per_block) {
  if (!c) return fail(-1, "null ctx");
  if (!tiling_supported(threads)) return fail(-2, "threads must be 64, 128 or 256");
  if (seglen < 1) return fail(-2, "seglen must be >= 1");
  if (shots_per_block != 0 && !group_supported(threads, shots_per_block))
    return fail(-2, "shots_per_block must be 0 (automatic), 1, or 2 or 4 with threads >= 128");
  c->tile.nt_threads = threads;
  c->tile.seglen = seglen;
  c->group_req = shots_per_block;
  c->tile_fixed = true;
  c->nshots = 0;   // lists must be rebuilt
  return 0;
}

SW_API int sw_autotune(sw_ctx* c, int32_t nshots, int32_t steps, int32_t* threads, int32_t* seglen,
                       int32_t* shots_per_block, double* ms) {
  if (!c || nshots < 1 || steps < 1) return fail(-1, "bad arguments");
  if (!c->model_set) return fail(-2, "model not set");
  HIPCHK(hipSetDevice(c->device));
  std::vector<int32_t> isrc(nshots);
  for (int s = 0; s < nshots; ++s) isrc[s] = c->nsrc > 0 ? s % c->nsrc : -1;
  const int spl = c->steps_per_launch;
  // two-step shapes: marching tiles for SH or P-SV alone, LDS tiles also with spin
  const bool two = spl >= 2 && uses_psv(c) != uses_sh(c);
  const bool three = three_steps(c);
  // whole launch units of every multi-step shape (pairs of two-step
  // launches; LDS tiles: 2K steps, so 24 steps for K = 2, 3, 4) when nt
  // leaves room, else whole pairs
  int n2 = std::max(4, (steps + 3) / 4 * 4);
  if (two && c->g.nt >= (steps + 23) / 24 * 24 + 4) n2 = std::max(24, (steps + 23) / 24 * 24);
  if (c->g.nt < steps + 2 || (two && c->g.nt < n2 + 4))
    return fail(-2, "nt=%d too short to time %d steps", c->g.nt, two ? n2 : steps);
  // times `n` steps (after a warm-up) of the prepared batch, in ms
  auto time_steps = [&](int warm, int n, float& t) -> int {
    int rc = sw_prepare(c, nshots, isrc.data());
    if (rc) return rc;
    if ((rc = launch_steps(c, warm))) return rc;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if ((rc = launch_steps(c, n))) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&t, c->ev0, c->ev1));
    return 0;
  };
  const int segs[] = {8, 16, 32, 48, 64, 128, 256, 400, 500, 512};
  // one-step launches: timed and tuned only where they run the forward
  // steps (physics without multi-step kernels, one step per launch).  Where
  // multi-step launches run the forward, the one-step shape serves remainder
  // steps and the gradient's adjoint, whose fastest shape the forward's
  // timing does not find (16-shot 1000x600 gradient: 0.21 s with the
  // heuristic shape, 0.24 s with the one-step forward's autotuned one), so
  // it stays the heuristic's, chosen again for each prepared batch.
  c->steps_per_launch = 1;
  double best = 1e300;
  Tiling bt = c->tile;
  int bg = c->group_req;
  if (two) {
    c->steps_per_launch = spl;
    c->tile_fixed = false;
    c->group_req = 0;
    heuristic_tiling(c, nshots);
    bt = c->tile;
  }
  for (int nt : {64, 128, 256}) {
    if (two) break;
    for (int grp : {1, 2, 4}) {
      if (!group_supported(nt, grp) || nshots % grp != 0) continue;
      for (int seg : segs) {
        if (seg > 2 * c->g.nx) continue;
        c->tile.nt_threads = nt;
        c->tile.seglen = seg;
        c->group_req = grp;
        c->tile_fixed = true;
        float t = 0;
        const int rc = time_steps(2, steps, t);
        if (rc) {
          c->steps_per_launch = spl;
          return rc;
        }
        if (t < best) {
          best = t;
          bt = c->tile;
          bg = grp;
        }
      }
    }
  }
  c->steps_per_launch = spl;
  double best_ms = best / steps;
  if (!two) {
    c->tile = bt;
    c->group_req = bg;
    c->tile_fixed = true;
  }
  if (two) {
    // two-step launches: their own widths (up to 512), groups and segments
    double best2 = 1e300;
    Tiling bt2 = c->tile2;
    int bg2 = c->group2_req;
    for (int nt : {64, 128, 256, 320, 384, 512, kLdsTileThreads}) {
      if (uses_spin(c) && nt != kLdsTileThreads) continue;
      // three steps per launch: marching widths 256, 384 and 512, one shot
      // per workgroup (LDS tiles as for two)
      if (three ? (nt != kLdsTileThreads && !tiling3_supported(nt)) : (nt == 384)) continue;
      for (int grp : {1, 2, 4}) {
        if (three && nt != kLdsTileThreads ? grp != 1 : (!group2_supported(nt, grp) || nshots % grp != 0)) continue;
        // besides the fixed lengths: segments that make 1-4 whole rounds of
        // resident workgroups (equal lengths; partial rounds leave CUs idle)
        std::vector<int> cand(std::begin(segs), std::end(segs));
        if (nt != kLdsTileThreads) {
          const int halo = three ? 9 : 6;
          const int64_t res = three ? step3_blocks_per_cu(nt) : step2_blocks_per_cu(nt, grp);
          const int64_t base = int64_t(nshots / grp) * ((c->g.nz + nt - 2 * halo - 1) / (nt - 2 * halo));
          const int64_t ncu = c->ncu > 0 ? c->ncu : 256;
          for (int k = 1; k <= 4 && res > 0; ++k) {
            const int64_t ns = k * ncu * res / base;
            if (ns >= 1 && (k * ncu * res) % base == 0) cand.push_back(int((c->g.nx + ns - 1) / ns));
          }
        }
        for (int seg : cand) {
          if (seg > 2 * c->g.nx) continue;
          // LDS tiles: the kernel fixes its tile; steps per launch 2..4 (the
          // first segment lengths stand for K = 2, 3, 4)
          const int lk = nt != kLdsTileThreads ? 0 : seg == segs[0] ? 2 : seg == segs[1] ? 3 : seg == segs[2] ? 4 : -1;
          if (lk < 0 || lk > lds_max_steps(c)) continue;
          c->tile2.nt_threads = nt;
          c->tile2.seglen = lk ? lds_tile_rows(lk) : seg;
          c->tile2.lds_steps = lk;
          c->tile2.group = grp;
          c->group2_req = grp;
          c->tile2_fixed = true;
          float t = 0;
          const int rc = time_steps(4, n2, t);
          if (rc) return rc;
          if (t < best2) {
            best2 = t;
            bt2 = c->tile2;
            bg2 = grp;
          }
        }
      }
    }
    c->tile2 = bt2;
    c->group2_req = bg2;
    c->tile2_fixed = true;
    best_ms = best2 / n2;
  }
  c->nshots = 0;
  if (threads) *threads = bt.nt_threads;
  if (seglen) *seglen = bt.seglen;
  if (shots_per_block) *shots_per_block = bt.group;
  if (ms) *ms = best_ms;
  return 0;
}

SW_API int sw_set_tiling2(sw_ctx* c, int32_t threads, int32_t seglen, int32_t shots_per_block) {
  if (!c) return fail(-1, "null ctx");
  if (!tiling2_supported(threads) && !tiling3_supported(threads))
    return fail(-2, "threads must be 64, 128, 256, 320, 384 or 512 (marching tiles) or %d (LDS tiles)",
                kLdsTileThreads);
  // LDS tiles: `seglen` is the steps per launch (2..4, with spin 2..3; 0:
  // the default); the kernel fixes its tile
  int lds_steps = 0;
  if (threads == kLdsTileThreads) {
    lds_steps = seglen == 0 ? lds_default_steps(c) : seglen;
    if (lds_steps < 2 || lds_steps > lds_max_steps(c))
      return fail(-2, "LDS tiles run 2 to %d steps per launch%s", lds_max_steps(c), uses_spin(c) ? " with spin" : "");
    seglen = lds_tile_rows(lds_steps);
  }
  if (seglen < 1) return fail(-2, "seglen must be >= 1");
  if (shots_per_block > 1 && !group2_supported(threads, shots_per_block))
    return fail(-2, "shots_per_block must be 0 (automatic), 1, or 2 or 4 with threads >= 128 and at most "
                    "1024 threads per workgroup");
  c->tile2.nt_threads = threads;
  c->tile2.seglen = seglen;
  c->tile2.lds_steps = lds_steps;
  c->tile2.group = shots_per_block > 0 ? shots_per_block : 1;
  c->group2_req = shots_per_block;
  c->tile2_fixed = true;
  c->nshots = 0;   // lists must be rebuilt
  return 0;
}
SW_API int sw_get_tiling2(const sw_ctx* c, int32_t* threads, int32_t* seglen, int32_t* shots_per_block) {
  if (!c || !threads || !seglen) return fail(-1, "null argument");
  *threads = c->two_ok ? c->tile2.nt_threads : 0;
  *seglen = !c->two_ok ? 0 : tile2_lds(c->tile2) ? c->tile2.lds_steps : c->tile2.seglen;
  if (shots_per_block) *shots_per_block = c->two_ok ? c->tile2.group : 0;
  return 0;
}
SW_API int sw_get_tiling(const sw_ctx* c, int32_t* threads, int32_t* seglen, int32_t* shots_per_block) {
  if (!c || !threads || !seglen) return fail(-1, "null argument");
  *threads = c->tile.nt_threads;
  *seglen = c->tile.seglen;
  if (shots_per_block) *shots_per_block = c->tile.group;
  return 0;
}

SW_API int sw_release_cached_memory(int32_t hip_device, int64_t* bytes_freed) {
  const size_t n = cache_trim(hip_device);
  if (bytes_freed) *bytes_freed = int64_t(n);
  return 0;
}

SW_API int sw_bench_copy(int32_t device, int64_t nbytes, int32_t iters, double* avg_ms) {
  if (nbytes <= 0 || nbytes % 16 || iters < 1 || !avg_ms) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(device));
  float *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipMalloc(&a, size_t(nbytes)));
  HIPCHK(hipMalloc(&b, size_t(nbytes)));
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(launch_fill(a, 1.0f, nbytes / 4, s));
  HIPCHK(launch_copy(a, b, nbytes / 4, s));
  HIPCHK(hipEventRecord(e0, s));
  for (int q = 0; q < iters; ++q) HIPCHK(launch_copy(a, b, nbytes / 4, s));
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *avg_ms = double(ms) / iters;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

SW_API int sw_get_timing(const sw_ctx* c, sw_timing* out) {
  if (!c || !out) return fail(-1, "null argument");
  *out = c->timing;
  return 0;
}

SW_API int sw_set_division(sw_ctx* c, int32_t mode) {
  if (!c || (mode != SW_DIV_EXACT && mode != SW_DIV_RECIPROCAL)) return fail(-1, "division mode must be 0 or 1");
  c->recip_div = mode == SW_DIV_RECIPROCAL;
  return 0;
}

SW_API int sw_set_finite_check(sw_ctx* c, int32_t on) {
  if (!c) return fail(-1, "null ctx");
  c->finite_check = on != 0;
  return 0;
}

SW_API int sw_get_graph_state(const sw_ctx* c, int32_t* state, int32_t* captured, char* msg, int32_t msg_len) {
  if (!c || !state) return fail(-1, "null argument");
  *state = c->graphs_failed ? -1 : (c->graphs_off || getenv("SEISWG_NO_GRAPHS")) ? 0 : 1;
  if (captured) *captured = int32_t(c->graphs.size());
  if (msg && msg_len > 0) snprintf(msg, size_t(msg_len), "%s", c->graph_error.c_str());
  return 0;
}

}  // extern "C"

namespace {

// Device scratch of sw_smooth_noalloc: the weight table, the two 1-D weight
// sums and the x-pass result, for grids of at most kNoallocCells cells with
// nx, nz <= kNoallocAxis (5,292,032 floats, 21.2 MB per GPU, held until the
// process ends).  Host code never passes `scratch` itself to a kernel or a
// copy: its name is a host-side stand-in, and each GPU's copy has its own
// address, looked up once per device with that device current.
constexpr int kNoallocAxis = 16384;
constexpr int64_t kNoallocCells = int64_t(5) << 20;
constexpr int64_t kNoallocFloats = 3 * int64_t(kNoallocAxis) + kNoallocCells;
static __device__ __attribute__((used)) __align__(256) float scratch[kNoallocFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
// host tables of the call in progress (static storage: no allocation); the
// mutex also serialises the host side of concurrent calls
float g_noalloc_tab[3 * kNoallocAxis];

// the address of `scratch` on `device` (current): looked up on the first call
int noalloc_scratch(int device, float** out) {
  if (device < 0 || device >= kMaxDevices) return fail(-2, "device %d beyond the %d the library tracks", device, kMaxDevices);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(-3, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[device];
  return 0;
}

}  // namespace

extern "C" {

SW_API int sw_smooth_noalloc(sw_ctx* c, int32_t field, float sigma) {
  if (!c) return fail(-1, "null ctx");
  if (field != SW_F_K_MU && field != SW_F_K_LAM && field != SW_F_K_RHO) return fail(-2, "only kernel fields smooth");
  const int nx = c->g.nx, nz = c->g.nz;
  if (nx > kNoallocAxis || nz > kNoallocAxis || int64_t(nx) * nz > kNoallocCells)
    return fail(-2, "sw_smooth_noalloc: grid %d x %d beyond its fixed scratch (at most %lld cells, %d per axis); "
                    "use sw_smooth", nx, nz, (long long)kNoallocCells, kNoallocAxis);
  float* k = model_ptr(c, field);
  if (!k) return fail(-2, "kernel field %d not computed", field);
  HIPCHK(hipSetDevice(c->device));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  float* scr = nullptr;
  int rc = noalloc_scratch(c->device, &scr);
  if (rc) return rc;
  float* d_w = scr;
  float* d_sx = scr + kNoallocAxis;
  float* d_sz = scr + 2 * kNoallocAxis;
  float* d_gtmp = scr + 3 * kNoallocAxis;
  float* h_w = g_noalloc_tab;
  float* h_sx = g_noalloc_tab + kNoallocAxis;
  float* h_sz = g_noalloc_tab + 2 * kNoallocAxis;
  const int nmax = std::max(nx, nz);
  const int nw = smooth_tables(nx, nz, sigma, h_w, h_sx, h_sz);
  HIPCHK(hipMemcpyAsync(d_w, h_w, size_t(nmax) * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_sx, h_sx, size_t(nx) * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_sz, h_sz, size_t(nz) * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_gauss_x(d_w, nw, k, d_gtmp, nx, nz, c->stream));
  HIPCHK(launch_gauss_z(d_w, nw, d_gtmp, d_sx, d_sz, k, nx, nz, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));   // the host tables are reused by the next call
  return 0;
}

SW_API int sw_smooth(sw_ctx* c, int32_t field, float sigma) {
  if (!c) return fail(-1, "null ctx");
  HIPCHK(hipSetDevice(c->device));
  float* k = model_ptr(c, field);
  if (field != SW_F_K_MU && field != SW_F_K_LAM && field != SW_F_K_RHO) return fail(-2, "only kernel fields smooth");
  if (!k) return fail(-2, "kernel field %d not computed", field);
  const int nx = c->g.nx, nz = c->g.nz, nmax = std::max(nx, nz);
  std::vector<float> w(nmax), sx(nx), sz(nz);
  const int nw = smooth_tables(nx, nz, sigma, w.data(), sx.data(), sz.data());
  HIPCHK(c->wtab.ensure(nmax));
  HIPCHK(c->sumx.ensure(nx));
  HIPCHK(c->sumz.ensure(nz));
  HIPCHK(c->gtmp.ensure(size_t(c->npt)));
  HIPCHK(hipMemcpyAsync(c->wtab.p, w.data(), w.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->sumx.p, sx.data(), sx.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->sumz.p, sz.data(), sz.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_gauss_x(c->wtab.p, nw, k, c->gtmp.p, nx, nz, c->stream));
  HIPCHK(launch_gauss_z(c->wtab.p, nw, c->gtmp.p, c->sumx.p, c->sumz.p, k, nx, nz, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}


}  // extern "C"

namespace {

// ---- phases of the checkpointed adjoint gradient, shared by the one-problem
// (sw_gradient) and the many-problem (sw_multi_gradient) drivers

int grad_begin(sw_ctx* c, int ckpt_every) {
  if (uses_spin(c) || (uses_sh(c) == uses_psv(c)))
    return fail(-2, "the gradient needs exactly one of sh / psv, without spin");
  if (ckpt_every < 1) return fail(-2, "ckpt_every must be >= 1");
  if (!c->model_set) return fail(-2, "model not set");
  const bool sh = uses_sh(c);
  const int nt = c->g.nt;
  GradPlan& P = c->gp;
  P.K = std::min(ckpt_every, std::max(nt, 1));
  P.nck = (nt + P.K - 1) / P.K;
  P.ncomp = sh ? 1 : 2;
  P.naux = sh ? NAUX_SH : NAUX_PSV;
  const size_t npt = size_t(c->npt);
  for (DBuf<float>* k : {&c->k_rho, &c->kmu, &c->k_lam}) {
    HIPCHK(k->ensure(npt));
    HIPCHK(launch_fill(k->p, 0.0f, int64_t(npt), c->stream));
  }
  // misfit time taper (solver.rs:441-455), f32 on the host
  c->h_taper.assign(std::max(nt, 1), 1.0f);
  const float PI_F = 3.14159265358979323846f, dt = c->g.dt;
  const float t_end = float(nt - 1) * dt, tw = t_end / 10.0f, t_min = tw, t_max = t_end - tw;
  for (int n = 0; n < nt; ++n) {
    const float t = float(n) * dt;
    c->h_taper[n] = t <= t_min ? 0.5f + 0.5f * std::cos(PI_F * (t_min - t) / tw)
                    : t >= t_max ? 0.5f + 0.5f * std::cos(PI_F * (t_max - t) / tw) : 1.0f;
  }
  HIPCHK(c->gtaper.ensure(c->h_taper.size()));
  HIPCHK(hipMemcpyAsync(c->gtaper.p, c->h_taper.data(), c->h_taper.size() * 4, hipMemcpyHostToDevice, c->stream));
  c->h_comps[0] = sh ? SW_OBS_Y : SW_OBS_X;
  c->h_comps[1] = SW_OBS_Z;
  HIPCHK(c->gcomps.ensure(2));
  HIPCHK(hipMemcpyAsync(c->gcomps.p, c->h_comps, sizeof c->h_comps, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// a batch of nb shots: forward buffers, source and receiver injection lists,
// checkpoint storage
int grad_batch_setup(sw_ctx* c, int nb, const int32_t* isrc) {
  GradPlan& P = c->gp;
  P.kgroup = 1;   // merged launches: one shot per workgroup
  const int nrec = c->nrec;
  std::vector<std::vector<Entry>> inj(nb), ainj(nb);
  for (int s = 0; s < nb; ++s) {
    if (isrc[s] >= c->nsrc) return fail(-2, "source %d out of range (nsrc=%d)", isrc[s], c->nsrc);
    inj[s] = source_entries(c, isrc[s]);
    for (int r = 0; r < nrec; ++r) {
      const uint32_t k = uint32_t(c->rec_k[r]);
      if (int64_t(k) >= c->npt) continue;
      ainj[s].push_back({int(k / uint32_t(c->g.nz)), int(k % uint32_t(c->g.nz)), s * nrec + r});
    }
  }
  const size_t sstride = size_t(c->nsrc) * c->g.nt;
  c->val_x = c->stf.p; c->val_y = c->stf.p + sstride; c->val_z = c->stf.p + 2 * sstride; c->nt_val = c->g.nt;
  int rc = prepare_batch(c, nb, inj, true);
  if (rc) return rc;
  if ((rc = upload_inj(c, ainj, c->ainj_off, c->ainj))) return rc;
  P.nb = nb;
  P.state_floats = size_t(nb) * c->nslots * size_t(c->npt);
  HIPCHK(c->ckpt.ensure(std::max<size_t>(size_t(P.nck) * P.state_floats, 1)));
  return 0;
}

int grad_save_ckpt(sw_ctx* c, int q) {
  if (q > 0)
    HIPCHK(hipMemcpyAsync(c->ckpt.p + size_t(q) * c->gp.state_floats, c->state[c->cur].p, c->gp.state_floats * 4,
                          hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// adjoint sources, per-shot misfits (host copy in c->gp.mis after a sync) and
// the backward-pass buffers
int grad_adjsrc_and_bwd_setup(sw_ctx* c, const float* obs_ref_batch) {
  GradPlan& P = c->gp;
  const int nb = P.nb, nrec = c->nrec, nt = c->g.nt;
  const size_t npt = size_t(c->npt);
  const size_t per_shot_obs = size_t(P.ncomp) * nrec * nt;
  HIPCHK(c->gref.ensure(std::max<size_t>(size_t(nb) * per_shot_obs, 1)));
  HIPCHK(hipMemcpyAsync(c->gref.p, obs_ref_batch, size_t(nb) * per_shot_obs * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c->adj_val.ensure(std::max<size_t>(size_t(P.ncomp) * nb * nrec * nt, 1)));
  HIPCHK(c->gmis.ensure(size_t(nb) + size_t(nb) * P.ncomp * std::max(nrec, 1)));
  HIPCHK(launch_adjsrc(c->obs.p, c->gref.p, c->gtaper.p, c->adj_val.p, c->gmis.p, nb, P.ncomp, c->gcomps.p, nrec, nt,
                       c->g.dt, c->stream));
  P.mis.assign(nb, 0.0);
  HIPCHK(hipMemcpyAsync(P.mis.data(), c->gmis.p, size_t(nb) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  // forward velocities of a recomputed segment: time levels n0+1..n0+K, slot n % K
  HIPCHK(c->aux.ensure(std::max<size_t>(size_t(P.K) * nb * P.naux * npt, 1)));
  HIPCHK(c->kshot.ensure(size_t(nb) * NK * npt));
  HIPCHK(launch_fill(c->kshot.p, 0.0f, int64_t(nb) * NK * int64_t(npt), c->stream));
  for (int p = 0; p < 2; ++p) {
    HIPCHK(c->adj[p].ensure(std::max<size_t>(P.state_floats, 1)));
    HIPCHK(launch_fill(c->adj[p].p, 0.0f, int64_t(P.state_floats), c->stream));
  }
  HIPCHK(c->aclock.ensure(2));
  c->aux_period = P.K;
  HIPCHK(hipStreamSynchronize(c->stream));   // obs_ref_batch and P.mis are host memory
  if (c->finite_check)
    for (int s = 0; s < nb; ++s)
      if (!std::isfinite(P.mis[s]))
        return fail(-5, "misfit of shot %d is not finite (%g): the forward wavefield diverged (time step above the "
                        "stability limit of this model) or an input is not finite", s, P.mis[s]);
  return 0;
}

// segment q of the backward pass: restore its checkpoint and set the clocks;
// returns the segment's step range [n0, n1) through the arguments
int grad_segment_setup(sw_ctx* c, int q, int& n0, int& n1) {
  GradPlan& P = c->gp;
  n0 = q * P.K;
  n1 = std::min(c->g.nt, n0 + P.K);
  c->cur = n0 & 1;
  if (q == 0) HIPCHK(launch_fill(c->state[c->cur].p, 0.0f, int64_t(P.state_floats), c->stream));
  else
    HIPCHK(hipMemcpyAsync(c->state[c->cur].p, c->ckpt.p + size_t(q) * P.state_floats, P.state_floats * 4,
                          hipMemcpyDeviceToDevice, c->stream));
  c->h_clk[0] = n0; c->h_clk[1] = 0; c->h_clk[2] = n1 - 1; c->h_clk[3] = 0;
  HIPCHK(hipMemcpyAsync(c->clock.p, c->h_clk, 2 * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->aclock.p, c->h_clk + 2, 2 * sizeof(int), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// kernel fields after a gradient
int check_grad_finite(sw_ctx* c) {
  int rc = 0;
  if ((rc = check_finite(c, c->k_rho.p, c->npt, "K_rho")) || (rc = check_finite(c, c->kmu.p, c->npt, "K_mu")) ||
      (rc = check_finite(c, c->k_lam.p, c->npt, "K_lambda")))
    return rc;
  return 0;
}

// sum the batch's kernel accumulators (one per kgroup shots) into the
// context's K fields (fixed order)
int grad_reduce(sw_ctx* c) {
  const size_t npt = size_t(c->npt);
  const int nb = c->gp.nb / c->gp.kgroup;
  HIPCHK(launch_reduce_shots(c->kshot.p + size_t(K_RHO) * npt, c->k_rho.p, nb, int64_t(npt), NK * int64_t(npt),
                             c->stream));
  HIPCHK(launch_reduce_shots(c->kshot.p + size_t(K_MU) * npt, c->kmu.p, nb, int64_t(npt), NK * int64_t(npt),
                             c->stream));
  HIPCHK(launch_reduce_shots(c->kshot.p + size_t(K_LAM) * npt, c->k_lam.p, nb, int64_t(npt), NK * int64_t(npt),
                             c->stream));
  return 0;
}

}  // namespace

extern "C" {

SW_API int sw_gradient(sw_ctx* c, int32_t nshots, const int32_t* isrc, const float* obs_ref, int32_t ckpt_every,
                       int32_t shots_per_launch, double* misfit) {
  if (!c || nshots < 1 || !isrc || !obs_ref) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(c->device));
  int rc = grad_begin(c, ckpt_every);
  if (rc) return rc;
  const GradPlan& P = c->gp;
  const int nt = c->g.nt;
  const int S = (shots_per_launch <= 0 || shots_per_launch > nshots) ? nshots : shots_per_launch;
  const size_t per_shot_obs = size_t(P.ncomp) * c->nrec * nt;
  for (int b0 = 0; b0 < nshots; b0 += S) {
    const int nb = std::min(S, nshots - b0);
    if ((rc = grad_batch_setup(c, nb, isrc + b0))) return rc;
    c->gp.kgroup = c->tile.group;
    // forward with checkpoints (traces recorded)
    for (int q = 0; q < P.nck; ++q) {
      if ((rc = grad_save_ckpt(c, q))) return rc;
      if ((rc = launch_steps(c, std::min(P.K, nt - q * P.K), M_FWD))) return rc;
    }
    if ((rc = grad_adjsrc_and_bwd_setup(c, obs_ref + size_t(b0) * per_shot_obs))) return rc;
    // backward: per segment, recompute the forward steps keeping aux, then the adjoint steps
    for (int q = P.nck - 1; q >= 0; --q) {
      int n0, n1;
      if ((rc = grad_segment_setup(c, q, n0, n1))) return rc;
      if ((rc = launch_steps(c, n1 - n0, M_FWD_AUX))) return rc;
      if ((rc = launch_steps(c, n1 - n0, M_ADJ))) return rc;
      HIPCHK(hipStreamSynchronize(c->stream));   // h_clk is reused by the next segment
    }
    if ((rc = grad_reduce(c))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (misfit)
      for (int s = 0; s < nb; ++s) misfit[b0 + s] = P.mis[s];
  }
  return check_grad_finite(c);
}

}  // extern "C"

struct sw_multi {
  std::vector<sw_ctx*> p;
  int device = 0;
  hipStream_t stream = nullptr;
  // gradients: problems with at least this many shot-cells per step run on
  // their own launches; the smaller ones share merged launches.  Default 0:
  // own launches for all (with three running at once they measured faster
  // than merged launches from 0.03 M shot-cells per step on: 8 equal
  // problems of 128x128 x 2 shots 0.207 s against 0.274 s)
  int64_t merge_limit = 0;
  // problems on their own launches run concurrently, each on its own host
  // thread and stream (their launches and host work overlap): this many
  // GPU-filling problems at a time, smaller ones by their share of that
  int workers = 3;
  DBuf<StepArgs> args;
  DBuf<int2> work;
  double loop_ms = 0;
  int64_t launches = 0;
};

namespace {

// Swaps every problem's stream for the shared one while a multi call runs, so
// all per-problem copies and the merged launches are ordered on one stream.
struct SharedStream {
  sw_multi* m;
  std::vector<hipStream_t> saved;
  explicit SharedStream(sw_multi* x) : m(x) {
    for (sw_ctx* c : m->p) { saved.push_back(c->stream); c->stream = m->stream; }
  }
  ~SharedStream() {
    for (size_t q = 0; q < m->p.size(); ++q) m->p[q]->stream = saved[q];
  }
};

// nlaunch merged launches of `mode`; problem q takes part while its clock is
// in [win[q].x, win[q].y).  One launch per (kernel kind, tile width) group;
// SH groups go before P-SV groups, as the reference dispatches them.
int multi_steps(sw_multi* m, StepMode mode, int nlaunch, const std::vector<int2>& win) {
  struct Ent { int kind, nt; StepArgs a; };
  std::vector<Ent> ents;
  const bool aux = mode == M_FWD_AUX;
  for (size_t q = 0; q < m->p.size(); ++q) {
    sw_ctx* c = m->p[q];
    if (win[q].x >= win[q].y) continue;
    StepArgs a = make_args(c, mode);
    a.nblocks = c->nshots * c->tile.nstrips * c->tile.nsegs;   // merged launches run one shot per workgroup
    a.it_lo = win[q].x;
    a.it_hi = win[q].y;
    const bool sh = step_sh_on(c), psv = step_psv_on(c), spin = uses_spin(c);
    const int nt = c->tile.nt_threads;
    if (mode == M_ADJ) {
      ents.push_back({sh ? MK_ADJ_SH : MK_ADJ_PSV, nt, a});
      continue;
    }
    if (sh) {
      StepArgs b = a;
      b.advance = psv ? 0 : 1;
      ents.push_back({aux ? MK_SH_AUX : MK_SH, nt, b});
    }
    if (psv) ents.push_back({aux ? MK_PSV_AUX : (spin ? MK_SPIN : MK_PSV), nt, a});
  }
  if (ents.empty() || nlaunch <= 0) return 0;
  auto rank = [](int kind) { return (kind == MK_SH || kind == MK_SH_AUX || kind == MK_ADJ_SH) ? 0 : 1; };
  std::stable_sort(ents.begin(), ents.end(), [&](const Ent& x, const Ent& y) {
    return rank(x.kind) != rank(y.kind) ? rank(x.kind) < rank(y.kind)
                                         : (x.kind != y.kind ? x.kind < y.kind : x.nt < y.nt);
  });
  std::vector<StepArgs> hargs;
  std::vector<int2> hwork;
  struct Grp { int kind, nt, off, nb; };
  std::vector<Grp> grps;
  for (size_t e = 0; e < ents.size(); ++e) {
    if (grps.empty() || grps.back().kind != ents[e].kind || grps.back().nt != ents[e].nt)
      grps.push_back({ents[e].kind, ents[e].nt, int(hwork.size()), 0});
    hargs.push_back(ents[e].a);
    for (int b = 0; b < ents[e].a.nblocks; ++b) hwork.push_back(make_int2(int(e), b));
    grps.back().nb += ents[e].a.nblocks;
  }
  HIPCHK(m->args.ensure(hargs.size()));
  HIPCHK(m->work.ensure(hwork.size()));
  HIPCHK(hipMemcpyAsync(m->args.p, hargs.data(), hargs.size() * sizeof(StepArgs), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(m->work.p, hwork.data(), hwork.size() * sizeof(int2), hipMemcpyHostToDevice, m->stream));
  for (int l = 0; l < nlaunch; ++l)
    for (const Grp& g : grps) {
      HIPCHK(launch_multi(g.kind, g.nt, m->args.p, m->work.p + g.off, g.nb, m->stream));
      ++m->launches;
    }
  HIPCHK(hipStreamSynchronize(m->stream));   // hargs / hwork are host memory
  return 0;
}

}  // namespace

extern "C" {

SW_API int sw_multi_create(sw_ctx* const* problems, int32_t n, sw_multi** out) {
  if (!problems || n < 1 || !out) return fail(-1, "bad arguments");
  for (int q = 0; q < n; ++q) {
    if (!problems[q]) return fail(-1, "problem %d is null", q);
    if (problems[q]->device != problems[0]->device) return fail(-2, "problems must share one HIP device");
    if (!problems[q]->model_set) return fail(-2, "problem %d has no model", q);
    for (int r = 0; r < q; ++r)
      if (problems[r] == problems[q]) return fail(-2, "problem %d appears twice", q);
  }
  HIPCHK(hipSetDevice(problems[0]->device));
  sw_multi* m = new sw_multi;
  m->p.assign(problems, problems + n);
  m->device = problems[0]->device;
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete m;
    return fail(-3, "stream creation failed: %s", hipGetErrorString(e));
  }
  *out = m;
  return 0;
}

SW_API void sw_multi_destroy(sw_multi* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) {
    (void)hipStreamSynchronize(m->stream);
    (void)hipStreamDestroy(m->stream);
  }
  delete m;
}

// Tile shapes of merged launches (one shot per workgroup, every problem's
// workgroups in each launch): the problem's own tile width, 16-row segments
// (measured on 8 equal problems of 0.03-0.26 M shot-cells per step: 16 rows
// 1.2-2.2x faster than the shapes a lone problem would get, 32-128 rows
// slower again).  Shapes the caller fixed are kept; for the others the
// automatic choice comes back with the next batch the context prepares.
struct MergedTiles {
  sw_multi* m;
  std::vector<int> auto_tiles;
  explicit MergedTiles(sw_multi* x) : m(x) {
    for (size_t q = 0; q < m->p.size(); ++q) {
      sw_ctx* c = m->p[q];
      if (c->tile_fixed) continue;
      auto_tiles.push_back(int(q));
      c->tile.nt_threads = c->g.nz >= 400 ? 256 : c->g.nz >= 100 ? 128 : 64;
      c->tile.seglen = std::min(16, c->g.nx);
      c->tile_fixed = true;
    }
  }
  ~MergedTiles() {
    for (int q : auto_tiles) m->p[q]->tile_fixed = false;
  }
};

SW_API int sw_multi_forward(sw_multi* m, const int32_t* nshots, const int32_t* const* isrc) {
  if (!m || !nshots || !isrc) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(m->device));
  SharedStream guard(m);
  MergedTiles tiles(m);
  const int n = int(m->p.size());
  int maxnt = 0;
  std::vector<int2> win(n);
  for (int q = 0; q < n; ++q) {
    int rc = sw_prepare(m->p[q], nshots[q], isrc[q]);
    if (rc) return rc;
    win[q] = make_int2(0, m->p[q]->g.nt);
    maxnt = std::max(maxnt, m->p[q]->g.nt);
  }
  hipEvent_t e0 = m->p[0]->ev0, e1 = m->p[0]->ev1;
  HIPCHK(hipEventRecord(e0, m->stream));
  int rc = multi_steps(m, M_FWD, maxnt, win);
  if (rc) return rc;
  HIPCHK(hipEventRecord(e1, m->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  m->loop_ms = ms;
  for (sw_ctx* c : m->p) {
    c->steps_done = c->g.nt;
    c->cur = c->g.nt & 1;
  }
  return 0;
}

namespace {
int multi_gradient_merged(sw_multi* m, const int32_t* nshots, const int32_t* const* isrc,
                          const float* const* obs_ref, int32_t ckpt_every, double* const* misfit) {
  HIPCHK(hipSetDevice(m->device));
  SharedStream guard(m);
  MergedTiles tiles(m);
  const int n = int(m->p.size());
  int rc = 0, max_nck = 0;
  hipEvent_t e0 = m->p[0]->ev0, e1 = m->p[0]->ev1;
  HIPCHK(hipEventRecord(e0, m->stream));
  for (int q = 0; q < n; ++q) {
    sw_ctx* c = m->p[q];
    if ((rc = grad_begin(c, ckpt_every))) return rc;
    if ((rc = grad_batch_setup(c, nshots[q], isrc[q]))) return rc;
    max_nck = std::max(max_nck, c->gp.nck);
  }
  const int K = m->p[0]->gp.K;
  for (sw_ctx* c : m->p)
    if (c->gp.K != K) return fail(-2, "ckpt_every exceeds nt of some problem; use ckpt_every <= min nt");
  // forward with checkpoints, all problems in lockstep
  std::vector<int2> win(n);
  for (int s = 0; s < max_nck; ++s) {
    int nl = 0;
    for (int q = 0; q < n; ++q) {
      sw_ctx* c = m->p[q];
      win[q] = make_int2(0, c->g.nt);
      if (s < c->gp.nck) {
        if ((rc = grad_save_ckpt(c, s))) return rc;
        nl = std::max(nl, std::min(K, c->g.nt - s * K));
      }
    }
    if ((rc = multi_steps(m, M_FWD, nl, win))) return rc;
    for (sw_ctx* c : m->p) {
      const int done = std::min(c->g.nt, s * K + nl);
      c->steps_done = done;
      c->cur = done & 1;
    }
  }
  for (int q = 0; q < n; ++q)
    if ((rc = grad_adjsrc_and_bwd_setup(m->p[q], obs_ref[q]))) return rc;
  // backward, segment by segment from the last
  for (int s = max_nck - 1; s >= 0; --s) {
    int nl = 0;
    for (int q = 0; q < n; ++q) {
      sw_ctx* c = m->p[q];
      if (s < c->gp.nck) {
        int n0, n1;
        if ((rc = grad_segment_setup(c, s, n0, n1))) return rc;
        win[q] = make_int2(n0, n1);
        nl = std::max(nl, n1 - n0);
      } else {
        win[q] = make_int2(0, 0);
      }
    }
    if ((rc = multi_steps(m, M_FWD_AUX, nl, win))) return rc;
    if ((rc = multi_steps(m, M_ADJ, nl, win))) return rc;
  }
  for (sw_ctx* c : m->p)
    if ((rc = grad_reduce(c))) return rc;
  for (sw_ctx* c : m->p)
    if ((rc = check_grad_finite(c))) return rc;
  HIPCHK(hipEventRecord(e1, m->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  m->loop_ms = ms;
  for (int q = 0; q < n; ++q)
    if (misfit && misfit[q])
      for (int s = 0; s < nshots[q]; ++s) misfit[q][s] = m->p[q]->gp.mis[s];
  return 0;
}
}  // namespace

SW_API int sw_multi_set_merge_limit(sw_multi* m, int64_t shot_cells) {
  if (!m || shot_cells < 0) return fail(-1, "bad arguments");
  m->merge_limit = shot_cells;
  return 0;
}

SW_API int sw_multi_set_workers(sw_multi* m, int32_t workers) {
  if (!m || workers < 1 || workers > 16) return fail(-1, "workers must be 1..16");
  m->workers = workers;
  return 0;
}

SW_API int sw_multi_gradient(sw_multi* m, const int32_t* nshots, const int32_t* const* isrc,
                             const float* const* obs_ref, int32_t ckpt_every, double* const* misfit) {
  if (!m || !nshots || !isrc || !obs_ref) return fail(-1, "bad arguments");
  HIPCHK(hipSetDevice(m->device));
  const auto t0 = std::chrono::steady_clock::now();
  const int n = int(m->p.size());
  std::vector<int> merged, solo;
  int rc = 0;
  for (int q = 0; q < n; ++q)
    (int64_t(nshots[q]) * m->p[q]->npt < m->merge_limit ? merged : solo).push_back(q);
  // problems that fill the GPU on their own: their own launches (own tile
  // shapes, two-step pairs), largest first, `workers` of them at a time on
  // their own streams.  Graph capture is off meanwhile: a capture in one
  // thread and another thread's allocations or synchronisations exclude each
  // other in HIP (measured: no loss without graphs at these sizes).
  std::stable_sort(solo.begin(), solo.end(), [&](int a, int b) {
    return int64_t(nshots[a]) * m->p[a]->npt * m->p[a]->g.nt > int64_t(nshots[b]) * m->p[b]->npt * m->p[b]->g.nt;
  });
  if (!solo.empty()) {
    // concurrency budget: `workers` problems that fill the GPU (at least
    // kFillShotCells shot-cells per step) at a time; a smaller problem counts
    // for its share of that, so latency-bound small problems run more at
    // once (small uneven batch, 48 problems of 128-448^2: 1.79x over one
    // problem after the other with 3 at a time, 2.21x with 12), while large
    // ones keep three (more would hold more run buffers without more speed)
    constexpr double kFillShotCells = double(1 << 22);
    const double budget = double(m->workers);
    std::vector<double> weight(solo.size());
    for (size_t i = 0; i < solo.size(); ++i)
      weight[i] = std::min(1.0, double(nshots[solo[i]]) * double(m->p[solo[i]]->npt) / kFillShotCells);
    std::mutex sched_mu;
    std::condition_variable sched_cv;
    double running = 0.0;
    int nrunning = 0;
    size_t next = 0;
    std::atomic<bool> failed{false};
    std::mutex err_mu;
    std::string first_err;
    int first_rc = 0;
    auto worker = [&]() {
      const hipError_t e = hipSetDevice(m->device);
      if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(err_mu);
        if (!failed.exchange(true)) {
          first_rc = -3;
          first_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
        }
        return;
      }
      for (;;) {
        size_t i;
        {
          // claim the next problem (largest first), then wait for room in the budget
          std::unique_lock<std::mutex> lk(sched_mu);
          i = next++;
          if (i >= solo.size() || failed.load()) return;
          sched_cv.wait(lk, [&] { return failed.load() || nrunning == 0 || running + weight[i] <= budget + 1e-9; });
          if (failed.load()) return;
          running += weight[i];
          ++nrunning;
        }
        const int q = solo[i];
        sw_ctx* c = m->p[q];
        const bool graphs = c->graphs_off;
        c->graphs_off = true;
        const int r = sw_gradient(c, nshots[q], isrc[q], obs_ref[q], ckpt_every, 0, misfit ? misfit[q] : nullptr);
        c->graphs_off = graphs;
        auto done = [&]() {
          {
            std::lock_guard<std::mutex> lk(sched_mu);
            running -= weight[i];
            --nrunning;
          }
          sched_cv.notify_all();
        };
        if (r) {
          {
            std::lock_guard<std::mutex> lk(err_mu);
            if (!failed.exchange(true)) {
              first_rc = r;
              first_err = g_err;
            }
          }
          done();
          return;
        }
        // its kernels are in K_RHO / K_MU / K_LAM: give the device memory of
        // the run (checkpoints, aux fields, adjoint state, per-shot kernels)
        // back before the next problem allocates its own.  sw_gradient has
        // synchronised the problem's stream, the only one these buffers were
        // used on: no device-wide wait, which would also wait for the other
        // workers' launches
        for (DBuf<float>* b : {&c->aux, &c->ckpt, &c->adj[0], &c->adj[1], &c->kshot, &c->spare}) b->release(true);
        done();
      }
    };
    // host threads: one per problem that may run at once (up to four per
    // unit of budget, so 12 by default)
    const int nw = m->workers == 1 ? 1 : std::max(1, std::min<int>(4 * m->workers, int(solo.size())));
    if (nw == 1) {
      worker();
    } else {
      std::vector<std::thread> th;
      for (int w = 0; w < nw; ++w) th.emplace_back(worker);
      for (auto& t : th) t.join();
    }
    if (failed.load()) return fail(first_rc, "%s", first_err.c_str());
  }
  if (!merged.empty()) {
    // the others in merged launches, over the subset
    std::vector<sw_ctx*> all = m->p, sub;
    std::vector<int32_t> ns;
    std::vector<const int32_t*> is;
    std::vector<const float*> ob;
    std::vector<double*> mi;
    for (int q : merged) {
      sub.push_back(all[q]);
      ns.push_back(nshots[q]);
      is.push_back(isrc[q]);
      ob.push_back(obs_ref[q]);
      mi.push_back(misfit ? misfit[q] : nullptr);
    }
    m->p = sub;
    rc = multi_gradient_merged(m, ns.data(), is.data(), ob.data(), ckpt_every, mi.data());
    m->p = all;
    if (rc) return rc;
  }
  m->loop_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

SW_API int sw_multi_timing(const sw_multi* m, double* ms, int64_t* launches) {
  if (!m) return fail(-1, "null argument");
  if (ms) *ms = m->loop_ms;
  if (launches) *launches = m->launches;
  return 0;
}

SW_API int sw_adjoint_plain(sw_ctx* c, int32_t nruns, const int32_t* isrc, const float* obs_ref,
                            int32_t adj_interval, float smooth, double* misfit) {
  if (!c || nruns < 1 || !isrc || !obs_ref) return fail(-1, "bad arguments");
  if (adj_interval <= 0) return fail(-2, "adj_interval must be > 0");
  const int nt = c->g.nt, nrec = c->nrec;
  const int nsa = nt / adj_interval;
  if (nsa == 0) return fail(-2, "adj_interval (%d) >= nt (%d)", adj_interval, nt);   // solver.rs:1311-1314
  HIPCHK(hipSetDevice(c->device));
  const size_t npt = size_t(c->npt);
  HIPCHK(c->kmu.ensure(npt));
  HIPCHK(launch_fill(c->kmu.p, 0.0f, int64_t(npt), c->stream));
  HIPCHK(c->fwd_store.ensure(size_t(nsa) * npt));
  const size_t nobs = size_t(nrec) * nt;
  HIPCHK(c->adj_src.ensure(std::max<size_t>(nobs, 1)));
  HIPCHK(c->ref.ensure(std::max<size_t>(nobs, 1)));
  // misfit taper (solver.rs:441-455) tabulated in f32 on the host
  std::vector<float> taper(nt);
  {
    const float PI_F = 3.14159265358979323846f;
    const float dt = c->g.dt;
    const float t_end = float(nt - 1) * dt, tw = t_end / 10.0f, t_min = tw, t_max = t_end - tw;
    for (int it = 0; it < nt; ++it) {
      const float t = float(it) * dt;
      if (t <= t_min) taper[it] = 0.5f + 0.5f * std::cos(PI_F * (t_min - t) / tw);
      else if (t >= t_max) taper[it] = 0.5f + 0.5f * std::cos(PI_F * (t_max - t) / tw);
      else taper[it] = 1.0f;
    }
  }
  HIPCHK(c->taper.ensure(std::max(nt, 1)));
  HIPCHK(hipMemcpyAsync(c->taper.p, taper.data(), size_t(nt) * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c->partial.ensure(std::max<size_t>((nobs + 255) / 256, 1)));
  HIPCHK(c->misfit_d.ensure(1));
  const float ndt = float(adj_interval) * c->g.dt;
  const size_t sstride = size_t(c->nsrc) * nt;
  const bool sh = uses_sh(c);
  struct ShOnly {
    sw_ctx* c;
    explicit ShOnly(sw_ctx* x) : c(x) { c->sh_only = true; }
    ~ShOnly() { c->sh_only = false; }
  } sh_only_scope(c);
  for (int r = 0; r < nruns; ++r) {
    if (isrc[r] >= c->nsrc) return fail(-2, "source %d out of range", isrc[r]);
    // forward run, keeping vy after every adj_interval-th step (solver.rs:1381-1414)
    std::vector<std::vector<Entry>> inj(1, source_entries(c, isrc[r]));
    c->val_x = c->stf.p; c->val_y = c->stf.p + sstride; c->val_z = c->stf.p + 2 * sstride; c->nt_val = nt;
    int rc = prepare_batch(c, 1, inj, true);
    if (rc) return rc;
    HIPCHK(launch_fill(c->fwd_store.p, 0.0f, int64_t(nsa) * int64_t(npt), c->stream));
    // blocks of adj_interval steps (multi-step launches, bit-identical to
    // one step per launch), vy kept after each
    for (int it = 0; it < nt;) {
      const int n = std::min(adj_interval - it % adj_interval, nt - it);
      if ((rc = launch_steps(c, n))) return rc;
      it += n;
      if (it % adj_interval == 0 && sh) {
        const int isa = nsa - it / adj_interval;
        HIPCHK(hipMemcpyAsync(c->fwd_store.p + size_t(isa) * npt, slot_ptr(c, S_VY, 0), npt * 4,
                              hipMemcpyDeviceToDevice, c->stream));
      }
    }
    // misfit + adjoint source (solver.rs:1425-1439)
    double mis = 0.0;
    if (sh && nobs) {
      HIPCHK(hipMemcpyAsync(c->ref.p, obs_ref + size_t(r) * nobs, nobs * 4, hipMemcpyHostToDevice, c->stream));
      int nb = 0;
      HIPCHK(launch_misfit(c->obs.p + 1 * nobs, c->ref.p, c->taper.p, c->adj_src.p, c->partial.p, nrec, nt,
                           c->stream, &nb));
      HIPCHK(launch_reduce_sum(c->partial.p, nb, c->misfit_d.p, c->stream));
      HIPCHK(hipMemcpyAsync(&mis, c->misfit_d.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      mis = std::sqrt(mis);
      if (c->finite_check && !std::isfinite(mis))
        return fail(-5, "misfit of run %d is not finite (%g): the forward wavefield diverged (time step above the "
                        "stability limit of this model) or an input is not finite", r, mis);
    } else if (nobs) {
      HIPCHK(launch_fill(c->adj_src.p, 0.0f, int64_t(nobs), c->stream));
    }
    if (misfit) misfit[r] = mis;
    // adjoint run: adjoint source injected at the receivers (adj_dsy,
    // fd2d.wgsl:562-568), imaging every adj_interval steps (solver.rs:1442-1498)
    std::vector<Entry> rin;
    for (int q = 0; q < nrec; ++q) {
      const uint32_t k = uint32_t(c->rec_k[q]);
      if (int64_t(k) >= c->npt) continue;
      rin.push_back({int(k / uint32_t(c->g.nz)), int(k % uint32_t(c->g.nz)), q});
    }
    std::vector<std::vector<Entry>> ainj(1, rin);
    c->val_x = c->adj_src.p; c->val_y = c->adj_src.p; c->val_z = c->adj_src.p; c->nt_val = nt;
    if ((rc = prepare_batch(c, 1, ainj, false))) return rc;
    // imaging after steps 0, I, 2I, ...: a block of one step, then blocks of I
    for (int it = 0; it < nt;) {
      const int n = std::min(it == 0 ? 1 : adj_interval, nt - it);
      if ((rc = launch_steps(c, n))) return rc;
      const int last = it + n - 1;   // the step after which the state now is
      it += n;
      if (last % adj_interval == 0 && sh) {
        const int isae = last / adj_interval;
        if (isae < nsa)
          HIPCHK(launch_imaging_sh(slot_ptr(c, S_VY, 0), c->fwd_store.p + size_t(isae) * npt, c->kmu.p, c->g.nx,
                                   c->g.nz, c->h8x, c->h8z, ndt, c->stream));
      }
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  const int rc = sw_smooth(c, SW_F_K_MU, smooth);
  return rc ? rc : check_finite(c, c->kmu.p, c->npt, "K_mu");
}

}  // extern "C"
