// Host side of the `seiswg <config.ini>` workflow (include/seiswg_host.h).
//
// Restates config.rs (INI parsing rules), main.rs (workflow dispatch) and the
// host parts of solver.rs (file formats, grid set-up, Ricker table, forward
// and adjoint drivers, writers) over the device ABI of seiswg_hip.h.  All
// shots of a forward workflow advance together in one batched launch per time
// step; outputs are written exactly as the reference writes them.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/seiswg_host.h"

namespace fs = std::filesystem;

namespace {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::string fmt(const char* f, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof buf, f, ap);
  va_end(ap);
  return buf;
}

// env_logger-style info lines on stderr; RUST_LOG=error|warn|off silences them
// (main.rs:34-36, integration.rs:83).
int log_level() {
  static int lvl = [] {
    const char* e = std::getenv("RUST_LOG");
    if (!e) return 3;
    std::string s(e);
    for (auto& ch : s) ch = char(std::tolower(ch));
    if (s == "off") return 0;
    if (s == "error") return 1;
    if (s == "warn") return 2;
    return 3;
  }();
  return lvl;
}
void log_info(const std::string& m) {
  if (log_level() >= 3) fprintf(stderr, "[INFO  seiswg] %s\n", m.c_str());
}
void log_warn(const std::string& m) {
  if (log_level() >= 2) fprintf(stderr, "[WARN  seiswg] %s\n", m.c_str());
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}
std::string lower(std::string s) {
  for (auto& ch : s) ch = char(std::tolower(static_cast<unsigned char>(ch)));
  return s;
}

// Grammar of Rust's f32/f64 FromStr: [+-]? (inf | infinity | nan | digits[.digits][exp] | .digits[exp])
bool rust_float_syntax(const std::string& s) {
  size_t p = 0, n = s.size();
  if (p < n && (s[p] == '+' || s[p] == '-')) ++p;
  std::string rest = lower(s.substr(p));
  if (rest == "inf" || rest == "infinity" || rest == "nan") return true;
  size_t d0 = p;
  while (p < n && std::isdigit(static_cast<unsigned char>(s[p]))) ++p;
  size_t ni = p - d0, nf = 0;
  if (p < n && s[p] == '.') {
    ++p;
    size_t f0 = p;
    while (p < n && std::isdigit(static_cast<unsigned char>(s[p]))) ++p;
    nf = p - f0;
  }
  if (ni + nf == 0) return false;
  if (p < n && (s[p] == 'e' || s[p] == 'E')) {
    ++p;
    if (p < n && (s[p] == '+' || s[p] == '-')) ++p;
    size_t e0 = p;
    while (p < n && std::isdigit(static_cast<unsigned char>(s[p]))) ++p;
    if (p == e0) return false;
  }
  return p == n;
}
bool parse_f64(const std::string& s, double& out) {
  if (!rust_float_syntax(s)) return false;
  out = std::strtod(s.c_str(), nullptr);
  return true;
}
bool parse_f32(const std::string& s, float& out) {
  if (!rust_float_syntax(s)) return false;
  out = std::strtof(s.c_str(), nullptr);
  return true;
}
bool parse_u32(const std::string& s, uint32_t& out) {
  size_t p = 0;
  if (!s.empty() && s[0] == '+') p = 1;
  if (p >= s.size()) return false;
  uint64_t v = 0;
  for (; p < s.size(); ++p) {
    if (!std::isdigit(static_cast<unsigned char>(s[p]))) return false;
    v = v * 10 + uint64_t(s[p] - '0');
    if (v > 0xFFFFFFFFull) return false;
  }
  out = uint32_t(v);
  return true;
}

// ---------------------------------------------------------------- INI
// configparser 3.1.0 rules as used by config.rs: sections and keys are
// lower-cased, '=' or ':' delimit, ';' and '#' start comments, a key without
// a delimiter has no value.
struct Ini {
  std::map<std::string, std::map<std::string, std::pair<bool, std::string>>> sec;
  bool get(const std::string& s, const std::string& k, std::string& v) const {
    auto a = sec.find(s);
    if (a == sec.end()) return false;
    auto b = a->second.find(k);
    if (b == a->second.end() || !b->second.first) return false;
    v = b->second.second;
    return true;
  }
};

Ini parse_ini_text(const std::string& text) {
  Ini ini;
  std::string section = "default";
  std::istringstream in(text);
  std::string raw;
  int num = 0;
  while (std::getline(in, raw)) {
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    size_t c = raw.find_first_of(";#");
    std::string line = trim(c == std::string::npos ? raw : raw.substr(0, c));
    ++num;
    if (line.empty()) continue;
    if (line[0] == '[') {
      size_t e = line.rfind(']');
      if (e == std::string::npos)
        throw Error(fmt("line %d: Found opening bracket for section name but no closing bracket", num - 1));
      section = lower(trim(line.substr(1, e - 1)));
      ini.sec[section];
      continue;
    }
    size_t d = line.find_first_of("=:");
    if (d != std::string::npos) {
      std::string key = lower(trim(line.substr(0, d)));
      if (key.empty()) throw Error(fmt("line %d:%zu: Key cannot be empty", num - 1, d));
      ini.sec[section][key] = {true, trim(line.substr(d + 1))};
    } else {
      ini.sec[section][lower(line)] = {false, ""};
    }
  }
  return ini;
}

struct Config {
  std::string workflow;
  uint32_t nt = 0;
  float dt = 0;
  bool left = false, right = false, bottom = false, top = false;
  uint32_t width = 20;
  float alpha = 0.015f;
  bool sh = false, psv = true, spin = false;
  uint32_t threads_per_block = 128;
  bool combine = false;
  uint32_t save_snapshot = 0, adj_interval = 10;
  float smooth = 5.0f;
  // workflow = inversion (section [inversion]; no reference counterpart)
  uint32_t inv_niter = 3, inv_tries = 4, inv_ckpt = 50;
  float inv_step = 0.02f;
  fs::path output, output_traces, model_true, model_init, sources, stations, traces;
  bool has_output_traces = false, has_model_true = false, has_model_init = false, has_traces = false;
};

// config.rs:62-192; relative paths resolve against `base` (the INI's
// directory, or "" for INI text given in memory as lib.rs:103 does)
Config load_config_text(const std::string& text, const fs::path& base) {
  Ini ini = parse_ini_text(text);
  auto required = [&](const char* s, const char* k) {
    std::string v;
    if (!ini.get(s, k, v)) throw Error(fmt("[%s] %s is required", s, k));
    return v;
  };
  auto get_bool = [&](const char* k, bool d) {
    std::string v;
    return ini.get("solver", k, v) ? lower(trim(v)) == "yes" : d;
  };
  auto get_u32_or = [&](const char* k, uint32_t d) {
    std::string v;
    uint32_t x;
    return (ini.get("solver", k, v) && parse_u32(trim(v), x)) ? x : d;
  };
  auto get_f32_or = [&](const char* k, float d) {
    std::string v;
    float x;
    return (ini.get("solver", k, v) && parse_f32(trim(v), x)) ? x : d;
  };
  auto resolve = [&](const std::string& raw) {
    fs::path p(trim(raw));
    return p.is_absolute() ? p : base / p;
  };
  Config c;
  c.workflow = lower(trim(required("workflow", "mode")));
  if (!parse_u32(trim(required("solver", "nt")), c.nt)) throw Error("[solver] nt must be an integer");
  if (!parse_f32(trim(required("solver", "dt")), c.dt)) throw Error("[solver] dt must be a float");
  c.left = get_bool("abs_left", false);
  c.right = get_bool("abs_right", false);
  c.bottom = get_bool("abs_bottom", false);
  c.top = get_bool("abs_top", false);
  c.width = get_u32_or("abs_width", 20);
  c.alpha = get_f32_or("abs_alpha", 0.015f);
  c.sh = get_bool("sh", false);
  c.psv = get_bool("psv", true);
  c.spin = get_bool("spin", false);
  c.threads_per_block = get_u32_or("threads_per_block", 128);
  c.combine = get_bool("combine_sources", false);
  c.save_snapshot = get_u32_or("save_snapshot", 0);
  c.adj_interval = get_u32_or("adj_interval", 10);
  c.smooth = get_f32_or("smooth", 5.0f);
  std::string v;
  {
    uint32_t u;
    float f;
    if (ini.get("inversion", "niter", v) && parse_u32(trim(v), u)) c.inv_niter = u;
    if (ini.get("inversion", "linesearch", v) && parse_u32(trim(v), u)) c.inv_tries = std::max(u, 1u);
    if (ini.get("inversion", "checkpoint", v) && parse_u32(trim(v), u)) c.inv_ckpt = std::max(u, 1u);
    if (ini.get("inversion", "step", v) && parse_f32(trim(v), f)) c.inv_step = f;
  }
  c.output = resolve(required("path", "output"));
  if (ini.get("path", "output_traces", v)) { c.output_traces = resolve(v); c.has_output_traces = true; }
  if (ini.get("path", "model_true", v)) { c.model_true = resolve(v); c.has_model_true = true; }
  if (ini.get("path", "model_init", v)) { c.model_init = resolve(v); c.has_model_init = true; }
  c.sources = resolve(required("path", "sources"));
  c.stations = resolve(required("path", "stations"));
  if (ini.get("path", "traces", v)) { c.traces = resolve(v); c.has_traces = true; }
  return c;
}

Config load_config(const fs::path& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(fmt("Cannot open %s: %s", path.string().c_str(), std::strerror(errno)));
  std::stringstream ss;
  ss << f.rdbuf();
  return load_config_text(ss.str(), path.parent_path());
}

// ---------------------------------------------------------------- files
using Vfs = std::map<std::string, std::vector<char>>;

// Where a workflow's files come from and go to.  The CLI reads and writes the
// filesystem.  The in-memory entry point (lib.rs:74-180) looks its inputs up
// in a path-keyed map first, by the resolved path or the path without a
// leading "./" (solver.rs:236-266), and hands every output back in another
// map keyed as the reference's Solver returns them (solver.rs:941-1119,
// 1298-1529) without touching the filesystem.
struct Io {
  const Vfs* in = nullptr;   // inputs by path, looked up before the filesystem
  Vfs* out = nullptr;        // outputs by path
  bool disk = true;          // read inputs missing from `in` from, and write outputs to, the filesystem
};

std::vector<char> read_all(const Io& io, const fs::path& p, bool& ok) {
  if (io.in) {
    const std::string key = p.string();
    auto it = io.in->find(key);
    if (it == io.in->end() && key.rfind("./", 0) == 0) it = io.in->find(key.substr(2));
    if (it != io.in->end()) {
      ok = true;
      return it->second;
    }
  }
  ok = false;
  if (!io.disk) return {};
  std::ifstream f(p, std::ios::binary);
  ok = bool(f);
  if (!ok) return {};
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void make_dirs(const Io& io, const fs::path& d) {
  if (!io.disk) return;
  std::error_code ec;
  fs::create_directories(d, ec);
  if (ec) throw Error(fmt("Cannot create %s: %s", d.string().c_str(), ec.message().c_str()));
}

// one output file: into the output map and/or onto the filesystem
void emit(const Io& io, const fs::path& p, const std::vector<char>& bytes) {
  if (io.out) (*io.out)[p.string()] = bytes;
  if (!io.disk) return;
  std::ofstream f(p, std::ios::binary | std::ios::trunc);
  if (!f) throw Error(fmt("Cannot write %s", p.string().c_str()));
  f.write(bytes.data(), std::streamsize(bytes.size()));
  if (!f) throw Error(fmt("Cannot write %s", p.string().c_str()));
}

// load_model_field (solver.rs:300-325): i32 npt, then npt f32; extra bytes ignored
std::vector<float> load_model_field(const Io& io, const fs::path& p) {
  bool ok;
  std::vector<char> b = read_all(io, p, ok);
  if (!ok) throw Error(fmt("Cannot read model file %s", p.string().c_str()));
  if (b.size() < 4) throw Error(fmt("Model file too small: %s", p.string().c_str()));
  int32_t n;
  std::memcpy(&n, b.data(), 4);
  const size_t expected = 4 + size_t(uint32_t(n)) * 4;
  if (n < 0 || b.size() < expected)
    throw Error(fmt("Model file %s too short (expected %zu bytes, got %zu)", p.string().c_str(), expected, b.size()));
  std::vector<float> v(static_cast<size_t>(n));
  if (n) std::memcpy(v.data(), b.data() + 4, size_t(n) * 4);
  return v;
}

// load_text_table (solver.rs:328-346)
std::vector<std::vector<double>> load_table(const Io& io, const fs::path& p) {
  bool ok;
  std::vector<char> b = read_all(io, p, ok);
  if (!ok) throw Error(fmt("Cannot read %s", p.string().c_str()));
  std::vector<std::vector<double>> rows;
  std::istringstream in(std::string(b.begin(), b.end()));
  std::string line;
  while (std::getline(in, line)) {
    std::string t = trim(line);
    if (t.empty() || t[0] == '#') continue;
    std::vector<double> row;
    std::istringstream ls(line);
    std::string tok;
    while (ls >> tok) {
      double x;
      row.push_back(parse_f64(tok, x) ? x : 0.0);
    }
    rows.push_back(row);
  }
  return rows;
}

// make_npy_bytes (solver.rs:361-389): NumPy v1.0, header padded to a
// multiple of 64 bytes
std::vector<char> npy_bytes(const float* data, size_t nrows, size_t ncols) {
  std::string dict = fmt("{'descr': '<f4', 'fortran_order': False, 'shape': (%zu, %zu), }", nrows, ncols);
  const size_t hl = ((dict.size() + 1 + 63) / 64) * 64;
  std::string header = dict;
  while (header.size() < hl - 1) header.push_back(' ');
  header.push_back('\n');
  std::vector<char> b;
  b.reserve(10 + hl + nrows * ncols * 4);
  static const char magic[] = "\x93NUMPY\x01\x00";   // 8 bytes: magic, version 1.0
  b.insert(b.end(), magic, magic + 8);
  const uint16_t h16 = uint16_t(hl);
  b.push_back(char(h16 & 0xff));
  b.push_back(char(h16 >> 8));
  b.insert(b.end(), header.begin(), header.end());
  const char* d = reinterpret_cast<const char*>(data);
  b.insert(b.end(), d, d + nrows * ncols * 4);
  return b;
}

// raw little-endian f32, with the model format's i32 count in front if given
std::vector<char> raw_bytes(const float* data, size_t n, const int32_t* header = nullptr) {
  std::vector<char> b;
  b.reserve(n * 4 + (header ? 4 : 0));
  if (header) b.insert(b.end(), reinterpret_cast<const char*>(header), reinterpret_cast<const char*>(header) + 4);
  const char* d = reinterpret_cast<const char*>(data);
  b.insert(b.end(), d, d + n * 4);
  return b;
}

// load_npy_f32 (solver.rs:401-428)
std::vector<float> load_npy(const Io& io, const fs::path& p) {
  bool ok;
  std::vector<char> b = read_all(io, p, ok);
  if (!ok) throw Error(fmt("Cannot load observed traces %s", p.string().c_str()));
  if (b.size() < 10 || std::memcmp(b.data(), "\x93NUMPY", 6) != 0)
    throw Error(fmt("Cannot load observed traces %s: Not a NumPy file", p.string().c_str()));
  const size_t hl = size_t(uint8_t(b[8])) | (size_t(uint8_t(b[9])) << 8);
  const size_t start = 10 + hl;
  if (b.size() < start) throw Error(fmt("Truncated .npy header in %s", p.string().c_str()));
  if ((b.size() - start) % 4) throw Error(fmt("Payload length not a multiple of 4 in %s", p.string().c_str()));
  std::vector<float> v((b.size() - start) / 4);
  if (!v.empty()) std::memcpy(v.data(), b.data() + start, v.size() * 4);
  return v;
}

uint32_t sat_u32(float v) {   // Rust `as u32`: NaN -> 0, saturating
  if (!(v == v) || v <= 0.0f) return 0;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return uint32_t(v);
}

void check(int rc) {
  if (rc != 0) throw Error(sw_last_error());
}

struct Problem {
  Config cfg;
  Io io;
  int32_t nx = 0, nz = 0;
  float dx = 0, dz = 0;
  int64_t npt = 0;
  std::vector<std::vector<float>> model;   // the loaded model fields (rho, vp, vs / Lame set)
  std::vector<float> x, z;                 // coordinates of the model files
  std::vector<uint32_t> src_k, rec_k;
  sw_ctx* ctx = nullptr;
  ~Problem() { sw_destroy(ctx); }
};

// Solver::new (solver.rs:578-933)
void solver_new(Problem& P) {
  const Config& c = P.cfg;
  fs::path model_dir;
  if (c.has_model_init) model_dir = c.model_init;
  else if (c.has_model_true) model_dir = c.model_true;
  else throw Error("No model path in config");
  const bool spin = c.spin;
  std::vector<float> rho = load_model_field(P.io, model_dir / "proc000000_rho.bin");
  const int64_t npt = int64_t(rho.size());
  std::vector<float> x = load_model_field(P.io, model_dir / "proc000000_x.bin");
  std::vector<float> z = load_model_field(P.io, model_dir / "proc000000_z.bin");
  if (int64_t(x.size()) != npt || int64_t(z.size()) != npt)
    throw Error(fmt("coordinate files hold %zu/%zu points, expected %lld", x.size(), z.size(), (long long)npt));
  int32_t fb = 0;
  check(seiswg_infer_grid(npt, x.data(), z.data(), &P.nx, &P.nz, &P.dx, &P.dz, &fb));
  if (fb)
    log_warn(fmt("grid rule round(sqrt(npt*lx/lz)) does not factor npt=%lld; using the exact shape %dx%d taken "
                 "from the x coordinates",
                 (long long)npt, P.nx, P.nz));
  log_info(fmt("Grid: nx=%d nz=%d npt=%lld dx=%.1f dz=%.1f", P.nx, P.nz, (long long)npt, P.dx, P.dz));
  P.npt = npt;
  std::vector<std::vector<float>> fields;
  const char* names_vs[] = {"rho", "vp", "vs"};
  const char* names_sp[] = {"rho", "lambda", "mu", "nu", "j", "lambda_c", "mu_c", "nu_c"};
  fields.push_back(std::move(rho));
  const int nf = spin ? 8 : 3;
  for (int q = 1; q < nf; ++q) {
    const char* nm = spin ? names_sp[q] : names_vs[q];
    fields.push_back(load_model_field(P.io, model_dir / (std::string("proc000000_") + nm + ".bin")));
    if (int64_t(fields.back().size()) < npt)
      throw Error(fmt("Model file proc000000_%s.bin holds %zu values, expected %lld", nm, fields.back().size(),
                      (long long)npt));
  }
  // sources (solver.rs:702-739)
  auto src_table = load_table(P.io, c.sources);
  const int nsrc = int(src_table.size());
  std::vector<float> sp(size_t(nsrc) * 4);
  for (int s = 0; s < nsrc; ++s) {
    const auto& r = src_table[s];
    auto g = [&](size_t n, double d) { return n < r.size() ? r[n] : d; };
    const float sx = float(g(0, 0.0)), sz = float(g(1, 0.0));
    sp[4 * s + 0] = float(g(3, 10.0));
    sp[4 * s + 1] = float(g(4, 0.1));
    sp[4 * s + 2] = float(g(5, 0.0));
    sp[4 * s + 3] = float(g(6, 1.0e10));
    P.src_k.push_back(seiswg_grid_index(sx, sz, P.dx, P.dz, uint32_t(P.nz)));
  }
  std::vector<float> stf(size_t(3) * nsrc * c.nt);
  check(seiswg_ricker_table(nsrc, int32_t(c.nt), c.dt, sp.data(), stf.data()));
  // stations (solver.rs:741-751)
  auto rec_table = load_table(P.io, c.stations);
  for (const auto& r : rec_table) {
    const float rx = float(r.size() > 0 ? r[0] : 0.0), rz = float(r.size() > 1 ? r[1] : 0.0);
    P.rec_k.push_back(seiswg_grid_index(rx, rz, P.dx, P.dz, uint32_t(P.nz)));
  }
  // device module
  sw_grid g{};
  g.nx = P.nx; g.nz = P.nz; g.nt = int32_t(c.nt);
  g.dx = P.dx; g.dz = P.dz; g.dt = c.dt;
  g.sh = c.sh; g.psv = c.psv; g.spin = c.spin;
  g.abs_left = c.left; g.abs_right = c.right; g.abs_bottom = c.bottom; g.abs_top = c.top;
  g.abs_width = int32_t(c.width); g.abs_alpha = c.alpha;
  check(sw_create(&g, 0, &P.ctx));
  std::vector<const float*> fp;
  for (auto& f : fields) fp.push_back(f.data());
  check(sw_set_model(P.ctx, spin ? SW_MODEL_LAME : SW_MODEL_VPVS, fp.data()));
  P.model = std::move(fields);
  P.x = std::move(x);
  P.z = std::move(z);
  std::vector<int32_t> si(P.src_k.begin(), P.src_k.end()), ri(P.rec_k.begin(), P.rec_k.end());
  check(sw_set_sources(P.ctx, nsrc, si.data(), stf.data()));
  check(sw_set_receivers(P.ctx, int32_t(ri.size()), ri.data()));
}

struct SnapWriter {
  const Io* io;
  fs::path out_dir;
  int last_shot;
  // frames by (shot, step, field), kept when outputs go to a map
  std::map<std::tuple<int, int, int>, std::vector<float>> frames;
  std::string err;
};

int snap_cb(void* user, int32_t shot, int32_t it, int32_t field, const float* data, int64_t npt) {
  auto* w = static_cast<SnapWriter*>(user);
  try {
    if (w->io->out) w->frames[{shot, it, field}].assign(data, data + npt);
    // later runs overwrite the same files (solver.rs:1144-1176)
    if (w->io->disk && shot == w->last_shot) {
      const char* nm = field == SW_F_VY ? "vy" : field == SW_F_VX ? "vx" : field == SW_F_VZ ? "vz" : "ry";
      Io disk_only{nullptr, nullptr, true};
      emit(disk_only, w->out_dir / fmt("proc%06d_%s.bin", it, nm), raw_bytes(data, size_t(npt)));
    }
  } catch (const std::exception& e) {
    w->err = e.what();
    return 1;
  }
  return 0;
}

// run_forward (solver.rs:941-1119)
void run_forward(Problem& P) {
  const Config& c = P.cfg;
  const int nsrc = int(P.src_k.size()), nrec = int(P.rec_k.size()), nt = int(c.nt);
  fs::path trace_dir = c.has_output_traces ? c.output_traces : c.output / "traces";
  make_dirs(P.io, trace_dir);
  if (c.save_snapshot > 0) make_dirs(P.io, c.output);
  std::vector<int32_t> runs;
  if (c.combine) runs.push_back(-1);
  else for (int s = 0; s < nsrc; ++s) runs.push_back(s);
  if (runs.empty()) return;
  for (int32_t r : runs)
    log_info(fmt("Forward run  isrc=%s  (%d time steps)", r < 0 ? "all" : std::to_string(r).c_str(), nt));
  SnapWriter w{&P.io, c.output, int(runs.size()) - 1, {}, ""};
  const int rc = sw_forward_batch(P.ctx, int32_t(runs.size()), runs.data(), int32_t(c.save_snapshot), snap_cb, &w);
  if (rc) throw Error(w.err.empty() ? std::string(sw_last_error()) : w.err);
  sw_timing tm{};
  sw_get_timing(P.ctx, &tm);
  log_info(fmt("time loop: %lld steps x %zu shots in %.1f ms (%.3g cell-updates/s)", (long long)tm.steps,
               runs.size(), tm.loop_ms, tm.loop_ms > 0 ? double(tm.cell_updates) / (tm.loop_ms * 1e-3) : 0.0));
  std::vector<float> comp(size_t(nrec) * nt);
  for (size_t s = 0; s < runs.size(); ++s) {
    const int idx = runs[s] < 0 ? 0 : runs[s];
    auto get = [&](int cc) { check(sw_read_obs(P.ctx, int32_t(s), cc, comp.data())); };
    auto npy = [&](const char* nm) { emit(P.io, trace_dir / fmt("%s_%06d.npy", nm, idx), npy_bytes(comp.data(), nrec, nt)); };
    auto raw = [&](const char* nm) { emit(P.io, trace_dir / fmt("%s_%06d.npy", nm, idx), raw_bytes(comp.data(), comp.size())); };
    if (c.sh) {
      get(SW_OBS_Y);
      npy("uy");
      raw("vy");
    }
    if (c.psv) {
      get(SW_OBS_X);
      npy("ux");
      raw("vx");
      get(SW_OBS_Z);
      npy("uz");
      raw("vz");
      if (c.spin) {
        get(SW_OBS_Y);
        npy("ry");
        get(SW_OBS_YI);
        npy("yi");
      }
    }
  }
  log_info("  → traces saved to " + trace_dir.string());
  // snapshot frames (solver.rs:1104-1117): "snapshot/nframes" = u32 count of
  // snapshot steps over all runs; "snapshot/<field>.f32" = that field's
  // frames, run after run, step after step
  if (P.io.out && !w.frames.empty()) {
    std::map<std::pair<int, int>, int> steps;
    std::map<std::string, std::vector<char>> per_field;
    for (const auto& f : w.frames) {
      const int shot = std::get<0>(f.first), field = std::get<2>(f.first);
      steps[{shot, std::get<1>(f.first)}] = 1;
      const char* nm = field == SW_F_VY ? "vy" : field == SW_F_VX ? "vx" : field == SW_F_VZ ? "vz" : "ry";
      std::vector<char>& b = per_field[nm];
      const char* d = reinterpret_cast<const char*>(f.second.data());
      b.insert(b.end(), d, d + f.second.size() * 4);
    }
    const uint32_t n = uint32_t(steps.size());
    (*P.io.out)["snapshot/nframes"] = std::vector<char>(reinterpret_cast<const char*>(&n),
                                                         reinterpret_cast<const char*>(&n) + 4);
    for (auto& f : per_field) (*P.io.out)["snapshot/" + f.first + ".f32"] = std::move(f.second);
  }
}

// run_adjoint (solver.rs:1298-1529)
void run_adjoint(Problem& P) {
  const Config& c = P.cfg;
  const int nsrc = int(P.src_k.size()), nrec = int(P.rec_k.size()), nt = int(c.nt);
  if (c.adj_interval == 0) throw Error("adj_interval must be > 0");
  const uint32_t nsa = c.nt / c.adj_interval;
  if (nsa == 0) throw Error(fmt("adj_interval (%u) >= nt (%u)", c.adj_interval, c.nt));
  make_dirs(P.io, c.output);
  if (!c.has_traces) throw Error("config [path] traces is required for the adjoint workflow");
  std::vector<int32_t> runs;
  std::vector<int> ref_idx;
  if (c.combine) { runs.push_back(-1); ref_idx.push_back(0); }
  else for (int s = 0; s < nsrc; ++s) { runs.push_back(s); ref_idx.push_back(s); }
  const size_t nobs = size_t(nrec) * nt;
  std::vector<float> ref(runs.size() * nobs, 0.0f);
  if (c.sh) {
    for (size_t r = 0; r < runs.size(); ++r) {
      const int i = c.combine ? 0 : ref_idx[r];
      fs::path p = c.traces / fmt("uy_%06d.npy", i);
      std::vector<float> v = load_npy(P.io, p);
      if (v.size() < nobs)
        throw Error(fmt("observed traces %s hold %zu values, expected %zu", p.string().c_str(), v.size(), nobs));
      std::copy(v.begin(), v.begin() + nobs, ref.begin() + r * nobs);
    }
  }
  for (size_t r = 0; r < runs.size(); ++r)
    log_info(fmt("Adjoint: source %zu/%zu (%d steps, %u checkpoints)…", r + 1, runs.size(), nt, nsa));
  std::vector<double> mis(std::max<size_t>(runs.size(), 1), 0.0);
  check(sw_adjoint_plain(P.ctx, int32_t(runs.size()), runs.data(), ref.data(), int32_t(c.adj_interval), c.smooth,
                         mis.data()));
  double total = 0.0;
  for (size_t r = 0; r < runs.size(); ++r) {
    log_info(fmt("Adjoint: source %zu misfit = %.6e", r + 1, mis[r]));
    total += mis[r];
  }
  log_info(fmt("Adjoint: total misfit = %.6e", total));
  log_info(fmt("Adjoint: smoothing kernel (sigma=%.1f)…", c.smooth));
  std::vector<float> kmu(size_t(P.npt));
  check(sw_read_field(P.ctx, 0, SW_F_K_MU, kmu.data()));
  const int32_t hdr = int32_t(P.npt);
  fs::path kp = c.output / "proc000000_kmu.bin";
  emit(P.io, kp, raw_bytes(kmu.data(), kmu.size()));
  emit(P.io, c.output / "gradient_kmu.bin", raw_bytes(kmu.data(), kmu.size(), &hdr));
  log_info("Adjoint: gradient saved to " + kp.string());
}

// workflow = inversion.  The reference lists it as planned (README.md:61) and
// stops on unknown workflows (main.rs:94); this is a steepest-descent loop on
// top of the batched exact gradient: per iteration the misfit gradient of all
// runs (sw_gradient, checkpointed), Gaussian smoothing of K_mu / K_lambda as
// in the adjoint workflow, the chain rule through vps2lm to vs (and vp for
// P-SV; rho stays), a step of at most `step` times the largest velocity, a
// backtracking line search (halving, `linesearch` trials) on the misfit of a
// forward run, and the accepted model written in the proc000000_*.bin format
// to output/iter_NNN.  The misfit is that of the gradient: the sum over runs,
// components, receivers and samples of (w_n (d - obs))^2, w the cosine taper
// of solver.rs:441-455.
double run_misfit(Problem& P, const std::vector<int32_t>& runs, const std::vector<int>& comps,
                  const std::vector<float>& ref) {
  const Config& c = P.cfg;
  const int nrec = int(P.rec_k.size()), nt = int(c.nt);
  check(sw_forward_batch(P.ctx, int32_t(runs.size()), runs.data(), 0, nullptr, nullptr));
  const float PI_F = 3.14159265358979323846f, dt = c.dt;
  const float t_end = float(nt - 1) * dt, tw = t_end / 10.0f, t_min = tw, t_max = t_end - tw;
  std::vector<float> w(static_cast<size_t>(nt));
  for (int n = 0; n < nt; ++n) {
    const float t = float(n) * dt;
    w[n] = t <= t_min ? 0.5f + 0.5f * std::cos(PI_F * (t_min - t) / tw)
           : t >= t_max ? 0.5f + 0.5f * std::cos(PI_F * (t_max - t) / tw) : 1.0f;
  }
  std::vector<float> d(size_t(nrec) * nt);
  double J = 0.0;
  const size_t per_comp = size_t(nrec) * nt;
  for (size_t r = 0; r < runs.size(); ++r)
    for (size_t k = 0; k < comps.size(); ++k) {
      check(sw_read_obs(P.ctx, int32_t(r), comps[k], d.data()));
      const float* o = ref.data() + (r * comps.size() + k) * per_comp;
      for (int q = 0; q < nrec; ++q)
        for (int n = 0; n < nt; ++n) {
          const double e = double(w[n] * (d[size_t(q) * nt + n] - o[size_t(q) * nt + n]));
          J += e * e;
        }
    }
  return J;
}

void write_model_dir(const Problem& P, const fs::path& dir) {
  make_dirs(P.io, dir);
  const int32_t hdr = int32_t(P.npt);
  const char* names[] = {"rho", "vp", "vs"};
  for (int q = 0; q < 3; ++q)
    emit(P.io, dir / (std::string("proc000000_") + names[q] + ".bin"), raw_bytes(P.model[q].data(), size_t(P.npt), &hdr));
  emit(P.io, dir / "proc000000_x.bin", raw_bytes(P.x.data(), size_t(P.npt), &hdr));
  emit(P.io, dir / "proc000000_z.bin", raw_bytes(P.z.data(), size_t(P.npt), &hdr));
}

void run_inversion(Problem& P) {
  const Config& c = P.cfg;
  if (c.spin || c.sh == c.psv) throw Error("the inversion needs exactly one of sh / psv, without spin");
  if (!c.has_traces) throw Error("config [path] traces is required for the inversion workflow");
  const int nsrc = int(P.src_k.size()), nrec = int(P.rec_k.size()), nt = int(c.nt);
  const size_t npt = size_t(P.npt);
  make_dirs(P.io, c.output);
  std::vector<int32_t> runs;
  if (c.combine) runs.push_back(-1);
  else for (int s = 0; s < nsrc; ++s) runs.push_back(s);
  const std::vector<int> comps = c.sh ? std::vector<int>{SW_OBS_Y} : std::vector<int>{SW_OBS_X, SW_OBS_Z};
  const char* comp_names[] = {"ux", "uy", "uz"};
  const size_t nobs = size_t(nrec) * nt;
  std::vector<float> ref(runs.size() * comps.size() * nobs, 0.0f);
  for (size_t r = 0; r < runs.size(); ++r)
    for (size_t k = 0; k < comps.size(); ++k) {
      const int idx = c.combine ? 0 : int(r);
      fs::path p = c.traces / fmt("%s_%06d.npy", comp_names[comps[k] == SW_OBS_X ? 0 : comps[k] == SW_OBS_Y ? 1 : 2],
                                  idx);
      std::vector<float> v = load_npy(P.io, p);
      if (v.size() < nobs)
        throw Error(fmt("observed traces %s hold %zu values, expected %zu", p.string().c_str(), v.size(), nobs));
      std::copy(v.begin(), v.begin() + nobs, ref.begin() + (r * comps.size() + k) * nobs);
    }
  std::vector<float>& rho = P.model[0];
  std::vector<float>& vp = P.model[1];
  std::vector<float>& vs = P.model[2];
  auto set_model = [&](const std::vector<float>& a, const std::vector<float>& b) {
    const float* f[3] = {rho.data(), a.data(), b.data()};
    check(sw_set_model(P.ctx, SW_MODEL_VPVS, f));
  };
  // misfit per accepted iteration, rewritten after each one
  std::string mlog;
  auto write_log = [&](uint32_t iter, double J) {
    mlog += std::to_string(iter) + " " + fmt("%.9e", J) + "\n";
    emit(P.io, c.output / "misfit.txt", std::vector<char>(mlog.begin(), mlog.end()));
  };
  double J0 = run_misfit(P, runs, comps, ref);
  log_info(fmt("Inversion: %zu run(s), initial misfit %.6e", runs.size(), J0));
  write_log(0, J0);
  std::vector<double> mis(runs.size());
  std::vector<float> kmu(npt), klam(npt), gvs(npt), gvp(npt), vs1(npt), vp1(npt);
  for (uint32_t iter = 1; iter <= c.inv_niter; ++iter) {
    check(sw_gradient(P.ctx, int32_t(runs.size()), runs.data(), ref.data(), int32_t(c.inv_ckpt), 0, mis.data()));
    check(sw_smooth(P.ctx, SW_F_K_MU, c.smooth));
    check(sw_read_field(P.ctx, 0, SW_F_K_MU, kmu.data()));
    if (c.psv) {
      check(sw_smooth(P.ctx, SW_F_K_LAM, c.smooth));
      check(sw_read_field(P.ctx, 0, SW_F_K_LAM, klam.data()));
    } else {
      std::fill(klam.begin(), klam.end(), 0.0f);
    }
    // chain rule through mu = rho vs^2, lambda = rho (vp^2 - 2 vs^2) (0 where vp <= vs)
    float gmax_s = 0.0f, gmax_p = 0.0f, vs_max = 0.0f, vp_max = 0.0f;
    for (size_t k = 0; k < npt; ++k) {
      const bool lam_on = vp[k] > vs[k];
      gvs[k] = kmu[k] * 2.0f * rho[k] * vs[k] + (lam_on ? klam[k] * (-4.0f) * rho[k] * vs[k] : 0.0f);
      gvp[k] = lam_on ? klam[k] * 2.0f * rho[k] * vp[k] : 0.0f;
      gmax_s = std::max(gmax_s, std::fabs(gvs[k]));
      gmax_p = std::max(gmax_p, std::fabs(gvp[k]));
      vs_max = std::max(vs_max, vs[k]);
      vp_max = std::max(vp_max, vp[k]);
    }
    if (!(gmax_s > 0.0f) && !(gmax_p > 0.0f)) {
      log_info("Inversion: zero gradient; stopping");
      break;
    }
    float step = c.inv_step;
    bool accepted = false;
    double J1 = J0;
    for (uint32_t tr = 0; tr < c.inv_tries && !accepted; ++tr, step *= 0.5f) {
      const float as = gmax_s > 0.0f ? step * vs_max / gmax_s : 0.0f;
      const float ap = (c.psv && gmax_p > 0.0f) ? step * vp_max / gmax_p : 0.0f;
      for (size_t k = 0; k < npt; ++k) {
        vs1[k] = std::max(vs[k] - as * gvs[k], 0.0f);
        vp1[k] = c.psv ? std::max(vp[k] - ap * gvp[k], 0.0f) : vp[k];
      }
      set_model(vp1, vs1);
      J1 = run_misfit(P, runs, comps, ref);
      log_info(fmt("Inversion: iteration %u, step %.4g: misfit %.6e (from %.6e)", iter, step, J1, J0));
      accepted = J1 < J0;
    }
    if (!accepted) {
      set_model(vp, vs);
      log_info("Inversion: no step lowers the misfit; stopping");
      break;
    }
    vs.swap(vs1);
    vp.swap(vp1);
    J0 = J1;
    write_model_dir(P, c.output / fmt("iter_%03u", iter));
    write_log(iter, J0);
    log_info(fmt("Inversion: iteration %u accepted, misfit %.6e; model written to %s", iter, J0,
                 (c.output / fmt("iter_%03u", iter)).string().c_str()));
  }
}

// Solver::new + the workflow of the config (main.rs:80-97)
void run_workflow(Problem& P) {
  log_info("Workflow: " + P.cfg.workflow);
  try {
    solver_new(P);
  } catch (const Error& e) {
    throw Error(std::string("Failed to initialise solver: ") + e.what());
  }
  if (P.cfg.workflow == "forward") run_forward(P);
  else if (P.cfg.workflow == "adjoint") run_adjoint(P);
  else if (P.cfg.workflow == "inversion") run_inversion(P);
  else throw Error(fmt("Workflow '%s' is not yet implemented in seiswg", P.cfg.workflow.c_str()));
  log_info("Done.");
}

void put_u32(Vfs& v, const char* key, uint32_t x) {
  v[key] = std::vector<char>(reinterpret_cast<const char*>(&x), reinterpret_cast<const char*>(&x) + 4);
}
void put_f32(Vfs& v, const char* key, float x) {
  v[key] = std::vector<char>(reinterpret_cast<const char*>(&x), reinterpret_cast<const char*>(&x) + 4);
}

}  // namespace

struct sw_vfs {
  Vfs files;
};

namespace sw {
void set_last_error(const std::string& msg);   // capi.cpp
}

extern "C" {

SW_API int seiswg_ricker_table(int32_t nsrc, int32_t nt, float dt, const float* params, float* out) {
  if (nsrc < 0 || nt < 0 || (nsrc > 0 && (!params || !out))) return -1;
  const float PI_F = 3.14159265358979323846f;
  const float RAD = PI_F / 180.0f;   // f32::to_radians
  const size_t stride = size_t(nsrc) * nt;
  for (int32_t s = 0; s < nsrc; ++s) {
    const float f0 = params[4 * s], t0 = params[4 * s + 1], ang = params[4 * s + 2], amp = params[4 * s + 3];
    const float pf = PI_F * f0;
    const float a = pf * pf;           // powi(2)
    const float ar = ang * RAD;
    const float sa = std::sin(ar), ca = std::cos(ar);
    for (int32_t it = 0; it < nt; ++it) {
      const float t = float(it) * dt;
      const float d = t - t0;
      const float st = -amp * (1.0f - 2.0f * a * d * d) * std::exp(-a * d * d);
      out[0 * stride + size_t(s) * nt + it] = st * sa;
      out[1 * stride + size_t(s) * nt + it] = st;
      out[2 * stride + size_t(s) * nt + it] = -st * ca;
    }
  }
  return 0;
}

SW_API int seiswg_infer_grid(int64_t npt, const float* x, const float* z, int32_t* nx, int32_t* nz, float* dx,
                             float* dz, int32_t* fallback) {
  if (npt < 4 || !x || !z || !nx || !nz || !dx || !dz) return -1;
  float xmax = -INFINITY, xmin = INFINITY, zmax = -INFINITY, zmin = INFINITY;
  for (int64_t k = 0; k < npt; ++k) {
    xmax = std::fmax(xmax, x[k]); xmin = std::fmin(xmin, x[k]);
    zmax = std::fmax(zmax, z[k]); zmin = std::fmin(zmin, z[k]);
  }
  const float lx = xmax - xmin, lz = zmax - zmin;
  auto u32r = [](double v) -> uint32_t { return (v != v || v <= 0) ? 0u : v >= 4294967295.0 ? 0xFFFFFFFFu : uint32_t(v); };
  uint32_t ax = u32r(std::round(std::sqrt(double(npt) * double(lx) / double(lz))));
  uint32_t az = u32r(std::round(std::sqrt(double(npt) * double(lz) / double(lx))));
  int32_t fb = 0;
  if (int64_t(ax) * int64_t(az) != npt) {
    // exact shape: z is the fast axis, so the first column is the run of equal x values
    int64_t run = 1;
    while (run < npt && x[run] == x[0]) ++run;
    if (npt % run != 0 || run < 2 || npt / run < 2) return -2;
    az = uint32_t(run);
    ax = uint32_t(npt / run);
    fb = 1;
  }
  if (ax < 2 || az < 2) return -2;
  *nx = int32_t(ax);
  *nz = int32_t(az);
  *dx = lx / float(ax - 1);
  *dz = lz / float(az - 1);
  if (fallback) *fallback = fb;
  return 0;
}

SW_API uint32_t seiswg_grid_index(float x, float z, float dx, float dz, uint32_t nz) {
  const uint32_t ix = sat_u32(std::round(x / dx));
  const uint32_t iz = sat_u32(std::round(z / dz));
  return ix * nz + iz;
}

SW_API int seiswg_run(const char* config_path) {
  try {
    if (!config_path) throw Error("no config path");
    log_info(fmt("Loading config from %s", config_path));
    Problem P;
    try {
      P.cfg = load_config(config_path);
    } catch (const Error& e) {
      throw Error(fmt("Failed to load %s: %s", config_path, e.what()));
    }
    run_workflow(P);
    return 0;
  } catch (const std::exception& e) {
    sw::set_last_error(e.what());
    fprintf(stderr, "Error: %s\n", e.what());
    return 1;
  }
}

SW_API sw_vfs* seiswg_vfs_new(void) {
  try {
    return new sw_vfs;
  } catch (const std::exception&) {
    return nullptr;
  }
}

SW_API void seiswg_vfs_free(sw_vfs* v) { delete v; }

SW_API int seiswg_vfs_put(sw_vfs* v, const char* key, const void* data, int64_t nbytes) {
  if (!v || !key || nbytes < 0 || (nbytes > 0 && !data)) {
    sw::set_last_error("seiswg_vfs_put: bad arguments");
    return -1;
  }
  try {
    const char* d = static_cast<const char*>(data);
    v->files[key] = std::vector<char>(d, d + nbytes);
  } catch (const std::exception& e) {
    sw::set_last_error(e.what());
    return -3;
  }
  return 0;
}

SW_API int64_t seiswg_vfs_get(const sw_vfs* v, const char* key, const void** data) {
  if (!v || !key) return -1;
  auto it = v->files.find(key);
  if (it == v->files.end()) return -1;
  if (data) *data = it->second.data();
  return int64_t(it->second.size());
}

SW_API int32_t seiswg_vfs_count(const sw_vfs* v) { return v ? int32_t(v->files.size()) : 0; }

SW_API const char* seiswg_vfs_key(const sw_vfs* v, int32_t i) {
  if (!v || i < 0 || size_t(i) >= v->files.size()) return nullptr;
  return std::next(v->files.begin(), i)->first.c_str();
}

SW_API int seiswg_run_simulation(const char* config_ini, const char* sources_dat, const char* stations_dat,
                                 const char* config_true_ini, const sw_vfs* files, sw_vfs* out) {
  try {
    if (!config_ini || !out) throw Error("seiswg_run_simulation: config text and output map are required");
    Vfs vfs;
    if (files) vfs = files->files;
    Problem P;
    P.cfg = load_config_text(config_ini, fs::path());
    // the tables given as text stand for the files the configs name (lib.rs:97-98)
    auto add_tables = [&](const Config& c) {
      if (sources_dat) vfs[c.sources.string()] = std::vector<char>(sources_dat, sources_dat + std::strlen(sources_dat));
      if (stations_dat)
        vfs[c.stations.string()] = std::vector<char>(stations_dat, stations_dat + std::strlen(stations_dat));
    };
    add_tables(P.cfg);
    // adjoint: a forward run on the true model first, whose traces join the
    // inputs (lib.rs:106-113)
    if ((P.cfg.workflow == "adjoint" || P.cfg.workflow == "inversion") && config_true_ini) {
      Problem T;
      T.cfg = load_config_text(config_true_ini, fs::path());
      add_tables(T.cfg);
      Vfs obs;
      T.io = Io{&vfs, &obs, false};
      run_workflow(T);
      for (auto& f : obs) vfs[f.first] = std::move(f.second);
    }
    Vfs result;
    P.io = Io{&vfs, &result, false};
    run_workflow(P);
    // the scalars lib.rs:126-131 / 166-170 return beside the files
    put_u32(result, "nt", P.cfg.nt);
    put_f32(result, "dt", P.cfg.dt);
    put_u32(result, "nrec", uint32_t(P.rec_k.size()));
    put_u32(result, "npt", uint32_t(P.npt));
    put_u32(result, "nx", uint32_t(P.nx));
    put_u32(result, "nz", uint32_t(P.nz));
    for (auto& f : result) out->files[f.first] = std::move(f.second);
    return 0;
  } catch (const std::exception& e) {
    sw::set_last_error(e.what());
    return 1;
  }
}

}  // extern "C"
