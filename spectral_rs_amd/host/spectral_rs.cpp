// spectral_rs.cpp -- host program mirroring the reference binary
// (src/main.rs:3-13, src/lib.rs:34-202, src/save/mod.rs:8-81) over the C ABI
// of include/spectral_rs.h.  Reads ./config.toml (or the path given as the
// first argument), validates it like Config::new, runs the loop of run() on
// the GPU and, when [output] write_output is set, writes the reference's
// output layout: output/dat_NNNNN/{x,y,u,gam}.npy and flds/{Jx,...,Dens}.npy.
//
// Keys beyond the reference's schema (all optional; the defaults reproduce
// the reference's hard-coded choices at src/lib.rs:124-126):
//   [params]   mass_ratio = 1.0, vth_ion = 1e-3, vth_lec = 1e-3, seed = 0,
//              precision = "f32" | "f64" (cargo feature dprec),
//              periodic = false (feature periodic), checked = true
//              (unchecked = !checked), left_edge = "assert" | "reflect",
//              sort_interval = 20 (steps between tile sorts of the particles;
//              0 = never), autotune = false (time the launch shapes first)
//   [output]   verbose = false (print the reference's per-step timer lines),
//              checkpoint_interval = 0 (every that many steps write
//              output/checkpoint_NNNNN.ckpt, the restartable state of
//              sp_checkpoint_save; 0 = never)
//   [setup]    restart = "path" (continue the run saved in that checkpoint
//              instead of initialising the particles; the loop resumes at
//              the saved step)
//
// With several config files (spectral_rs a/config.toml b/config.toml ...)
// the runs are independent members of one ensemble stepped together on the
// GPU (sp_ensemble_*); each writes its output under its config's directory
// (a/output/dat_NNNNN/...).  They must share t_final.
#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "spectral_rs.h"

namespace {

// ---- minimal TOML subset: [section], key = int | float | bool | "string" ----
struct Toml {
  std::map<std::string, std::string> kv;  // "section.key" -> raw value
  bool has(const std::string& k) const { return kv.count(k) != 0; }
};

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

bool parse_toml(const std::string& path, Toml& t, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "Could not open the config.toml file";
    return false;
  }
  std::string line, sec;
  int ln = 0;
  while (std::getline(f, line)) {
    ++ln;
    bool inq = false;
    for (size_t i = 0; i < line.size(); ++i) {  // strip comments outside strings
      if (line[i] == '"') inq = !inq;
      if (line[i] == '#' && !inq) {
        line = line.substr(0, i);
        break;
      }
    }
    line = trim(line);
    if (line.empty()) continue;
    if (line.front() == '[') {
      if (line.back() != ']') {
        err = "Could not parse Config file (line " + std::to_string(ln) + ")";
        return false;
      }
      sec = trim(line.substr(1, line.size() - 2));
      continue;
    }
    size_t eq = line.find('=');
    if (eq == std::string::npos) {
      err = "Could not parse Config file (line " + std::to_string(ln) + ")";
      return false;
    }
    t.kv[sec + "." + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  return true;
}

struct Config {  // src/lib.rs:34-66 plus the optional extensions
  uint32_t t_final = 0;
  int64_t size_x = 0, size_y = 0, delta = 0;
  double dt = 0, c = 0, gamma_inj = 0;
  int64_t dens = 0, n_pass = 0;
  bool track_prtls = false, write_output = false, verbose = false;
  int64_t track_interval = 0, output_interval = 1, stride = 1, istep = 1;
  double mass_ratio = 1.0, vth_ion = 1e-3, vth_lec = 1e-3;
  uint64_t seed = 0;
  bool f64 = false, periodic = false, checked = true, autotune = false;
  int left_edge = SP_LEFT_ASSERT;
  int64_t sort_interval = 20;
  int64_t checkpoint_interval = 0;
  std::string restart;             // checkpoint to continue from
  std::string out_dir = "output";  // where save_output writes
};

bool get_num(const Toml& t, const std::string& k, double& v, bool required, std::string& err) {
  if (!t.has(k)) {
    if (required) err = "Could not parse Config file: missing field `" + k.substr(k.find('.') + 1) + "`";
    return !required;
  }
  const std::string& s = t.kv.at(k);
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end) {
    err = "Could not parse Config file: bad value for `" + k + "`";
    return false;
  }
  return true;
}
bool get_int(const Toml& t, const std::string& k, int64_t& v, bool required, std::string& err) {
  double d = (double)v;
  if (!get_num(t, k, d, required, err)) return false;
  if (t.has(k) && (d != std::floor(d) || d < 0)) {
    err = "Could not parse Config file: `" + k + "` must be a non-negative integer";
    return false;
  }
  v = (int64_t)d;
  return true;
}
bool get_bool(const Toml& t, const std::string& k, bool& v, bool required, std::string& err) {
  if (!t.has(k)) {
    if (required) err = "Could not parse Config file: missing field `" + k.substr(k.find('.') + 1) + "`";
    return !required;
  }
  const std::string& s = t.kv.at(k);
  if (s == "true") v = true;
  else if (s == "false") v = false;
  else {
    err = "Could not parse Config file: bad bool for `" + k + "`";
    return false;
  }
  return true;
}
std::string get_str(const Toml& t, const std::string& k, const std::string& dflt) {
  if (!t.has(k)) return dflt;
  std::string s = t.kv.at(k);
  if (s.size() >= 2 && s.front() == '"' && s.back() == '"') s = s.substr(1, s.size() - 2);
  return s;
}

// Config::new (src/lib.rs:69-117)
bool load_config(const std::string& path, Config& c, std::string& err) {
  Toml t;
  if (!parse_toml(path, t, err)) return false;
  int64_t tf = 0;
  bool ok = get_int(t, "setup.t_final", tf, true, err) && get_int(t, "params.size_x", c.size_x, true, err) &&
            get_int(t, "params.size_y", c.size_y, true, err) && get_int(t, "params.delta", c.delta, true, err) &&
            get_num(t, "params.dt", c.dt, true, err) && get_num(t, "params.c", c.c, true, err) &&
            get_int(t, "params.dens", c.dens, true, err) && get_num(t, "params.gamma_inj", c.gamma_inj, true, err) &&
            get_int(t, "params.n_pass", c.n_pass, true, err) &&
            get_bool(t, "output.track_prtls", c.track_prtls, true, err) &&
            get_bool(t, "output.write_output", c.write_output, true, err) &&
            get_int(t, "output.track_interval", c.track_interval, true, err) &&
            get_int(t, "output.output_interval", c.output_interval, true, err) &&
            get_int(t, "output.stride", c.stride, true, err) && get_int(t, "output.istep", c.istep, true, err) &&
            get_bool(t, "output.verbose", c.verbose, false, err) &&
            get_num(t, "params.mass_ratio", c.mass_ratio, false, err) &&
            get_num(t, "params.vth_ion", c.vth_ion, false, err) && get_num(t, "params.vth_lec", c.vth_lec, false, err) &&
            get_bool(t, "params.periodic", c.periodic, false, err) && get_bool(t, "params.checked", c.checked, false, err) &&
            get_bool(t, "params.autotune", c.autotune, false, err) &&
            get_int(t, "params.sort_interval", c.sort_interval, false, err) &&
            get_int(t, "output.checkpoint_interval", c.checkpoint_interval, false, err);
  if (!ok) return false;
  c.restart = get_str(t, "setup.restart", "");
  int64_t seed = 0;
  if (!get_int(t, "params.seed", seed, false, err)) return false;
  c.seed = (uint64_t)seed;
  c.t_final = (uint32_t)tf;
  const std::string prec = get_str(t, "params.precision", "f32");
  if (prec != "f32" && prec != "f64") {
    err = "Could not parse Config file: precision must be \"f32\" or \"f64\"";
    return false;
  }
  c.f64 = prec == "f64";
  const std::string le = get_str(t, "params.left_edge", "assert");
  if (le != "assert" && le != "reflect") {
    err = "Could not parse Config file: left_edge must be \"assert\" or \"reflect\"";
    return false;
  }
  c.left_edge = le == "reflect" ? SP_LEFT_REFLECT : SP_LEFT_ASSERT;
  auto which = [](bool xo, bool yo) { return !xo && !yo ? "x & y" : (!xo ? "x" : "y"); };
  if (c.size_x % 2 != 0 || c.size_y % 2 != 0) {
    err = std::string("Number of cells in ") + which(c.size_x % 2 == 0, c.size_y % 2 == 0) +
          " direction must be even";
    return false;
  }
  if (c.delta * 2 >= c.size_x) {
    err = "Delta must be less than 1/2 the size of size_x";
    return false;
  }
  if (c.gamma_inj <= 1.0) {
    err = "Lorentz factor must be greater than 1";
    return false;
  }
  if (c.istep == 0 || c.size_x % c.istep != 0 || c.size_y % c.istep != 0) {
    err = std::string("Number of cells in ") +
          which(c.istep && c.size_x % c.istep == 0, c.istep && c.size_y % c.istep == 0) +
          " direction must be divisible by 2";
    return false;
  }
  if (c.output_interval == 0 || c.stride == 0) {
    err = "output_interval and stride must be positive";
    return false;
  }
  return true;
}

// ---- .npy writer (1-D little-endian float arrays, as the npy crate writes) ----
bool write_npy(const std::string& path, const void* data, size_t n, bool f64) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::string hdr = std::string("{'descr': '") + (f64 ? "<f8" : "<f4") +
                    "', 'fortran_order': False, 'shape': (" + std::to_string(n) + ",), }";
  size_t total = 10 + hdr.size() + 1;
  size_t pad = (64 - total % 64) % 64;
  hdr += std::string(pad, ' ') + "\n";
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const uint16_t hl = (uint16_t)hdr.size();
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&hl, 2, 1, f) == 1 &&
            std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
            std::fwrite(data, f64 ? 8 : 4, n, f) == n;
  return std::fclose(f) == 0 && ok;
}

bool mkdirs(const std::string& p) {
  std::string cur;
  std::stringstream ss(p);
  std::string part;
  while (std::getline(ss, part, '/')) {
    cur += part + "/";
    if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
  }
  return true;
}

template <class F> std::vector<F> strided(const std::vector<F>& v, size_t stride) {
  std::vector<F> out;
  for (size_t i = 0; i < v.size(); i += stride) out.push_back(v[i]);
  return out;
}

// save_output (src/save/mod.rs:25-81), species 0 only like the reference.
template <class F> int save_output(sp_ctx* ctx, const Config& cfg, uint32_t t, std::string& err) {
  if (t % cfg.output_interval != 0) return 0;
  char pre[64];
  std::snprintf(pre, sizeof pre, "/dat_%05u", (unsigned)(t / cfg.output_interval));
  const std::string prefix = cfg.out_dir + pre;
  if (!mkdirs(prefix + "/flds")) {
    err = "Unable to create output directory";
    return 1;
  }
  std::printf("saving prtls\n");
  int64_t n = 0;
  sp_species_count(ctx, 0, &n);
  std::vector<int32_t> ix(n), iy(n);
  std::vector<F> dx(n), dy(n), px(n), py(n), pz(n), psa(n);
  if (sp_get_species(ctx, 0, n, ix.data(), iy.data(), dx.data(), dy.data(), px.data(), py.data(), pz.data(),
                     psa.data()) != SP_OK) {
    err = sp_last_error(ctx);
    return 1;
  }
  std::vector<F> x(n), y(n), u(n);
  for (int64_t i = 0; i < n; ++i) {
    x[i] = (F)ix[i] + dx[i];  // save/mod.rs:39
    y[i] = (F)iy[i] + dy[i];
    u[i] = px[i] / (F)cfg.c;  // save/mod.rs:58
  }
  const bool f64 = sizeof(F) == 8;
  const size_t st = (size_t)cfg.stride;
  struct Out {
    const char* name;
    std::vector<F> v;
  } outs[4] = {{"x", strided(x, st)}, {"y", strided(y, st)}, {"u", strided(u, st)}, {"gam", strided(psa, st)}};
  for (auto& o : outs)
    if (!write_npy(prefix + "/" + o.name + ".npy", o.v.data(), o.v.size(), f64)) {
      err = std::string("Could not save ") + o.name + " data to file";
      return 1;
    }
  // save_fld_spatial (save/mod.rs:8-23): for each interior row, Nx values
  // starting at the row's first (ghost) column, as the reference's
  // skip(ij_ghosts).take(size_x) does.
  const int ids[10] = {SP_JX, SP_JY, SP_JZ, SP_EX, SP_EY, SP_EZ, SP_BX, SP_BY, SP_BZ, SP_DENS};
  const char* names[10] = {"Jx", "Jy", "Jz", "Ex", "Ey", "Ez", "Bx", "By", "Bz", "Dens"};
  const size_t P = (size_t)cfg.size_x + 2;
  std::vector<F> g(P * (cfg.size_y + 2)), o;
  for (int k = 0; k < 10; ++k) {
    if (sp_get_grid(ctx, ids[k], g.data(), g.size()) != SP_OK) {
      err = sp_last_error(ctx);
      return 1;
    }
    o.clear();
    for (int64_t r = 0; r < cfg.size_y; ++r)
      o.insert(o.end(), g.begin() + (r + 1) * P, g.begin() + (r + 1) * P + cfg.size_x);
    if (!write_npy(prefix + "/flds/" + names[k] + ".npy", o.data(), o.size(), f64)) {
      err = std::string("Could not save ") + names[k] + " data to file";
      return 1;
    }
  }
  return 0;
}

#define CHECK(call)                                                        \
  do {                                                                     \
    int rc_ = (call);                                                      \
    if (rc_ != SP_OK) {                                                    \
      std::printf("Application error: %s\n", sp_last_error(ctx));        \
      sp_destroy(ctx);                                                     \
      return 1;                                                            \
    }                                                                      \
  } while (0)

// Flds::new + Sim::new + the two Prtl::new of run() (src/lib.rs:119-142):
// context, species, pre-loop move & deposit; optionally the autotuner.  With
// a restart checkpoint the context takes the saved state instead.
int setup(const Config& cfg, sp_ctx** out) {
  sp_params p{};
  p.size_x = (int32_t)cfg.size_x;
  p.size_y = (int32_t)cfg.size_y;
  p.delta = (int32_t)cfg.delta;
  p.dens = (int32_t)cfg.dens;
  p.n_pass = (int32_t)cfg.n_pass;
  p.dt = cfg.dt;
  p.c = cfg.c;
  p.gamma_inj = cfg.gamma_inj;
  p.precision = cfg.f64 ? SP_F64 : SP_F32;
  p.periodic_x = cfg.periodic;
  p.checked = cfg.checked;
  p.left_edge = cfg.left_edge;
  p.sort_interval = (int32_t)cfg.sort_interval;
  sp_ctx* ctx = nullptr;
  if (sp_create(&p, &ctx) != SP_OK) {
    std::printf("Application error: %s\n", sp_last_error(nullptr));
    return 1;
  }
  if (!cfg.restart.empty()) {
    CHECK(sp_checkpoint_load(ctx, cfg.restart.c_str()));
    std::printf("restarted from %s\n", cfg.restart.c_str());
    *out = ctx;
    return 0;
  }
  std::printf("initialzing  prtls\n");
  int sid = 0;
  CHECK(sp_species_init(ctx, 1.0, cfg.mass_ratio, cfg.vth_ion, cfg.seed, &sid));  // ions (lib.rs:124)
  CHECK(sp_species_init(ctx, -1.0, 1.0, cfg.vth_lec, cfg.seed, &sid));           // electrons (lib.rs:126)
  if (cfg.autotune) {
    sp_tuning tu;
    CHECK(sp_autotune(ctx, 3, 1, &tu));
    std::printf("autotuned: chunk %d, sort tile %dx%d, FFT rows %d (radix %d), columns %d (radix %d)\n", tu.chunk,
                1 << tu.tile_w_log, 1 << tu.tile_h_log, tu.fft_rows, tu.fft_radix_x, tu.fft_cols, tu.fft_radix_y);
  }
  CHECK(sp_prestep(ctx));  // lib.rs:130-142
  *out = ctx;
  return 0;
}

int save(sp_ctx* ctx, const Config& cfg, uint32_t t) {
  if (!cfg.write_output) return 0;
  std::string err;
  const int rc = cfg.f64 ? save_output<double>(ctx, cfg, t, err) : save_output<float>(ctx, cfg, t, err);
  if (rc) std::printf("Application error: %s\n", err.c_str());
  return rc;
}

// run (src/lib.rs:119-202)
int run(const Config& cfg) {
  sp_ctx* ctx = nullptr;
  if (setup(cfg, &ctx)) return 1;
  sp_diag d0{};
  CHECK(sp_get_diag(ctx, &d0));  // 0, or the step a restart resumes at
  const uint32_t t0 = (uint32_t)d0.steps;
  CHECK(sp_set_timing(ctx, 1));
  double ms[4];
  for (uint32_t t = t0; t <= cfg.t_final; ++t) {
    if (save(ctx, cfg, t)) {
      sp_destroy(ctx);
      return 1;
    }
    if (cfg.checkpoint_interval > 0 && t > t0 && t % cfg.checkpoint_interval == 0) {
      char name[64];
      std::snprintf(name, sizeof name, "/checkpoint_%05u.ckpt", (unsigned)t);
      if (!mkdirs(cfg.out_dir)) {
        std::printf("Application error: Unable to create output directory\n");
        sp_destroy(ctx);
        return 1;
      }
      CHECK(sp_checkpoint_save(ctx, (cfg.out_dir + name).c_str()));
    }
    CHECK(sp_step(ctx, 1));
    if (cfg.verbose) {
      CHECK(sp_sync(ctx));
      CHECK(sp_get_timers(ctx, ms, 1));
      std::printf("%u\nmoving & dep prtl\n%.3fms\nsolving fields\n%.3fms\npushing prtl\n%.3fms\n", t, ms[0], ms[1],
                  ms[2]);
    }
  }
  CHECK(sp_sync(ctx));
  CHECK(sp_get_timers(ctx, ms, 1));
  std::printf("done: %u steps; dep %.3f ms, solve %.3f ms, push %.3f ms (totals over the un-reported steps)\n",
              cfg.t_final + 1 - t0, ms[0], ms[1], ms[2]);
  sp_destroy(ctx);
  return 0;
}

// Several runs as one ensemble: the same loop for every member, stepped
// together between the output steps of any member.
int run_ensemble(std::vector<Config>& cfgs) {
  std::vector<sp_ctx*> ctx;
  auto cleanup = [&](sp_ensemble* e) {
    if (e) sp_ensemble_destroy(e);
    for (sp_ctx* c : ctx) sp_destroy(c);
  };
  for (auto& c : cfgs) {
    sp_ctx* x = nullptr;
    if (setup(c, &x)) {
      cleanup(nullptr);
      return 1;
    }
    ctx.push_back(x);
  }
  sp_ensemble* e = nullptr;
  if (sp_ensemble_create(ctx.data(), (int)ctx.size(), 4, 2, &e) != SP_OK) {
    std::printf("Application error: %s\n", sp_ensemble_last_error(nullptr));
    cleanup(nullptr);
    return 1;
  }
  const uint32_t tf = cfgs[0].t_final;
  uint32_t t = 0;
  while (t <= tf) {
    uint32_t next = tf + 1;
    for (size_t k = 0; k < cfgs.size(); ++k) {
      const Config& c = cfgs[k];
      if (save(ctx[k], c, t)) {
        cleanup(e);
        return 1;
      }
      if (c.write_output) {
        const uint32_t nt = (t / c.output_interval + 1) * c.output_interval;
        if (nt < next) next = nt;
      }
    }
    if (sp_ensemble_step(e, (int)(next - t)) != SP_OK) {
      std::printf("Application error: %s\n", sp_ensemble_last_error(e));
      cleanup(e);
      return 1;
    }
    t = next;
  }
  int bad = -1;
  if (sp_ensemble_sync(e, &bad) != SP_OK) {
    std::printf("Application error: %s\n", sp_ensemble_last_error(e));
    cleanup(e);
    return 1;
  }
  std::printf("done: %zu runs x %u steps\n", cfgs.size(), tf + 1);
  cleanup(e);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {  // src/main.rs:3-13
  std::vector<Config> cfgs;
  const int nfiles = argc > 1 ? argc - 1 : 1;
  for (int i = 0; i < nfiles; ++i) {
    const std::string path = argc > 1 ? argv[1 + i] : "config.toml";
    Config cfg;
    std::string err;
    if (!load_config(path, cfg, err)) {
      std::printf("Problem parsing arguments: %s\n", err.c_str());
      return 1;
    }
    if (nfiles > 1) {  // each member writes beside its config file
      const size_t slash = path.find_last_of('/');
      cfg.out_dir = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/output";
      if (!cfg.restart.empty()) {
        std::printf("Problem parsing arguments: restart is for a single run, not an ensemble member\n");
        return 1;
      }
      if (!cfgs.empty() && cfg.t_final != cfgs[0].t_final) {
        std::printf("Problem parsing arguments: ensemble members must share t_final\n");
        return 1;
      }
    }
    cfgs.push_back(cfg);
  }
  return nfiles > 1 ? run_ensemble(cfgs) : run(cfgs[0]);
}
