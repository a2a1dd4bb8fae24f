// api.hip -- the C ABI of include/spectral_rs.h: context lifetime, buffers,
// host<->device copies, and the order of kernel launches that reproduces the
// reference's run() loop (src/lib.rs:119-202).
//
// Particle passes: run() calls boris_push at the end of iteration t and
// move_and_deposit at the start of iteration t+1 with nothing in between that
// touches particles (src/lib.rs:171-196).  sp_step therefore leaves the push
// of its last iteration pending and performs it fused into the next
// iteration's move&deposit pass; any other entry point (and sp_sync) first
// runs a pending push on its own, so every observable state matches the
// reference's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sp_internal.h"

using namespace sp;

namespace {

thread_local std::string g_create_error;

struct SoA {
  int32_t* cell = nullptr;
  void *dx = nullptr, *dy = nullptr, *px = nullptr, *py = nullptr, *pz = nullptr;
  int32_t* tag = nullptr;
};

struct SpeciesBuf {
  int64_t n = 0;
  double charge = 0, mass = 0;
  SoA cur, alt;  // alt is the target of the tile sort; the two swap after it
  // the species' own density scratch grid for the step, whose passes all
  // deposit before one fold adds every species' density to rho and dens
  void* dsp = nullptr;
};

struct TimerPair {
  hipEvent_t a, b;
  int phase;
};

}  // namespace

struct sp_ctx {
  sp_params p{};
  bool f64 = false;
  size_t esz = 4;
  hipStream_t st = nullptr;      // the stream every launch of this context goes to
  hipStream_t own_st = nullptr;  // created by sp_create; st differs from it inside an ensemble
  sp_ensemble* ens = nullptr;    // the ensemble this context belongs to, if any
  sp_tuning tune{};              // launch shapes (sp_set_tuning / sp_autotune)
  int64_t tune_gen = 0;          // bumped on every change of launch arguments (graph signatures)
  int device = 0;
  void* grid[11] = {};
  void* work[4] = {};
  void* dsp = nullptr;  // per-species density scratch grid (GridSet::dsp)
  void* spec[14] = {};
  FftDim fx{}, fy{};
  // spectra (bit = grid id) of which the step's real-field pipeline left only
  // the half plane kx = 0 .. nx/2 valid; full_spectrum() fills the rest
  uint32_t half_spec = 0;
  int* flags_d = nullptr;  // [0] invariant flags; [2..3] the slow-path particle counter (SLOW_COUNTER)
  int* flags_h = nullptr;
  // adaptive tile-sort schedule (sp_set_sort_policy): off while sort_max == 0
  int sort_min = 0, sort_max = 0;
  double sort_rise = 0;
  int64_t last_sort = -1;            // step whose particle pass ran right after the last sort
  int64_t nsorts = 0;                // tile sorts run (sp_diag.sorts)
  uint64_t* slow_h = nullptr;        // pinned snapshots of the running slow-path counter, one per step
  hipEvent_t slow_ev[SLOW_RING] = {};
  int64_t slow_step[SLOW_RING];     // step of each snapshot (-2: none)
  double slow_f0 = -1;               // slow fraction of step last_sort, once known
  int* tile_count = nullptr;   // per-tile particle counts of the tile sort
  int* tile_cursor = nullptr;  // per-tile write positions
  double* gauss_d = nullptr;   // Gauss-law sums of the last solve + block partials
  // learned denoiser (sp_set_denoiser): parameters, input/correction planes, sums, counters
  bool dn_on = false;
  float* dn_buf = nullptr;  // parameters + packed weights + r (4 planes)
  double* dn_sums = nullptr;
  int* dn_stats = nullptr;
  float dn_tol = 0, dn_gtol = 0;
  int dn_mask = 0xF;  // channels the correction is written to (bit c: Jx Jy Jz rho)
  int dn_backoff = 0;  // solves skipped after a failed check (sp_set_denoiser_backoff)
  bool dn_raw = false;  // the network reads the unfiltered deposit (it replaces the filter)
  float dn_hp = 0;      // spectrum check: high-pass energy bound relative to the filter's (0 = off)
  // step-loop graph (sp_set_graph): the captured period and the buffers it uses
  bool graph_on = false;
  hipGraphExec_t gexec = nullptr;
  std::vector<const void*> gsig;
  double* diag_d = nullptr;    // energy / momentum sums of sp_get_diag
  std::vector<SpeciesBuf> species;
  std::string err;
  bool push_pending = false;
  int64_t steps = 0;
  // phase timers (dep, solve, push, particle pass)
  bool timing = false;
  std::vector<TimerPair> open_pairs;
  std::vector<hipEvent_t> free_events;
  double t_ms[4] = {0, 0, 0, 0};
  size_t ngrid_elems() const { return (size_t)(p.size_x + 2) * (p.size_y + 2); }
  size_t nspec_elems() const { return (size_t)p.size_x * p.size_y; }
};

namespace {

sp_tuning default_tuning(bool f64, int /*dens*/) {
  sp_tuning t{};
  t.chunk = f64 ? 1024 : 2048;
  // FP32 sort tiles: 8x4 at every density.  4096x1024 shock at 32 ppc, whole
  // 2000-step run on the adaptive sort schedule: 5.15 ms/step with 8x4 against
  // 5.26 with 16x2, 5.42 with 8x2 and 16x4, 6.15 with 32x2.  At 4 ppc a 2-row
  // footprint 256+ cells wide would outgrow the deposit window once particles
  // spread a few rows (4.1 vs 0.86 ms per particle pass with 16x2).
  t.tile_w_log = 3;
  t.tile_h_log = f64 ? 3 : 2;
  t.fft_rows = 1;
  t.fft_cols = 8;
  t.fft_radix_x = 8;
  t.fft_radix_y = 8;
  return t;
}

int fail(sp_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(ctx, SP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

template <class T> Geo<T> geo(const sp_ctx* c) {
  Geo<T> g;
  g.nx = c->p.size_x;
  g.ny = c->p.size_y;
  g.pitch = c->p.size_x + 2;
  g.delta = c->p.delta;
  g.wall = c->p.size_x + 1 - c->p.delta;
  g.lwall = c->p.delta + 1;
  g.periodic_x = c->p.periodic_x;
  g.checked = c->p.checked;
  g.left_edge = c->p.left_edge;
  g.lds_ok = c->p.dt * c->p.c < 1.0;
  g.direct = c->p.sort_interval > 0;
  g.chunk = c->tune.chunk;
  g.abs_pos32 = (c->p.study & SP_STUDY_ABS_POS32) != 0;
  g.dt = (T)c->p.dt;
  g.c = (T)c->p.c;
  g.csqinv = T(1.) / (g.c * g.c);
  g.half_dt = T(0.5) * g.dt;
  return g;
}

template <class T> SpeciesDev<T> sdev(const sp_ctx* c, const SpeciesBuf& s, const SoA& a) {
  SpeciesDev<T> d;
  d.n = s.n;
  d.cell = a.cell;
  d.dx = (T*)a.dx;
  d.dy = (T*)a.dy;
  d.px = (T*)a.px;
  d.py = (T*)a.py;
  d.pz = (T*)a.pz;
  d.tag = a.tag;
  const T q = (T)s.charge, m = (T)s.mass;
  const T dt = (T)c->p.dt, cc = (T)c->p.c;
  d.charge = q;
  d.beta = (q / m) * T(0.5) * dt;          // src/prtls.rs:61
  d.alpha = (q / m) * T(0.5) * dt / cc;    // src/prtls.rs:62
  return d;
}
template <class T> SpeciesDev<T> sdev(const sp_ctx* c, const SpeciesBuf& s) { return sdev<T>(c, s, s.cur); }

GridSet gset(const sp_ctx* c) {
  GridSet g;
  for (int i = 0; i < 11; ++i) g.g[i] = c->grid[i];
  g.dsp = c->dsp;
  return g;
}

// Radix plan of one FFT dimension: the largest radices first, at most
// `maxr` (8 or 4), then 2 and 3.
bool plan_dim(FftDim& d, int n, int maxr = 8) {
  d.n = n;
  d.nrad = 0;
  d.radix_code = 0;
  int m = n;
  auto push = [&](int r) {
    if (d.nrad < 16) d.radix_code |= (uint64_t)r << (4 * d.nrad);
    d.nrad++;
    m /= r;
  };
  while (maxr >= 8 && m % 8 == 0) push(8);
  while (m % 4 == 0) push(4);
  while (m % 2 == 0) push(2);
  while (m % 3 == 0) push(3);
  return m == 1 && d.nrad <= 16;
}

// ---- phase timers ----
hipEvent_t take_event(sp_ctx* c) {
  if (!c->free_events.empty()) {
    hipEvent_t e = c->free_events.back();
    c->free_events.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
struct PhaseScope {
  sp_ctx* c;
  int phase;
  hipEvent_t a = nullptr;
  PhaseScope(sp_ctx* c_, int ph) : c(c_), phase(ph) {
    if (c->timing) {
      a = take_event(c);
      if (a) hipEventRecord(a, c->st);
    }
  }
  ~PhaseScope() {
    if (c->timing && a) {
      hipEvent_t b = take_event(c);
      if (b) {
        hipEventRecord(b, c->st);
        c->open_pairs.push_back({a, b, phase});
      }
    }
  }
};
void harvest_timers(sp_ctx* c) {  // after a stream sync
  for (auto& tp : c->open_pairs) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, tp.a, tp.b) == hipSuccess) c->t_ms[tp.phase] += ms;
    c->free_events.push_back(tp.a);
    c->free_events.push_back(tp.b);
  }
  c->open_pairs.clear();
}

// ---- launch sequences, templated on Float ----
template <class T> void zero_currents(sp_ctx* c) {
  T* g[5] = {(T*)c->grid[SP_JX], (T*)c->grid[SP_JY], (T*)c->grid[SP_JZ], (T*)c->grid[SP_RHO],
             (T*)c->grid[SP_DENS]};
  fill_many<T>(g, 5, T(0), (int64_t)c->ngrid_elems(), c->st);
}

template <class T> void fold(sp_ctx* c) {
  T* g[5] = {(T*)c->grid[SP_JX], (T*)c->grid[SP_JY], (T*)c->grid[SP_JZ], (T*)c->grid[SP_RHO],
             (T*)c->grid[SP_DENS]};
  launch_fold_ghosts<T>(g, 5, c->p.size_x, c->p.size_y, c->st);
}

// The spectra the step's pipeline leaves on the half plane kx = 0 .. nx/2:
// the forward transforms of J and rho (the forward columns write that half
// only).  E and B are transformed back in place after the update, so their
// buffers no longer hold spectra afterwards (the reference's in-place inverse,
// src/flds/mod.rs:280-290, leaves its E and B spectral arrays transformed
// back as well).
constexpr uint32_t HALF_SPECTRA = 1u << SP_JX | 1u << SP_JY | 1u << SP_JZ | 1u << SP_RHO;

// Makes spectrum `id` whole again (the other half as conjugate mirrors of the
// half plane) if the pipeline left only the half plane.
void full_spectrum(sp_ctx* c, int id) {
  if (!((c->half_spec >> id) & 1u)) return;
  if (c->f64) launch_mirror_fill<double>(c->spec[id], c->p.size_x, c->p.size_y, c->st);
  else launch_mirror_fill<float>(c->spec[id], c->p.size_x, c->p.size_y, c->st);
  c->half_spec &= ~(1u << id);
}
void full_spectra(sp_ctx* c, uint32_t mask) {
  for (int id = 0; id < 14; ++id)
    if ((mask >> id) & 1u) full_spectrum(c, id);
}

template <class T> void flush_push(sp_ctx* c) {
  if (!c->push_pending) return;
  PhaseScope ps(c, 2);
  const Geo<T> g = geo<T>(c);
  const GridSet gs = gset(c);
  for (auto& s : c->species) launch_push<T>(sdev<T>(c, s), g, gs, c->flags_d, c->st);
  c->push_pending = false;
}

DnTensors dn_tensors(const sp_ctx* c);

template <class T> void fields_update(sp_ctx* c) {
  const int nx = c->p.size_x, ny = c->p.size_y;
  // mod.rs:90-94 (Jx Jy Jz rho are consecutive grid ids)
  static_assert(SP_JY == SP_JX + 1 && SP_JZ == SP_JX + 2 && SP_RHO == SP_JX + 3, "grid ids");
  launch_binomial<T>(&c->grid[SP_JX], c->work, 4, nx, ny, c->p.n_pass, c->st);
  if constexpr (sizeof(T) == 4) {
    if (c->dn_on) {  // the learned currents, or the fallback: the filtered ones
      float* g4[4] = {(float*)c->grid[SP_JX], (float*)c->grid[SP_JY], (float*)c->grid[SP_JZ], (float*)c->grid[SP_RHO]};
      // after the filter's pointer swap the work grids hold its input, the
      // unfiltered deposit (one filter launch: n_pass <= 8, checked when set)
      float* w4[4] = {(float*)c->work[0], (float*)c->work[1], (float*)c->work[2], (float*)c->work[3]};
      const bool raw = c->dn_raw && c->p.n_pass > 0;
      dn_apply_sim(dn_params(c->dn_buf), dn_tensors(c), raw ? w4 : g4, g4, nx, ny, c->p.dens, c->dn_sums, c->gauss_d,
                   c->dn_tol, c->dn_gtol, c->dn_hp, c->dn_mask, c->dn_backoff, c->dn_stats, c->st);
      launch_update_ghosts<float>(g4, 4, nx, ny, c->st);
    }
  }
  // Every field transformed here is real and nx, ny are even (sp_create), so
  // the step runs a real-field pipeline: rows two at a time into half spectra,
  // columns kx = 0 .. nx/2, the update on that half plane (and B_wrk on both
  // halves), and back.  The other half of the stored spectra is filled, as
  // exact conjugate mirrors, only when read through the API (full_spectrum).
  // (With even sizes the Nyquist column nx/2 and row ny/2 that the update
  // zeroes, mod.rs:230-266, are their own mirrors.)
  {  // mod.rs:97-103, 122-124: 7 forward transforms, spatial->spectral fused
    FftBatch b{};
    const int ids[7] = {SP_JX, SP_JY, SP_JZ, SP_EX, SP_EY, SP_EZ, SP_RHO};
    b.count = 7;
    for (int i = 0; i < 7; ++i) {
      b.spec[i] = c->spec[ids[i]];
      b.grid[i] = c->grid[ids[i]];
      b.scale[i] = 1.0;
    }
    b.scale[6] = (double)(T(1.0) / (T)c->p.dens);  // mod.rs:80-84
    launch_fft_rows_r2c<T>(b, c->fx, ny, c->st);
    launch_fft_cols<T>(b, c->fy, nx, 0, 1.0, c->st, nx / 2 + 1, false);
  }
  c->half_spec |= HALF_SPECTRA;
  // mod.rs:105-278, with the forward spectra (transforms of real fields) kept
  // exactly Hermitian (fields.hip)
  const int gblk = launch_spectral_update<T>(c->spec, geo<T>(c), c->p.dens, c->gauss_d, c->st, true, false);
  {  // mod.rs:280-307: 6 inverse transforms, spectral->spatial + ghosts fused
    FftBatch b{};
    const int ids[6] = {SP_BX, SP_BY, SP_BZ, SP_EX, SP_EY, SP_EZ};
    b.count = 6;
    b.gauss = c->gauss_d;  // the inverse columns' first block sums the update's Gauss partials
    b.gauss_nblk = gblk;
    for (int i = 0; i < 6; ++i) {
      b.spec[i] = c->spec[ids[i]];
      b.grid[i] = c->grid[ids[i]];
      b.scale[i] = 1.0;
    }
    const T norm = T(1) / (T)((int64_t)nx * ny);  // fft_2d.rs:170
    launch_fft_cols<T>(b, c->fy, nx, 1, 1.0, c->st, nx / 2 + 1, false);
    launch_fft_rows_c2r<T>(b, c->fx, ny, (double)norm, c->st);
  }
}

// FP64 sums for sp_get_diag into diag_d: [0] sum E^2, [1] sum B^2 over the
// interior, then per species [sum (gamma - 1), sum px, sum py, sum pz].
template <class T> void diag_sums(sp_ctx* c) {
  const int nd = DIAG_FIXED + 4 * (int)c->species.size();
  fill<double>(c->diag_d, 0.0, nd, c->st);
  const int nx = c->p.size_x, ny = c->p.size_y;
  const bool s32 = (c->p.study & SP_STUDY_SUM32) != 0;
  launch_sumsq3<T>((const T*)c->grid[SP_EX], (const T*)c->grid[SP_EY], (const T*)c->grid[SP_EZ], nx, ny,
                   c->diag_d + 0, c->st, s32);
  launch_sumsq3<T>((const T*)c->grid[SP_BX], (const T*)c->grid[SP_BY], (const T*)c->grid[SP_BZ], nx, ny,
                   c->diag_d + 1, c->st, s32);
  const double csqinv = 1.0 / ((double)c->p.c * c->p.c);
  for (size_t s = 0; s < c->species.size(); ++s)
    launch_kinetic<T>(sdev<T>(c, c->species[s]), csqinv, c->diag_d + DIAG_FIXED + 4 * s, c->st, s32);
}

// Buffers of the in-loop denoiser after the parameters in dn_buf: the packed
// BF16 weights and the correction r (4 channels; the fused kernel reads its
// input from the grids and keeps the rest in LDS).
DnTensors dn_tensors(const sp_ctx* c) {
  DnTensors t{};
  t.packed = (short*)(c->dn_buf + DN_NPARAMS);  // 16-byte aligned: DN_NPARAMS % 4 == 0
  t.r = c->dn_buf + DN_NPARAMS + DN_PACK_SHORTS / 2;
  return t;
}
static_assert(DN_NPARAMS % 4 == 0 && DN_PACK_SHORTS % 8 == 0, "alignment of the packed weights");

// ---- step-loop graph ----
// A period of steps whose buffer swaps cancel: the filter swaps grid and work
// buffers once a step, a sort swaps a species' SoA buffers, so 2 sort
// intervals (2 steps without sorting) return every pointer to where it was.
int graph_period(const sp_ctx* c) { return c->p.sort_interval > 0 ? 2 * c->p.sort_interval : 2; }

std::vector<const void*> graph_signature(const sp_ctx* c) {
  std::vector<const void*> v(c->grid, c->grid + 11);
  v.insert(v.end(), c->work, c->work + 4);
  for (const auto& s : c->species) {
    v.push_back(s.cur.cell);
    v.push_back(s.alt.cell);
    v.push_back((const void*)(intptr_t)s.n);
  }
  return v;
}

void drop_graph(sp_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->gsig.clear();
}

template <class T> void one_step(sp_ctx* c);

// Runs one period through the graph, capturing it first when there is none
// for the current buffers.  Capture advances the host-side state (step count,
// swaps) while recording; a replay advances only the step count, its swaps
// cancelling.  The fills inside the step are kernels (no memset nodes).
template <class T> int graph_period_run(sp_ctx* c) {
  const int P = graph_period(c);
  const std::vector<const void*> sig = graph_signature(c);
  if (!c->gexec || sig != c->gsig) {
    drop_graph(c);
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < P; ++i) one_step<T>(c);
    hipGraph_t graph = nullptr;
    HIPCHK(c, hipStreamEndCapture(c->st, &graph));
    const hipError_t e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      c->gexec = nullptr;
      return fail(c, SP_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    if (graph_signature(c) != sig) {  // cannot happen by the period's construction
      drop_graph(c);
      return fail(c, SP_EHIP, "captured period does not return the buffers to their start");
    }
    c->gsig = sig;
  } else {
    c->steps += P;
  }
  HIPCHK(c, hipGraphLaunch(c->gexec, c->st));
  return SP_OK;
}

// Tile-sorts every species (particle order only; no value changes).
template <class T> void sort_species(sp_ctx* c) {
  for (auto& s : c->species) {
    launch_tile_sort<T>(sdev<T>(c, s, s.cur), sdev<T>(c, s, s.alt), c->p.size_x, c->p.size_y, c->tune.tile_w_log,
                        c->tune.tile_h_log, c->tile_count, c->tile_cursor, c->st);
    std::swap(s.cur, s.alt);
  }
}

bool sort_due(const sp_ctx* c) { return c->p.sort_interval > 0 && c->steps % c->p.sort_interval == 0; }

// The adaptive schedule applies to a context stepped on its own, without the
// step-loop graph (captured periods and ensembles replay a fixed sort phase).
bool adaptive_sort(const sp_ctx* c) { return c->sort_max > 0 && !c->ens && !c->graph_on; }

// Snapshot of the slow-path counter after step `step`'s particle passes
// (stream-ordered copy into pinned memory; read back two steps later).
void slow_snapshot(sp_ctx* c, int64_t step) {
  const int k = (int)(((step % SLOW_RING) + SLOW_RING) % SLOW_RING);
  hipMemcpyAsync(&c->slow_h[k], c->flags_d + SLOW_COUNTER, sizeof(uint64_t), hipMemcpyDeviceToHost, c->st);
  hipEventRecord(c->slow_ev[k], c->st);
  c->slow_step[k] = step;
}

// Particles of step `step` that took the general path, as a fraction of all
// particles; -1 when the snapshots of `step` and `step - 1` are gone.
double slow_fraction(sp_ctx* c, int64_t step) {
  const int k1 = (int)(((step % SLOW_RING) + SLOW_RING) % SLOW_RING);
  const int k0 = (int)((((step - 1) % SLOW_RING) + SLOW_RING) % SLOW_RING);
  if (c->slow_step[k1] != step || c->slow_step[k0] != step - 1) return -1;
  hipEventSynchronize(c->slow_ev[k1]);
  int64_t n = 0;
  for (const auto& sp : c->species) n += sp.n;
  if (n <= 0) return 0;
  return (double)(c->slow_h[k1] - c->slow_h[k0]) / (double)n;
}

// Whether step c->steps sorts first.  Fixed schedule: every sort_interval
// steps.  Adaptive: at least sort_min and at most sort_max steps after the
// last sort, and in between when the fraction of particles the particle pass
// took through its general path (chunks whose cells spread beyond the LDS
// windows, or near a boundary) has risen by more than sort_rise since the pass
// right after the last sort.  The fraction is read two steps late so that the
// host does not wait for the step it is about to issue; it is a function of
// the particle state only, so the schedule is the same in every run.
bool sort_now(sp_ctx* c) {
  if (!adaptive_sort(c) || c->p.sort_interval <= 0) return sort_due(c);
  const int64_t s = c->steps;
  if (c->last_sort < 0) return true;
  const int64_t k = s - c->last_sort;
  if (k < c->sort_min) return false;
  if (k >= c->sort_max) return true;
  if (c->slow_f0 < 0) c->slow_f0 = slow_fraction(c, c->last_sort);
  if (c->slow_f0 < 0 || s - 2 <= c->last_sort) return false;
  const double f = slow_fraction(c, s - 2);
  return f >= 0 && f - c->slow_f0 > c->sort_rise;
}

void note_sort(sp_ctx* c) {
  c->nsorts += 1;
  c->last_sort = c->steps;
  c->slow_f0 = -1;
}

// One iteration of the loop body (src/lib.rs:155-199), with the push left
// pending (see the file comment).
template <class T> void one_step(sp_ctx* c) {
  const Geo<T> g = geo<T>(c);
  const GridSet gs = gset(c);
  {
    PhaseScope ps(c, 0);
    if (sort_now(c)) {
      sort_species<T>(c);
      note_sort(c);
    }
    zero_currents<T>(c);
    {
      // each species deposits its density into its own scratch grid; one
      // fold then adds them to rho and dens in species order (the same
      // arithmetic as a fold after each pass)
      PhaseScope pp(c, 3);
      for (auto& s : c->species) {
        GridSet gss = gs;
        gss.dsp = s.dsp;
        if (c->push_pending) launch_push_move_deposit<T>(sdev<T>(c, s), g, gss, c->flags_d, c->st, false);
        else launch_move_deposit<T>(sdev<T>(c, s), g, gss, c->flags_d, c->st, false);
      }
      void* d[MAX_SPECIES];
      double q[MAX_SPECIES];
      int ns = 0;
      for (auto& s : c->species) {
        d[ns] = s.dsp;
        q[ns++] = s.charge;
      }
      fold_species_densities<T>(gs, g, d, q, ns, c->st);
    }
    if (adaptive_sort(c)) slow_snapshot(c, c->steps);
    c->push_pending = false;
    fold<T>(c);
  }
  {
    PhaseScope ps(c, 1);
    fields_update<T>(c);
  }
  c->push_pending = true;
  c->steps += 1;
}

const char* flag_text(int f) {
  if (f & F_LEFT) return "assertion failed: min ix > delta (src/prtls.rs:163-170): a particle reached the left edge";
  if (f & F_WALL) return "assertion failed: max ix <= wall_loc + 1 (src/prtls.rs:155-162)";
  if (f & F_CELL) return "assertion failed: ix > 0 && iy > 0 && ijp1 + ix + 1 < len (src/prtls.rs:100-101, src/lib.rs:324-353)";
  return "non-finite particle momentum after the Boris push";
}

// Wait for the stream; report and clear deferred invariant flags.
int sync_check(sp_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->st));
  harvest_timers(c);
  HIPCHK(c, hipMemcpy(c->flags_h, c->flags_d, sizeof(int), hipMemcpyDeviceToHost));
  if (*c->flags_h) {
    const int f = *c->flags_h;
    *c->flags_h = 0;
    HIPCHK(c, hipMemcpy(c->flags_d, c->flags_h, sizeof(int), hipMemcpyHostToDevice));
    return fail(c, SP_EINVARIANT, flag_text(f));
  }
  return SP_OK;
}

int flush(sp_ctx* c) {
  if (c->f64) flush_push<double>(c);
  else flush_push<float>(c);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int valid_species(sp_ctx* c, int id) {
  if (!c) return SP_EINVAL;
  if (id < 0 || id >= (int)c->species.size()) return fail(c, SP_EINVAL, "unknown species id");
  return SP_OK;
}

int alloc_species(sp_ctx* c, SpeciesBuf& s, int64_t n) {
  s.n = n;
  const size_t nb = (size_t)(n > 0 ? n : 1);
  const size_t gb = c->ngrid_elems() * c->esz;
  HIPCHK(c, hipMalloc(&s.dsp, gb));
  HIPCHK(c, hipMemsetAsync(s.dsp, 0, gb, c->st));
  for (SoA* a : {&s.cur, &s.alt}) {
    HIPCHK(c, hipMalloc(&a->cell, nb * 4));
    HIPCHK(c, hipMalloc(&a->tag, nb * 4));
    void** f[5] = {&a->dx, &a->dy, &a->px, &a->py, &a->pz};
    for (auto* p : f) HIPCHK(c, hipMalloc(p, nb * c->esz));
  }
  return SP_OK;
}

void free_species(SpeciesBuf& s) {
  hipFree(s.dsp);
  for (SoA* a : {&s.cur, &s.alt}) {
    hipFree(a->cell);
    hipFree(a->tag);
    void* f[5] = {a->dx, a->dy, a->px, a->py, a->pz};
    for (auto* p : f) hipFree(p);
  }
}

}  // namespace

namespace {

// ---- tuning and the autotuner ----
int check_tuning(sp_ctx* c, const sp_tuning& t) {
  if (c->f64 ? t.chunk != 1024 : (t.chunk != 512 && t.chunk != 1024 && t.chunk != 2048))
    return fail(c, SP_EINVAL, c->f64 ? "chunk must be 1024 in FP64" : "chunk must be 512, 1024 or 2048");
  if (t.tile_w_log < TILE_LOG_MIN || t.tile_w_log > TILE_LOG_MAX || t.tile_h_log < TILE_LOG_MIN ||
      t.tile_h_log > TILE_LOG_MAX)
    return fail(c, SP_EINVAL, "tile logs must be in 1..5");
  if (t.fft_rows < 1 || t.fft_rows > 64 || t.fft_cols < 1 || t.fft_cols > 64)
    return fail(c, SP_EINVAL, "fft_rows and fft_cols must be in 1..64");
  if ((t.fft_radix_x != 8 && t.fft_radix_x != 4) || (t.fft_radix_y != 8 && t.fft_radix_y != 4))
    return fail(c, SP_EINVAL, "FFT radix caps must be 8 or 4");
  return SP_OK;
}

void apply_tuning(sp_ctx* c, const sp_tuning& t) {
  c->tune = t;
  // the twiddle table depends on n only; its pass-order copy on the plan
  plan_dim(c->fx, c->p.size_x, t.fft_radix_x);
  plan_dim(c->fy, c->p.size_y, t.fft_radix_y);
  if (c->f64) {
    fft_pass_twiddles<double>(c->fx, c->st);
    fft_pass_twiddles<double>(c->fy, c->st);
  } else {
    fft_pass_twiddles<float>(c->fx, c->st);
    fft_pass_twiddles<float>(c->fy, c->st);
  }
  c->fx.per_block = t.fft_rows;
  c->fy.per_block = t.fft_cols;
  c->tune_gen++;
  drop_graph(c);
}

// A scratch context with the parameters, particles and fields of c.
int clone_ctx(sp_ctx* c, sp_ctx** out) {
  sp_ctx* t = nullptr;
  if (sp_create(&c->p, &t) != SP_OK) return fail(c, SP_EHIP, std::string("scratch context: ") + sp_last_error(nullptr));
  auto bail = [&](const std::string& m) {
    sp_destroy(t);
    return fail(c, SP_EHIP, "scratch context: " + m);
  };
  for (const auto& s : c->species) {
    SpeciesBuf d;
    d.charge = s.charge;
    d.mass = s.mass;
    if (alloc_species(t, d, s.n) != SP_OK) {
      free_species(d);
      return bail(t->err);
    }
    t->species.push_back(d);
    if (s.n == 0) continue;
    const size_t n = (size_t)s.n;
    hipMemcpyAsync(d.cur.cell, s.cur.cell, n * 4, hipMemcpyDeviceToDevice, c->st);
    hipMemcpyAsync(d.cur.tag, s.cur.tag, n * 4, hipMemcpyDeviceToDevice, c->st);
    const void* src[5] = {s.cur.dx, s.cur.dy, s.cur.px, s.cur.py, s.cur.pz};
    void* dst[5] = {d.cur.dx, d.cur.dy, d.cur.px, d.cur.py, d.cur.pz};
    for (int k = 0; k < 5; ++k) hipMemcpyAsync(dst[k], src[k], n * c->esz, hipMemcpyDeviceToDevice, c->st);
  }
  for (int i = 0; i < 11; ++i)
    hipMemcpyAsync(t->grid[i], c->grid[i], c->ngrid_elems() * c->esz, hipMemcpyDeviceToDevice, c->st);
  for (int i = 0; i < 14; ++i)
    hipMemcpyAsync(t->spec[i], c->spec[i], c->nspec_elems() * 2 * c->esz, hipMemcpyDeviceToDevice, c->st);
  t->half_spec = c->half_spec;
  const hipError_t e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) return bail(hipGetErrorString(e));
  *out = t;
  return SP_OK;
}

// Mean time of `trials` runs of fn on t's stream, after one untimed run.
template <class F> float time_ms(sp_ctx* t, int trials, F&& fn) {
  fn();
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return 1e30f;
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return 1e30f;
  }
  hipEventRecord(a, t->st);
  for (int i = 0; i < trials; ++i) fn();
  hipEventRecord(b, t->st);
  float ms = 1e30f;
  if (hipEventSynchronize(b) == hipSuccess && hipGetLastError() == hipSuccess) hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms / (float)trials;
}

template <class T> void particle_trial(sp_ctx* t) {
  zero_currents<T>(t);
  const Geo<T> g = geo<T>(t);
  const GridSet gs = gset(t);
  for (auto& s : t->species) launch_push_move_deposit<T>(sdev<T>(t, s), g, gs, t->flags_d, t->st);
}

template <class T> sp_tuning autotune_on(sp_ctx* t, int trials) {
  sp_tuning best = t->tune;
  // chunk x sort tile: the particle pass after a sort with that tile
  {
    std::vector<int> chunks = sizeof(T) == 8 ? std::vector<int>{1024} : std::vector<int>{512, 1024, 2048};
    std::vector<std::pair<int, int>> tiles = {{2, 2}, {3, 2}, {3, 3}, {4, 3}, {4, 4}, {5, 3}};
    if (t->p.sort_interval <= 0) tiles = {{t->tune.tile_w_log, t->tune.tile_h_log}};
    float bt = 1e30f;
    sp_tuning cand = best;
    for (int k : chunks)
      for (const auto& tl : tiles) {
        cand.chunk = k;
        cand.tile_w_log = tl.first;
        cand.tile_h_log = tl.second;
        apply_tuning(t, cand);
        if (t->p.sort_interval > 0) sort_species<T>(t);
        const float ms = time_ms(t, trials, [&] { particle_trial<T>(t); });
        if (ms < bt) {
          bt = ms;
          best.chunk = k;
          best.tile_w_log = tl.first;
          best.tile_h_log = tl.second;
        }
      }
  }
  const int maxlen = fft_max_len<T>();
  // radix x rows per workgroup, then radix x columns per workgroup: the field solve
  {
    float bt = 1e30f;
    sp_tuning cand = best;
    for (int r : {8, 4})
      for (int rows = 1; rows <= 64 && rows * t->p.size_x <= maxlen && rows <= t->p.size_y; rows *= 2) {
        cand.fft_radix_x = r;
        cand.fft_rows = rows;
        apply_tuning(t, cand);
        const float ms = time_ms(t, trials, [&] { fields_update<T>(t); });
        if (ms < bt) {
          bt = ms;
          best.fft_radix_x = r;
          best.fft_rows = rows;
        }
      }
  }
  {
    float bt = 1e30f;
    sp_tuning cand = best;
    for (int r : {8, 4})
      for (int cols = 1; cols <= 64 && cols * t->p.size_y <= maxlen && cols <= t->p.size_x; cols *= 2) {
        cand.fft_radix_y = r;
        cand.fft_cols = cols;
        apply_tuning(t, cand);
        const float ms = time_ms(t, trials, [&] { fields_update<T>(t); });
        if (ms < bt) {
          bt = ms;
          best.fft_radix_y = r;
          best.fft_cols = cols;
        }
      }
  }
  return best;
}

std::mutex g_tune_mu;
std::map<std::vector<int64_t>, sp_tuning> g_tune_cache;

std::vector<int64_t> tune_key(const sp_ctx* c) {
  return {c->p.size_x, c->p.size_y, c->p.delta, c->p.dens, (int64_t)c->species.size(), c->f64 ? 1 : 0,
          c->p.sort_interval, c->p.n_pass};
}

}  // namespace

// ---- ensembles ----
namespace {
// The host-side state of a context that one step changes: grid/work pointers
// (the filter swaps them), species buffers (a sort swaps them), push state.
struct MemberState {
  void* grid[11];
  void* work[4];
  std::vector<std::pair<SoA, SoA>> sp;
  bool push_pending;
};
MemberState save_state(const sp_ctx* c) {
  MemberState m;
  std::memcpy(m.grid, c->grid, sizeof(m.grid));
  std::memcpy(m.work, c->work, sizeof(m.work));
  for (const auto& s : c->species) m.sp.emplace_back(s.cur, s.alt);
  m.push_pending = c->push_pending;
  return m;
}
void load_state(sp_ctx* c, const MemberState& m) {
  std::memcpy(c->grid, m.grid, sizeof(m.grid));
  std::memcpy(c->work, m.work, sizeof(m.work));
  for (size_t i = 0; i < c->species.size(); ++i) {
    c->species[i].cur = m.sp[i].first;
    c->species[i].alt = m.sp[i].second;
  }
  c->push_pending = m.push_pending;
}
}  // namespace

struct sp_ensemble {
  std::vector<sp_ctx*> m;
  std::vector<int> branch_of;
  int nb = 1, unit = 2;
  bool graph_on = true;
  hipStream_t st = nullptr;
  std::vector<hipStream_t> bst;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  struct Entry {
    std::vector<int64_t> sig;
    std::vector<MemberState> post;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Entry> kept;
  int captures = 0;
  std::string err;
};

namespace {

int efail(sp_ensemble* e, int code, const std::string& msg) {
  e->err = msg;
  return code;
}

void drop_kept(sp_ensemble* e) {
  for (auto& k : e->kept)
    if (k.exec) hipGraphExecDestroy(k.exec);
  e->kept.clear();
}

// Everything the launches of the next `unit` steps depend on.
std::vector<int64_t> ens_signature(const sp_ensemble* e) {
  std::vector<int64_t> v;
  for (const sp_ctx* c : e->m) {
    for (const void* p : graph_signature(c)) v.push_back((int64_t)(intptr_t)p);
    v.push_back(c->push_pending);
    v.push_back(c->tune_gen);
    v.push_back(c->dn_on);
    for (int j = 0; j < e->unit; ++j)
      v.push_back(c->p.sort_interval > 0 && (c->steps + j) % c->p.sort_interval == 0);
  }
  return v;
}

// One step of every member: fork into the branch streams, each member's
// launches on its branch, join back into the ensemble stream.
void ens_step_forked(sp_ensemble* e) {
  hipEventRecord(e->fork, e->st);
  for (int b = 0; b < e->nb; ++b) hipStreamWaitEvent(e->bst[b], e->fork, 0);
  for (size_t i = 0; i < e->m.size(); ++i) {
    sp_ctx* c = e->m[i];
    c->st = e->bst[e->branch_of[i]];
    if (c->f64) one_step<double>(c);
    else one_step<float>(c);
    c->st = e->st;
  }
  for (int b = 0; b < e->nb; ++b) {
    hipEventRecord(e->join[b], e->bst[b]);
    hipStreamWaitEvent(e->st, e->join[b], 0);
  }
}

// `unit` steps of every member as one graph launch; captured when no kept
// graph matches the members' state.
int ens_graph_unit(sp_ensemble* e) {
  const std::vector<int64_t> sig = ens_signature(e);
  for (const auto& k : e->kept) {
    if (k.sig != sig) continue;
    const hipError_t r = hipGraphLaunch(k.exec, e->st);
    if (r != hipSuccess) return efail(e, SP_EHIP, std::string("hipGraphLaunch: ") + hipGetErrorString(r));
    for (size_t i = 0; i < e->m.size(); ++i) {
      load_state(e->m[i], k.post[i]);
      e->m[i]->steps += e->unit;
    }
    return SP_OK;
  }
  if (e->kept.size() >= 16) drop_kept(e);
  std::vector<MemberState> pre;
  std::vector<int64_t> steps0;
  for (sp_ctx* c : e->m) {
    pre.push_back(save_state(c));
    steps0.push_back(c->steps);
  }
  auto undo = [&](const std::string& what, hipError_t r) {
    for (size_t i = 0; i < e->m.size(); ++i) {
      load_state(e->m[i], pre[i]);
      e->m[i]->steps = steps0[i];
    }
    return efail(e, SP_EHIP, what + ": " + hipGetErrorString(r));
  };
  hipError_t r = hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal);
  if (r != hipSuccess) return undo("hipStreamBeginCapture", r);
  for (int u = 0; u < e->unit; ++u) ens_step_forked(e);
  hipGraph_t graph = nullptr;
  r = hipStreamEndCapture(e->st, &graph);
  if (r != hipSuccess) return undo("hipStreamEndCapture", r);
  sp_ensemble::Entry k;
  k.sig = sig;
  r = hipGraphInstantiate(&k.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (r != hipSuccess) return undo("hipGraphInstantiate", r);
  for (sp_ctx* c : e->m) k.post.push_back(save_state(c));
  e->captures++;
  r = hipGraphLaunch(k.exec, e->st);
  e->kept.push_back(std::move(k));
  if (r != hipSuccess) return efail(e, SP_EHIP, std::string("hipGraphLaunch: ") + hipGetErrorString(r));
  return SP_OK;
}

}  // namespace

extern "C" {

int sp_create(const sp_params* params, sp_ctx** out) {
  if (!params || !out) return fail(nullptr, SP_EINVAL, "null argument");
  const sp_params& p = *params;
  // Config::new validation (src/lib.rs:75-113)
  if (p.size_x <= 0 || p.size_y <= 0) return fail(nullptr, SP_EINVAL, "grid sizes must be positive");
  if (p.size_x % 2 != 0 || p.size_y % 2 != 0) {
    const char* w = (p.size_x % 2 != 0 && p.size_y % 2 != 0) ? "x & y" : (p.size_x % 2 != 0 ? "x" : "y");
    return fail(nullptr, SP_EINVAL, std::string("Number of cells in ") + w + " direction must be even");
  }
  if (p.delta < 0 || p.delta * 2 >= p.size_x)
    return fail(nullptr, SP_EINVAL, "Delta must be less than 1/2 the size of size_x");
  if (!(p.gamma_inj > 1.0)) return fail(nullptr, SP_EINVAL, "Lorentz factor must be greater than 1");
  if (p.dens < 1 || p.n_pass < 0) return fail(nullptr, SP_EINVAL, "dens must be >= 1 and n_pass >= 0");
  if (!(p.dt > 0) || !(p.c > 0)) return fail(nullptr, SP_EINVAL, "dt and c must be positive");
  if (p.precision != SP_F32 && p.precision != SP_F64) return fail(nullptr, SP_EINVAL, "unknown precision");
  if (p.left_edge != SP_LEFT_ASSERT && p.left_edge != SP_LEFT_REFLECT)
    return fail(nullptr, SP_EINVAL, "unknown left_edge rule");
  if (p.study & ~(SP_STUDY_ABS_POS32 | SP_STUDY_TWIDDLE32 | SP_STUDY_SUM32))
    return fail(nullptr, SP_EINVAL, "unknown study flag");
  if (p.study && p.precision != SP_F32) return fail(nullptr, SP_EINVAL, "study variants are of the FP32 path");
  // limits of the GPU layout: packed cells, LDS-resident 1D transforms
  if (p.size_x + 2 > 65535 || p.size_y + 2 > 32767)
    return fail(nullptr, SP_EINVAL, "grid too large for the packed (iy << 16 | ix) cell index");
  const int maxn = p.precision == SP_F64 ? fft_max_len<double>() : fft_max_len<float>();
  FftDim fx{}, fy{};
  if (!plan_dim(fx, p.size_x) || !plan_dim(fy, p.size_y) || p.size_x > maxn || p.size_y > maxn)
    return fail(nullptr, SP_EINVAL,
                "size_x and size_y must be 2^a 3^b and at most " + std::to_string(maxn) + " for the GPU FFT");

  sp_ctx* c = new sp_ctx();
  c->p = p;
  c->f64 = p.precision == SP_F64;
  c->esz = c->f64 ? 8 : 4;
  c->fx = fx;
  c->fy = fy;
  auto bail = [&](int rc) {
    g_create_error = c->err;
    sp_destroy(c);
    return rc;
  };
#define CK(call)                                                                     \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      c->err = std::string(#call) + ": " + hipGetErrorString(e_);                    \
      return bail(SP_EHIP);                                                          \
    }                                                                                \
  } while (0)
  CK(hipGetDevice(&c->device));
  CK(hipStreamCreateWithFlags(&c->own_st, hipStreamNonBlocking));
  c->st = c->own_st;
  c->tune = default_tuning(c->f64, p.dens);
  c->fx.per_block = c->tune.fft_rows;
  c->fy.per_block = c->tune.fft_cols;
  const size_t gb = c->ngrid_elems() * c->esz, sb = c->nspec_elems() * 2 * c->esz;
  for (auto& g : c->grid) CK(hipMalloc(&g, gb));
  for (auto& g : c->work) CK(hipMalloc(&g, gb));
  CK(hipMalloc(&c->dsp, gb));
  CK(hipMemsetAsync(c->dsp, 0, gb, c->st));
  for (auto& s : c->spec) CK(hipMalloc(&s, sb));
  for (auto& g : c->grid) CK(hipMemsetAsync(g, 0, gb, c->st));
  for (auto& g : c->work) CK(hipMemsetAsync(g, 0, gb, c->st));
  for (auto& s : c->spec) CK(hipMemsetAsync(s, 0, sb, c->st));
  CK(hipMalloc(&c->fx.tw, (size_t)fx.n * 2 * c->esz));
  CK(hipMalloc(&c->fy.tw, (size_t)fy.n * 2 * c->esz));
  CK(hipMalloc(&c->fx.twp, (size_t)fx.n * 2 * c->esz));
  CK(hipMalloc(&c->fy.twp, (size_t)fy.n * 2 * c->esz));
  if (c->f64) {
    fft_make_twiddles<double>(c->fx, c->st);
    fft_make_twiddles<double>(c->fy, c->st);
  } else {
    const bool tw32 = (p.study & SP_STUDY_TWIDDLE32) != 0;
    fft_make_twiddles<float>(c->fx, c->st, tw32);
    fft_make_twiddles<float>(c->fy, c->st, tw32);
  }
  {
    // room for the smallest sort tile a tuning may pick
    const int tmin = 1 << TILE_LOG_MIN;
    const size_t ntiles = (size_t)((p.size_x + tmin - 1) / tmin) * ((p.size_y + tmin - 1) / tmin);
    CK(hipMalloc(&c->tile_count, tile_count_size(ntiles) * sizeof(int)));
    CK(hipMalloc(&c->tile_cursor, ntiles * sizeof(int)));
  }
  CK(hipMalloc(&c->flags_d, 4 * sizeof(int)));
  {
    const size_t gn = gauss_doubles(p.size_x, p.size_y);
    CK(hipMalloc(&c->gauss_d, gn * sizeof(double)));
    CK(hipMemsetAsync(c->gauss_d, 0, gn * sizeof(double), c->st));
    CK(hipMalloc(&c->diag_d, DIAG_DOUBLES * sizeof(double)));
  }
  CK(hipHostMalloc((void**)&c->flags_h, sizeof(int), hipHostMallocDefault));
  CK(hipMemsetAsync(c->flags_d, 0, 4 * sizeof(int), c->st));
  CK(hipStreamSynchronize(c->st));
  *c->flags_h = 0;
  std::fill(c->slow_step, c->slow_step + SLOW_RING, (int64_t)-2);
#undef CK
  *out = c;
  return SP_OK;
}

int sp_destroy(sp_ctx* c) {
  if (!c) return SP_OK;
  if (c->ens) return fail(c, SP_EINVAL, "the context belongs to an ensemble: destroy the ensemble first");
  if (c->st) hipStreamSynchronize(c->st);
  drop_graph(c);
  hipFree(c->dn_buf);
  hipFree(c->dn_sums);
  hipFree(c->dn_stats);
  for (auto& s : c->species) free_species(s);
  for (auto g : c->grid) hipFree(g);
  for (auto g : c->work) hipFree(g);
  hipFree(c->dsp);
  for (auto s : c->spec) hipFree(s);
  hipFree(c->fx.tw);
  hipFree(c->fy.tw);
  hipFree(c->fx.twp);
  hipFree(c->fy.twp);
  hipFree(c->flags_d);
  hipFree(c->tile_count);
  hipFree(c->tile_cursor);
  hipFree(c->gauss_d);
  hipFree(c->diag_d);
  if (c->flags_h) hipHostFree(c->flags_h);
  if (c->slow_h) hipHostFree(c->slow_h);
  for (auto e : c->slow_ev)
    if (e) hipEventDestroy(e);
  for (auto& tp : c->open_pairs) { hipEventDestroy(tp.a); hipEventDestroy(tp.b); }
  for (auto e : c->free_events) hipEventDestroy(e);
  if (c->own_st) hipStreamDestroy(c->own_st);
  delete c;
  return SP_OK;
}

const char* sp_last_error(const sp_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int sp_species_add(sp_ctx* c, double charge, double mass, int64_t n, const int32_t* ix, const int32_t* iy,
                   const void* dx, const void* dy, const void* px, const void* py, const void* pz, int* id) {
  if (!c) return SP_EINVAL;
  if (n < 0 || (n > 0 && (!ix || !iy || !dx || !dy || !px || !py || !pz)))
    return fail(c, SP_EINVAL, "bad species arrays");
  if (!(mass != 0.0)) return fail(c, SP_EINVAL, "mass must be non-zero");
  if (n > INT32_MAX) return fail(c, SP_EINVAL, "species too large for int32 tags");
  if ((int)c->species.size() >= MAX_SPECIES) return fail(c, SP_EINVAL, "too many species");
  int rc = flush(c);
  if (rc) return rc;
  std::vector<int32_t> cell(n), tag(n);
  for (int64_t i = 0; i < n; ++i) {
    if (ix[i] < 0 || ix[i] > c->p.size_x + 1 || iy[i] < 0 || iy[i] > c->p.size_y + 1)
      return fail(c, SP_EINVAL, "particle cell outside the padded grid");
    cell[i] = make_cell(ix[i], iy[i]);
    tag[i] = (int32_t)i;
  }
  SpeciesBuf s;
  s.charge = charge;
  s.mass = mass;
  rc = alloc_species(c, s, n);
  if (rc) { free_species(s); return rc; }
  if (n > 0) {
    HIPCHK(c, hipMemcpy(s.cur.cell, cell.data(), n * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(s.cur.tag, tag.data(), n * 4, hipMemcpyHostToDevice));
    const void* src[5] = {dx, dy, px, py, pz};
    void* dst[5] = {s.cur.dx, s.cur.dy, s.cur.px, s.cur.py, s.cur.pz};
    for (int k = 0; k < 5; ++k) HIPCHK(c, hipMemcpy(dst[k], src[k], n * c->esz, hipMemcpyHostToDevice));
  }
  c->species.push_back(s);
  if (id) *id = (int)c->species.size() - 1;
  return SP_OK;
}

int sp_species_init(sp_ctx* c, double charge, double mass, double vth, uint64_t seed, int* id) {
  if (!c) return SP_EINVAL;
  if (!(mass != 0.0) || !(vth >= 0.0)) return fail(c, SP_EINVAL, "bad mass or vth");
  int rc = flush(c);
  if (rc) return rc;
  const int64_t n = (int64_t)c->p.dens * (c->p.size_x - 2 * c->p.delta) * c->p.size_y;  // src/lib.rs:258-260
  if (n > INT32_MAX) return fail(c, SP_EINVAL, "species too large for int32 tags");
  if ((int)c->species.size() >= MAX_SPECIES) return fail(c, SP_EINVAL, "too many species");
  SpeciesBuf s;
  s.charge = charge;
  s.mass = mass;
  rc = alloc_species(c, s, n);
  if (rc) { free_species(s); return rc; }
  const int sid = (int)c->species.size();
  if (c->f64) launch_init_species<double>(sdev<double>(c, s), geo<double>(c), vth, c->p.gamma_inj, seed, sid, c->st);
  else launch_init_species<float>(sdev<float>(c, s), geo<float>(c), vth, c->p.gamma_inj, seed, sid, c->st);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->species.push_back(s);
  if (id) *id = sid;
  return SP_OK;
}

int sp_species_count(const sp_ctx* c, int id, int64_t* n) {
  if (!c || !n || id < 0 || id >= (int)c->species.size()) return SP_EINVAL;
  *n = c->species[id].n;
  return SP_OK;
}

int sp_get_species(sp_ctx* c, int id, int64_t n, int32_t* ix, int32_t* iy, void* dx, void* dy, void* px, void* py,
                   void* pz, void* psa) {
  int rc = valid_species(c, id);
  if (rc) return rc;
  const SpeciesBuf& s = c->species[id];
  if (n != s.n) return fail(c, SP_EINVAL, "count does not match the species size");
  rc = flush(c);
  if (rc) return rc;
  rc = sync_check(c);
  if (rc) return rc;
  if (n == 0) return SP_OK;
  std::vector<int32_t> cell(n), tag(n);
  HIPCHK(c, hipMemcpy(cell.data(), s.cur.cell, n * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(tag.data(), s.cur.tag, n * 4, hipMemcpyDeviceToHost));
  // values in tag order; the momenta are kept whenever psa is asked for, so
  // gamma never reads a caller array that may be NULL
  const size_t e = c->esz;
  std::vector<unsigned char> buf(n * e), mom[3];
  void* outs[5] = {dx, dy, px, py, pz};
  void* srcs[5] = {s.cur.dx, s.cur.dy, s.cur.px, s.cur.py, s.cur.pz};
  for (int k = 0; k < 5; ++k) {
    const bool keep = psa && k >= 2;
    if (!outs[k] && !keep) continue;
    HIPCHK(c, hipMemcpy(buf.data(), srcs[k], n * e, hipMemcpyDeviceToHost));
    unsigned char* o = (unsigned char*)outs[k];
    if (keep) {
      mom[k - 2].resize(n * e);
      o = mom[k - 2].data();
    }
    for (int64_t i = 0; i < n; ++i) std::memcpy(o + (size_t)tag[i] * e, buf.data() + (size_t)i * e, e);
    if (keep && outs[k]) std::memcpy(outs[k], o, n * e);
  }
  for (int64_t i = 0; i < n; ++i) {
    if (ix) ix[tag[i]] = cell_ix(cell[i]);
    if (iy) iy[tag[i]] = cell_iy(cell[i]);
  }
  if (psa) {  // src/prtls.rs:440 in Float
    for (int64_t i = 0; i < n; ++i) {
      if (c->f64) {
        const double* X = (const double*)mom[0].data(); const double* Y = (const double*)mom[1].data();
        const double* Z = (const double*)mom[2].data();
        const double cc = c->p.c, csq = 1. / (cc * cc);
        ((double*)psa)[i] = std::sqrt(1.0 + (X[i] * X[i] + Y[i] * Y[i] + Z[i] * Z[i]) * csq);
      } else {
        const float* X = (const float*)mom[0].data(); const float* Y = (const float*)mom[1].data();
        const float* Z = (const float*)mom[2].data();
        const float cc = (float)c->p.c, csq = 1.f / (cc * cc);
        ((float*)psa)[i] = std::sqrt(1.0f + (X[i] * X[i] + Y[i] * Y[i] + Z[i] * Z[i]) * csq);
      }
    }
  }
  return SP_OK;
}

int sp_get_species_raw(sp_ctx* c, int id, int64_t n, int32_t* cell, int32_t* tag, void* dx, void* dy, void* px,
                       void* py, void* pz) {
  int rc = valid_species(c, id);
  if (rc) return rc;
  const SpeciesBuf& s = c->species[id];
  if (n != s.n) return fail(c, SP_EINVAL, "count does not match the species size");
  rc = flush(c);
  if (rc) return rc;
  rc = sync_check(c);
  if (rc) return rc;
  if (n == 0) return SP_OK;
  void* outs[7] = {cell, tag, dx, dy, px, py, pz};
  const void* srcs[7] = {s.cur.cell, s.cur.tag, s.cur.dx, s.cur.dy, s.cur.px, s.cur.py, s.cur.pz};
  for (int k = 0; k < 7; ++k)
    if (outs[k]) HIPCHK(c, hipMemcpy(outs[k], srcs[k], n * (k < 2 ? 4 : c->esz), hipMemcpyDeviceToHost));
  return SP_OK;
}

static int grid_io(sp_ctx* c, int grid, void* host, size_t count, bool to_dev, bool spectrum) {
  if (!c || !host) return SP_EINVAL;
  const int nmax = spectrum ? 14 : 11;
  if (grid < 0 || grid >= nmax) return fail(c, SP_EINVAL, "unknown grid id");
  const size_t need = spectrum ? 2 * c->nspec_elems() : c->ngrid_elems();
  if (count != need) return fail(c, SP_EINVAL, "count must be " + std::to_string(need));
  int rc = flush(c);
  if (rc) return rc;
  rc = sync_check(c);
  if (rc) return rc;
  void* dev = spectrum ? c->spec[grid] : c->grid[grid];
  if (spectrum && to_dev) c->half_spec &= ~(1u << grid);  // the caller's values are whole
  if (spectrum && !to_dev) {
    full_spectrum(c, grid);
    HIPCHK(c, hipStreamSynchronize(c->st));
  }
  if (to_dev) HIPCHK(c, hipMemcpy(dev, host, need * c->esz, hipMemcpyHostToDevice));
  else HIPCHK(c, hipMemcpy(host, dev, need * c->esz, hipMemcpyDeviceToHost));
  return SP_OK;
}

int sp_get_grid(sp_ctx* c, int g, void* out, size_t n) { return grid_io(c, g, out, n, false, false); }
int sp_set_grid(sp_ctx* c, int g, const void* in, size_t n) { return grid_io(c, g, (void*)in, n, true, false); }
int sp_get_spectrum(sp_ctx* c, int g, void* out, size_t n) { return grid_io(c, g, out, n, false, true); }
int sp_set_spectrum(sp_ctx* c, int g, const void* in, size_t n) { return grid_io(c, g, (void*)in, n, true, true); }

#define ENTRY_PRELUDE(c)        \
  if (!(c)) return SP_EINVAL;   \
  {                             \
    int rc_ = flush(c);         \
    if (rc_) return rc_;        \
  }

int sp_zero_currents(sp_ctx* c) {
  ENTRY_PRELUDE(c);
  if (c->f64) zero_currents<double>(c);
  else zero_currents<float>(c);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_move_and_deposit(sp_ctx* c, int id) {
  ENTRY_PRELUDE(c);
  int rc = valid_species(c, id);
  if (rc) return rc;
  PhaseScope ps(c, 0);
  if (c->f64) launch_move_deposit<double>(sdev<double>(c, c->species[id]), geo<double>(c), gset(c), c->flags_d, c->st);
  else launch_move_deposit<float>(sdev<float>(c, c->species[id]), geo<float>(c), gset(c), c->flags_d, c->st);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_deposit_ghosts(sp_ctx* c) {
  ENTRY_PRELUDE(c);
  if (c->f64) fold<double>(c);
  else fold<float>(c);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_fields_update(sp_ctx* c) {
  ENTRY_PRELUDE(c);
  PhaseScope ps(c, 1);
  if (c->f64) fields_update<double>(c);
  else fields_update<float>(c);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_boris_push(sp_ctx* c, int id) {
  ENTRY_PRELUDE(c);
  int rc = valid_species(c, id);
  if (rc) return rc;
  PhaseScope ps(c, 2);
  if (c->f64) launch_push<double>(sdev<double>(c, c->species[id]), geo<double>(c), gset(c), c->flags_d, c->st);
  else launch_push<float>(sdev<float>(c, c->species[id]), geo<float>(c), gset(c), c->flags_d, c->st);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_filter(sp_ctx* c, int grid, int passes) {
  ENTRY_PRELUDE(c);
  if (grid < 0 || grid >= 11 || passes < 0) return fail(c, SP_EINVAL, "bad grid id or pass count");
  if (c->f64) launch_binomial<double>(&c->grid[grid], c->work, 1, c->p.size_x, c->p.size_y, passes, c->st);
  else launch_binomial<float>(&c->grid[grid], c->work, 1, c->p.size_x, c->p.size_y, passes, c->st);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

}  // extern "C"

// ---- sp_filter without allocation ----
// The filter's workspace (the grid each fused launch writes) lives in a
// file-scope device array instead of a context's work grid: one padded grid
// of up to SCRATCH_BYTES, the 4096 x 1024 grid in either precision.  Each
// device holds its own copy (33.7 MB) for the life of the process; its
// address is taken once per device with that device current.  Every call on
// a device shares that device's copy, so calls that overlap on one device
// (other streams or host threads) race.
namespace {
constexpr size_t SCRATCH_BYTES = (size_t)4210688 * 8;  // >= 4098 x 1026 doubles
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch[SCRATCH_MAX_DEVICES] = {};

int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(nullptr, SP_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(nullptr, SP_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, SP_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return SP_OK;
}
}  // namespace

extern "C" {

int sp_filter_noalloc(void* grid, int32_t size_x, int32_t size_y, int precision, int passes, void* stream) {
  if (!grid) return fail(nullptr, SP_EINVAL, "null grid");
  if (size_x <= 0 || size_y <= 0 || size_x % 2 != 0 || size_y % 2 != 0)
    return fail(nullptr, SP_EINVAL, "grid sizes must be positive and even");
  if (precision != SP_F32 && precision != SP_F64) return fail(nullptr, SP_EINVAL, "unknown precision");
  if (passes < 0) return fail(nullptr, SP_EINVAL, "negative pass count");
  const size_t esz = precision == SP_F64 ? 8 : 4;
  const size_t bytes = (size_t)(size_x + 2) * (size_t)(size_y + 2) * esz;
  if (bytes > SCRATCH_BYTES)
    return fail(nullptr, SP_EINVAL,
                "grid of " + std::to_string(bytes) + " bytes exceeds the " + std::to_string(SCRATCH_BYTES) +
                    "-byte workspace of sp_filter_noalloc");
  void* work = nullptr;
  int rc = scratch_address(&work);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  void* g = grid;
  if (precision == SP_F64) launch_binomial<double>(&g, &work, 1, size_x, size_y, passes, st);
  else launch_binomial<float>(&g, &work, 1, size_x, size_y, passes, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(nullptr, SP_EHIP, std::string("filter launch: ") + hipGetErrorString(e));
  if (g != grid) {  // an odd number of launches left the result in the workspace
    e = hipMemcpyAsync(grid, g, bytes, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return fail(nullptr, SP_EHIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  }
  return SP_OK;
}

int sp_fft2d(sp_ctx* c, int spec, int inverse) {
  ENTRY_PRELUDE(c);
  if (spec < 0 || spec >= 14) return fail(c, SP_EINVAL, "unknown spectrum id");
  full_spectrum(c, spec);  // the transform reads (inverse) or rewrites (forward) all of it
  FftBatch b{};
  b.count = 1;
  b.spec[0] = c->spec[spec];
  b.grid[0] = nullptr;
  b.scale[0] = 1.0;
  const int nx = c->p.size_x, ny = c->p.size_y;
  if (c->f64) {
    const double norm = inverse ? 1.0 / (double)((int64_t)nx * ny) : 1.0;
    launch_fft_rows<double>(b, c->fx, ny, inverse, LOAD_COMPLEX, STORE_COMPLEX, 1.0, c->st);
    launch_fft_cols<double>(b, c->fy, nx, inverse, norm, c->st);
  } else {
    const float norm = inverse ? 1.0f / (float)((int64_t)nx * ny) : 1.0f;
    launch_fft_rows<float>(b, c->fx, ny, inverse, LOAD_COMPLEX, STORE_COMPLEX, 1.0, c->st);
    launch_fft_cols<float>(b, c->fy, nx, inverse, (double)norm, c->st);
  }
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_spectral_update(sp_ctx* c) {
  ENTRY_PRELUDE(c);
  // the per-mode update reads every mode of J, rho and E
  full_spectra(c, HALF_SPECTRA);
  if (c->f64) launch_spectral_update<double>(c->spec, geo<double>(c), c->p.dens, c->gauss_d, c->st, false);
  else launch_spectral_update<float>(c->spec, geo<float>(c), c->p.dens, c->gauss_d, c->st, false);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_prestep(sp_ctx* c) {
  ENTRY_PRELUDE(c);
  PhaseScope ps(c, 0);
  if (c->p.sort_interval > 0) {
    if (c->f64) sort_species<double>(c);
    else sort_species<float>(c);
    note_sort(c);
  }
  for (auto& s : c->species) {
    if (c->f64) launch_move_deposit<double>(sdev<double>(c, s), geo<double>(c), gset(c), c->flags_d, c->st);
    else launch_move_deposit<float>(sdev<float>(c, s), geo<float>(c), gset(c), c->flags_d, c->st);
  }
  if (adaptive_sort(c)) slow_snapshot(c, c->steps - 1);
  if (c->f64) fold<double>(c);
  else fold<float>(c);
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_step(sp_ctx* c, int n) {
  if (!c) return SP_EINVAL;
  if (n < 0) return fail(c, SP_EINVAL, "negative step count");
  for (int i = 0; i < n;) {
    const int P = graph_period(c);
    if (c->graph_on && !c->timing && c->push_pending && c->steps % P == 0 && n - i >= P) {  // fixed sort phase
      const int rc = c->f64 ? graph_period_run<double>(c) : graph_period_run<float>(c);
      if (rc) return rc;
      i += P;
    } else {
      if (c->f64) one_step<double>(c);
      else one_step<float>(c);
      ++i;
    }
  }
  if (n > 0) c->half_spec |= HALF_SPECTRA;  // graph replays skip the host side of fields_update
  HIPCHK(c, hipGetLastError());
  return SP_OK;
}

int sp_set_denoiser(sp_ctx* c, const float* params, int n, float max_rel, float max_gauss) {
  ENTRY_PRELUDE(c);
  c->tune_gen++;  // captured steps hold the denoiser's settings as kernel arguments
  drop_graph(c);
  if (!params) {
    c->dn_on = false;
    return SP_OK;
  }
  if (c->f64) return fail(c, SP_EINVAL, "the denoiser runs in FP32 contexts only");
  if (n != DN_NPARAMS) return fail(c, SP_EINVAL, "expected SP_DN_NPARAMS parameters");
  if (!(max_rel > 0) || !(max_gauss >= 0)) return fail(c, SP_EINVAL, "max_rel must be positive, max_gauss >= 0");
  if (!c->dn_buf) {
    const size_t planes = DN_IN * c->nspec_elems();
    HIPCHK(c, hipMalloc(&c->dn_buf, (DN_NPARAMS + DN_PACK_SHORTS / 2 + planes) * sizeof(float)));
    HIPCHK(c, hipMalloc(&c->dn_sums, DN_SUMS * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->dn_stats, DN_STATS * sizeof(int)));
  }
  HIPCHK(c, hipMemcpyAsync(c->dn_buf, params, DN_NPARAMS * sizeof(float), hipMemcpyHostToDevice, c->st));
  fill<int>(c->dn_stats, 0, DN_STATS, c->st);
  HIPCHK(c, hipStreamSynchronize(c->st));
  c->dn_tol = max_rel;
  c->dn_gtol = max_gauss;
  c->dn_on = true;
  return SP_OK;
}

int sp_set_denoiser_input(sp_ctx* c, int raw, float max_hp) {
  if (!c) return SP_EINVAL;
  if (!(max_hp >= 0)) return fail(c, SP_EINVAL, "max_hp must be >= 0");
  if (raw && c->p.n_pass > 8) return fail(c, SP_EINVAL, "the unfiltered input needs n_pass <= 8 (one filter launch)");
  c->dn_raw = raw != 0;
  c->dn_hp = max_hp;
  c->tune_gen++;  // captured steps hold these as kernel arguments
  drop_graph(c);
  return SP_OK;
}

int sp_set_denoiser_channels(sp_ctx* c, int mask) {
  if (!c) return SP_EINVAL;
  if (mask < 0 || mask > 0xF) return fail(c, SP_EINVAL, "channel mask must be in 0..15");
  c->dn_mask = mask;
  c->tune_gen++;  // captured steps hold the mask as a kernel argument
  drop_graph(c);
  return SP_OK;
}

int sp_get_denoiser_stats(sp_ctx* c, int* applied, int* fallbacks) {
  if (!c || !applied || !fallbacks) return SP_EINVAL;
  int h[2] = {0, 0};
  if (c->dn_stats) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(h, c->dn_stats, sizeof(h), hipMemcpyDeviceToHost));
  }
  *applied = h[0];
  *fallbacks = h[1];
  return SP_OK;
}

int sp_set_denoiser_backoff(sp_ctx* c, int solves) {
  if (!c) return SP_EINVAL;
  if (solves < 0) return fail(c, SP_EINVAL, "the backoff must be >= 0 solves");
  c->dn_backoff = solves;
  c->tune_gen++;  // captured steps hold it as a kernel argument
  drop_graph(c);
  return SP_OK;
}

int sp_get_denoiser_skipped(sp_ctx* c, int* skipped) {
  if (!c || !skipped) return SP_EINVAL;
  int h[DN_STATS] = {};
  if (c->dn_stats) {
    HIPCHK(c, hipStreamSynchronize(c->st));
    HIPCHK(c, hipMemcpy(h, c->dn_stats, sizeof(h), hipMemcpyDeviceToHost));
  }
  *skipped = h[DN_STAT_SKIPPED];
  return SP_OK;
}

int sp_set_graph(sp_ctx* c, int on) {
  if (!c) return SP_EINVAL;
  if (on && c->ens) return fail(c, SP_EINVAL, "the ensemble captures its members' steps");
  c->graph_on = on != 0;
  if (!c->graph_on) drop_graph(c);
  return SP_OK;
}

int sp_set_sort_policy(sp_ctx* c, int min_interval, int max_interval, double max_slow_rise) {
  if (!c) return SP_EINVAL;
  if (max_interval == 0) {
    c->sort_max = 0;
    return SP_OK;
  }
  if (c->p.sort_interval <= 0) return fail(c, SP_EINVAL, "the adaptive sort schedule needs sort_interval > 0");
  if (min_interval < 1 || max_interval < min_interval || !(max_slow_rise >= 0) || !(max_slow_rise <= 1))
    return fail(c, SP_EINVAL, "need 1 <= min_interval <= max_interval and 0 <= max_slow_rise <= 1");
  if (!c->slow_h) {
    HIPCHK(c, hipHostMalloc((void**)&c->slow_h, SLOW_RING * sizeof(uint64_t), hipHostMallocDefault));
    for (auto& e : c->slow_ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  c->sort_min = min_interval;
  c->sort_max = max_interval;
  c->sort_rise = max_slow_rise;
  std::fill(c->slow_step, c->slow_step + SLOW_RING, (int64_t)-2);
  c->slow_f0 = -1;
  return SP_OK;
}

int sp_sync(sp_ctx* c) {
  if (!c) return SP_EINVAL;
  int rc = flush(c);
  if (rc) return rc;
  return sync_check(c);
}

int sp_set_timing(sp_ctx* c, int on) {
  if (!c) return SP_EINVAL;
  if (on && c->ens) return fail(c, SP_EINVAL, "no phase timers inside an ensemble");
  c->timing = on != 0;
  return SP_OK;
}

int sp_get_timers(sp_ctx* c, double* ms, int reset) {
  if (!c || !ms) return SP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->st));
  harvest_timers(c);
  for (int i = 0; i < 4; ++i) ms[i] = c->t_ms[i];
  if (reset)
    for (double& t : c->t_ms) t = 0;
  return SP_OK;
}

int sp_get_diag(sp_ctx* c, sp_diag* out) {
  if (!c || !out) return SP_EINVAL;
  int rc = flush(c);
  if (rc) return rc;
  std::memset(out, 0, sizeof(*out));
  const int ns = (int)c->species.size();
  const int nd = DIAG_FIXED + 4 * ns;
  std::vector<double> h((size_t)nd + 4);
  if (c->f64) diag_sums<double>(c);
  else diag_sums<float>(c);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(h.data(), c->diag_d, (size_t)nd * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(c, hipMemcpyAsync(h.data() + nd, c->gauss_d, 4 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  rc = sync_check(c);
  if (rc) return rc;
  const double dens = (double)c->p.dens, cc = c->p.c;
  out->energy_e = 0.5 * h[0];
  out->energy_b = 0.5 * h[1];
  for (int s = 0; s < ns; ++s) {
    const double m = c->species[s].mass;
    const double* k = h.data() + DIAG_FIXED + 4 * s;
    if (s < 8) out->energy_kin[s] = m * cc * cc * k[0] / dens;
    for (int d = 0; d < 3; ++d) out->momentum[d] += m * k[1 + d] / dens;
  }
  const double* g = h.data() + nd;
  const double rho2 = g[2];
  out->gauss_pre = rho2 > 0 ? std::sqrt(g[0] / rho2) : std::sqrt(g[0]);
  out->gauss_post = rho2 > 0 ? std::sqrt(g[1] / rho2) : std::sqrt(g[1]);
  out->gauss_post_max = g[3];
  out->steps = c->steps;
  out->sorts = c->nsorts;
  return SP_OK;
}


// ---- checkpoint (src/save/mod.rs:25-81 writes species 0 and the fields only;
// a restart also needs every species and the persistent spectral B_wrk,
// src/flds/mod.rs:31-33) ----
namespace {

struct CkptHeader {
  char magic[8];  // "SPCKPT\0\1"
  int32_t esz;    // 4 (SP_F32) or 8 (SP_F64)
  int32_t nspecies;
  sp_params params;
  sp_tuning tuning;
  int64_t steps;
  double gauss_post_max;
};
const char CKPT_MAGIC[8] = {'S', 'P', 'C', 'K', 'P', 'T', 0, 1};
constexpr size_t CKPT_CHUNK = (size_t)64 << 20;  // host staging per copy

bool same_params(const sp_params& a, const sp_params& b) {
  return a.size_x == b.size_x && a.size_y == b.size_y && a.delta == b.delta && a.dens == b.dens &&
         a.n_pass == b.n_pass && a.dt == b.dt && a.c == b.c && a.gamma_inj == b.gamma_inj &&
         a.precision == b.precision && a.periodic_x == b.periodic_x && a.checked == b.checked &&
         a.left_edge == b.left_edge && a.sort_interval == b.sort_interval && a.study == b.study;
}

// device <-> file in chunks through one host buffer
int dev_to_file(sp_ctx* c, FILE* f, const void* dev, size_t bytes, std::vector<unsigned char>& stage) {
  for (size_t off = 0; off < bytes; off += CKPT_CHUNK) {
    const size_t k = std::min(CKPT_CHUNK, bytes - off);
    HIPCHK(c, hipMemcpy(stage.data(), (const unsigned char*)dev + off, k, hipMemcpyDeviceToHost));
    if (std::fwrite(stage.data(), 1, k, f) != k) return fail(c, SP_EINVAL, "checkpoint: write failed");
  }
  return SP_OK;
}
// `check` (may be empty) sees every staged chunk of an int32 array before it
// is copied to the device and rejects the file by returning false.
int file_to_dev(sp_ctx* c, FILE* f, void* dev, size_t bytes, std::vector<unsigned char>& stage,
                const std::function<bool(const int32_t*, size_t)>& check = {}) {
  for (size_t off = 0; off < bytes; off += CKPT_CHUNK) {
    const size_t k = std::min(CKPT_CHUNK, bytes - off);
    if (std::fread(stage.data(), 1, k, f) != k) return fail(c, SP_EINVAL, "checkpoint: file is truncated");
    if (check && !check((const int32_t*)stage.data(), k / 4))
      return fail(c, SP_EINVAL, "checkpoint: particle cells or tags out of range");
    HIPCHK(c, hipMemcpy((unsigned char*)dev + off, stage.data(), k, hipMemcpyHostToDevice));
  }
  return SP_OK;
}

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};

}  // namespace

int sp_checkpoint_save(sp_ctx* c, const char* path) {
  if (!c || !path) return SP_EINVAL;
  int rc = flush(c);
  if (rc) return rc;
  full_spectra(c, HALF_SPECTRA);  // the file holds whole spectra
  rc = sync_check(c);
  if (rc) return rc;
  CkptHeader h{};
  std::memcpy(h.magic, CKPT_MAGIC, 8);
  h.esz = (int32_t)c->esz;
  h.nspecies = (int32_t)c->species.size();
  h.params = c->p;
  h.tuning = c->tune;
  h.steps = c->steps;
  HIPCHK(c, hipMemcpy(&h.gauss_post_max, c->gauss_d + 3, sizeof(double), hipMemcpyDeviceToHost));
  FileCloser fc{std::fopen(path, "wb")};
  if (!fc.f) return fail(c, SP_EINVAL, std::string("checkpoint: cannot open ") + path);
  if (std::fwrite(&h, sizeof(h), 1, fc.f) != 1) return fail(c, SP_EINVAL, "checkpoint: write failed");
  std::vector<unsigned char> stage(CKPT_CHUNK);
  for (const SpeciesBuf& s : c->species) {
    const double cm[2] = {s.charge, s.mass};
    if (std::fwrite(cm, sizeof(cm), 1, fc.f) != 1 || std::fwrite(&s.n, sizeof(s.n), 1, fc.f) != 1)
      return fail(c, SP_EINVAL, "checkpoint: write failed");
    // device order (the tile-sorted order the next passes read) with the tags
    const void* arr[7] = {s.cur.cell, s.cur.tag, s.cur.dx, s.cur.dy, s.cur.px, s.cur.py, s.cur.pz};
    for (int k = 0; k < 7; ++k)
      if ((rc = dev_to_file(c, fc.f, arr[k], (size_t)s.n * (k < 2 ? 4 : c->esz), stage))) return rc;
  }
  for (int g = 0; g < 11; ++g)
    if ((rc = dev_to_file(c, fc.f, c->grid[g], c->ngrid_elems() * c->esz, stage))) return rc;
  for (int g = 0; g < 14; ++g)
    if ((rc = dev_to_file(c, fc.f, c->spec[g], c->nspec_elems() * 2 * c->esz, stage))) return rc;
  if (std::fflush(fc.f) != 0) return fail(c, SP_EINVAL, "checkpoint: write failed");
  return SP_OK;
}

int sp_checkpoint_load(sp_ctx* c, const char* path) {
  if (!c || !path) return SP_EINVAL;
  if (c->ens) return fail(c, SP_EINVAL, "the context belongs to an ensemble");
  if (!c->species.empty()) return fail(c, SP_EINVAL, "checkpoint: load into a context without species");
  int rc = sync_check(c);
  if (rc) return rc;
  FileCloser fc{std::fopen(path, "rb")};
  if (!fc.f) return fail(c, SP_EINVAL, std::string("checkpoint: cannot open ") + path);
  CkptHeader h{};
  if (std::fread(&h, sizeof(h), 1, fc.f) != 1 || std::memcmp(h.magic, CKPT_MAGIC, 8) != 0)
    return fail(c, SP_EINVAL, "checkpoint: not a checkpoint file");
  if (h.esz != (int32_t)c->esz || !same_params(h.params, c->p))
    return fail(c, SP_EINVAL, "checkpoint: parameters differ from the context's");
  if (h.nspecies < 0 || h.nspecies > MAX_SPECIES) return fail(c, SP_EINVAL, "checkpoint: bad species count");
  if ((rc = check_tuning(c, h.tuning))) return rc;
  std::vector<unsigned char> stage(CKPT_CHUNK);
  for (int i = 0; i < h.nspecies; ++i) {
    double cm[2];
    int64_t n = 0;
    if (std::fread(cm, sizeof(cm), 1, fc.f) != 1 || std::fread(&n, sizeof(n), 1, fc.f) != 1 || n < 0 ||
        n > INT32_MAX) {
      for (auto& s : c->species) free_species(s);
      c->species.clear();
      return fail(c, SP_EINVAL, "checkpoint: bad species record");
    }
    SpeciesBuf s;
    s.charge = cm[0];
    s.mass = cm[1];
    rc = alloc_species(c, s, n);
    if (rc) {
      free_species(s);
      for (auto& t : c->species) free_species(t);
      c->species.clear();
      return rc;
    }
    c->species.push_back(s);
    void* arr[7] = {s.cur.cell, s.cur.tag, s.cur.dx, s.cur.dy, s.cur.px, s.cur.py, s.cur.pz};
    // cells inside the padded grid (the kernels index the grids with them),
    // tags a permutation of 0..n-1 (the getters place values by tag)
    const int nxp = c->p.size_x + 1, nyp = c->p.size_y + 1;
    auto cells_ok = [&](const int32_t* v, size_t m) {
      for (size_t j = 0; j < m; ++j) {
        const int ix = cell_ix(v[j]), iy = cell_iy(v[j]);
        if (ix > nxp || iy < 0 || iy > nyp) return false;
      }
      return true;
    };
    std::vector<bool> seen((size_t)n, false);
    auto tags_ok = [&](const int32_t* v, size_t m) {
      for (size_t j = 0; j < m; ++j) {
        if (v[j] < 0 || v[j] >= n || seen[(size_t)v[j]]) return false;
        seen[(size_t)v[j]] = true;
      }
      return true;
    };
    for (int k = 0; k < 7; ++k) {
      const std::function<bool(const int32_t*, size_t)> chk =
          k == 0 ? std::function<bool(const int32_t*, size_t)>(cells_ok)
                 : (k == 1 ? std::function<bool(const int32_t*, size_t)>(tags_ok) : nullptr);
      if ((rc = file_to_dev(c, fc.f, arr[k], (size_t)n * (k < 2 ? 4 : c->esz), stage, chk))) {
        for (auto& t : c->species) free_species(t);
        c->species.clear();
        return rc;
      }
    }
  }
  for (int g = 0; g < 11 && !rc; ++g) rc = file_to_dev(c, fc.f, c->grid[g], c->ngrid_elems() * c->esz, stage);
  for (int g = 0; g < 14 && !rc; ++g) rc = file_to_dev(c, fc.f, c->spec[g], c->nspec_elems() * 2 * c->esz, stage);
  c->half_spec = 0;  // the file holds whole spectra
  if (rc) {
    for (auto& t : c->species) free_species(t);
    c->species.clear();
    return rc;
  }
  apply_tuning(c, h.tuning);
  c->steps = h.steps;
  c->push_pending = false;
  // the adaptive sort schedule counts from here (the file's order is the
  // device order it was saved in); a fixed schedule follows from the step
  c->last_sort = c->steps;
  c->slow_f0 = -1;
  std::fill(c->slow_step, c->slow_step + SLOW_RING, (int64_t)-2);
  drop_graph(c);
  HIPCHK(c, hipMemcpy(c->gauss_d + 3, &h.gauss_post_max, sizeof(double), hipMemcpyHostToDevice));
  return SP_OK;
}

int sp_set_tuning(sp_ctx* c, const sp_tuning* t) {
  if (!c || !t) return SP_EINVAL;
  int rc = check_tuning(c, *t);
  if (rc) return rc;
  rc = flush(c);
  if (rc) return rc;
  apply_tuning(c, *t);
  return SP_OK;
}

int sp_get_tuning(const sp_ctx* c, sp_tuning* t) {
  if (!c || !t) return SP_EINVAL;
  *t = c->tune;
  return SP_OK;
}

int sp_autotune(sp_ctx* c, int trials, int use_cache, sp_tuning* out) {
  ENTRY_PRELUDE(c);
  if (trials < 1 || trials > 100) return fail(c, SP_EINVAL, "trials must be in 1..100");
  if (c->species.empty()) return fail(c, SP_EINVAL, "the autotuner needs the context's species");
  const std::vector<int64_t> key = tune_key(c);
  if (use_cache) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      apply_tuning(c, it->second);
      if (out) *out = c->tune;
      return SP_OK;
    }
  }
  sp_ctx* t = nullptr;
  int rc = clone_ctx(c, &t);
  if (rc) return rc;
  t->tune = c->tune;
  const sp_tuning best = c->f64 ? autotune_on<double>(t, trials) : autotune_on<float>(t, trials);
  const hipError_t e = hipStreamSynchronize(t->st);
  sp_destroy(t);
  if (e != hipSuccess) return fail(c, SP_EHIP, std::string("autotune: ") + hipGetErrorString(e));
  rc = check_tuning(c, best);
  if (rc) return rc;
  apply_tuning(c, best);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
  }
  if (out) *out = best;
  return SP_OK;
}

int sp_ensemble_create(sp_ctx* const* members, int n, int branches, int unit, sp_ensemble** out) {
  g_create_error.clear();
  if (!members || !out || n < 1) return fail(nullptr, SP_EINVAL, "null argument or empty ensemble");
  if (branches < 1 || branches > 16) return fail(nullptr, SP_EINVAL, "branches must be in 1..16");
  if (unit < 1 || unit > 64) return fail(nullptr, SP_EINVAL, "unit must be in 1..64");
  for (int i = 0; i < n; ++i) {
    sp_ctx* c = members[i];
    if (!c) return fail(nullptr, SP_EINVAL, "null member");
    if (c->ens) return fail(nullptr, SP_EINVAL, "a member already belongs to an ensemble");
    if (c->timing || c->graph_on) return fail(nullptr, SP_EINVAL, "members must have timing and graph off");
    if (c->device != members[0]->device) return fail(nullptr, SP_EINVAL, "members must share one device");
    for (int j = 0; j < i; ++j)
      if (members[j] == c) return fail(nullptr, SP_EINVAL, "a context appears twice");
  }
  sp_ensemble* e = new sp_ensemble();
  e->m.assign(members, members + n);
  e->nb = std::min(branches, n);
  e->unit = unit;
  auto bail = [&](const char* what, hipError_t r) {
    for (auto s : e->bst) hipStreamDestroy(s);
    for (auto v : e->join) hipEventDestroy(v);
    if (e->fork) hipEventDestroy(e->fork);
    if (e->st) hipStreamDestroy(e->st);
    delete e;
    return fail(nullptr, SP_EHIP, std::string(what) + ": " + hipGetErrorString(r));
  };
  hipError_t r = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking);
  if (r != hipSuccess) return bail("hipStreamCreateWithFlags", r);
  for (int b = 0; b < e->nb; ++b) {
    hipStream_t s;
    hipEvent_t v;
    r = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (r != hipSuccess) return bail("hipStreamCreateWithFlags", r);
    e->bst.push_back(s);
    r = hipEventCreateWithFlags(&v, hipEventDisableTiming);
    if (r != hipSuccess) return bail("hipEventCreateWithFlags", r);
    e->join.push_back(v);
  }
  r = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
  if (r != hipSuccess) return bail("hipEventCreateWithFlags", r);
  // branches balanced by particle count: largest member to the lightest branch
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  auto load = [&](int i) {
    int64_t k = 0;
    for (const auto& s : members[i]->species) k += s.n;
    return k + (int64_t)members[i]->ngrid_elems();
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return load(a) > load(b); });
  std::vector<int64_t> w(e->nb, 0);
  e->branch_of.assign(n, 0);
  for (int i : order) {
    const int b = (int)(std::min_element(w.begin(), w.end()) - w.begin());
    e->branch_of[i] = b;
    w[b] += load(i);
  }
  for (sp_ctx* c : e->m) {
    r = hipStreamSynchronize(c->st);
    if (r != hipSuccess) return bail("hipStreamSynchronize", r);
  }
  for (sp_ctx* c : e->m) {
    c->st = e->st;
    c->ens = e;
  }
  *out = e;
  return SP_OK;
}

int sp_ensemble_destroy(sp_ensemble* e) {
  if (!e) return SP_OK;
  hipStreamSynchronize(e->st);
  drop_kept(e);
  for (sp_ctx* c : e->m) {
    c->st = c->own_st;
    c->ens = nullptr;
  }
  for (auto s : e->bst) hipStreamDestroy(s);
  for (auto v : e->join) hipEventDestroy(v);
  hipEventDestroy(e->fork);
  hipStreamDestroy(e->st);
  delete e;
  return SP_OK;
}

const char* sp_ensemble_last_error(const sp_ensemble* e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int sp_ensemble_set_graph(sp_ensemble* e, int on) {
  if (!e) return SP_EINVAL;
  e->graph_on = on != 0;
  if (!e->graph_on) drop_kept(e);
  return SP_OK;
}

int sp_ensemble_step(sp_ensemble* e, int n) {
  if (!e) return SP_EINVAL;
  if (n < 0) return efail(e, SP_EINVAL, "negative step count");
  for (sp_ctx* c : e->m)
    if (c->timing || c->graph_on) return efail(e, SP_EINVAL, "members must have timing and graph off");
  for (int i = 0; i < n;) {
    if (e->graph_on && n - i >= e->unit) {
      const int rc = ens_graph_unit(e);
      if (rc) return rc;
      i += e->unit;
    } else {
      ens_step_forked(e);
      ++i;
    }
  }
  if (n > 0)
    for (sp_ctx* c : e->m) c->half_spec |= HALF_SPECTRA;  // graph replays skip the host side of fields_update
  const hipError_t r = hipGetLastError();
  if (r != hipSuccess) return efail(e, SP_EHIP, hipGetErrorString(r));
  return SP_OK;
}

int sp_ensemble_sync(sp_ensemble* e, int* bad_member) {
  if (!e) return SP_EINVAL;
  if (bad_member) *bad_member = -1;
  for (size_t i = 0; i < e->m.size(); ++i) {
    sp_ctx* c = e->m[i];
    int rc = flush(c);
    if (!rc) rc = sync_check(c);
    if (rc) {
      if (bad_member) *bad_member = (int)i;
      return efail(e, rc, "member " + std::to_string(i) + ": " + c->err);
    }
  }
  return SP_OK;
}

int sp_ensemble_stats(const sp_ensemble* e, int* kept, int* captures) {
  if (!e) return SP_EINVAL;
  if (kept) *kept = (int)e->kept.size();
  if (captures) *captures = e->captures;
  return SP_OK;
}

}  // extern "C"
