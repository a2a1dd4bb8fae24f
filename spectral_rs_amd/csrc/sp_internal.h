// sp_internal.h -- device-side types shared by the HIP translation units.
//
// Layout in HBM (one context = one simulation on one GPU):
//   * grids: padded (Ny+2) x (Nx+2) row-major arrays of Float, one per field
//     (Ex Ey Ez Bx By Bz Jx Jy Jz rho dens) plus a filter work grid;
//   * spectra: Ny x Nx interleaved complex, one per field plus B_wrk x3;
//   * particles: per species SoA of `cell` (int32, iy << 16 | ix), dx, dy,
//     px, py, pz (Float) and tag (int32, index in the order the species was
//     added, carried through cell sorts).  gamma is never stored: it is
//     sqrt(1 + |p|^2 csqinv), the exact expression the reference stores in psa.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/spectral_rs.h"

namespace sp {

enum Flags : int {
  F_CELL = 1,   // ix or iy outside the interior (src/prtls.rs:100-101, src/lib.rs:324-325)
  F_WALL = 2,   // max ix > wall + 1 (src/prtls.rs:155-162)
  F_LEFT = 4,   // min ix <= delta (src/prtls.rs:163-170)
  F_FINITE = 8  // non-finite momentum or gamma (no reference assert; reported as invariant)
};

// The flags buffer of a context holds 4 ints: [0] the flags above, [2..3] a
// running 64-bit count of the particles the particle pass took through its
// general path (chunks that were not quiet), which drives the adaptive
// tile-sort schedule; SLOW_RING snapshots of it are kept on the host.
constexpr int SLOW_COUNTER = 2;
constexpr int SLOW_RING = 8;

template <class T> struct Geo {
  int nx, ny, pitch;       // pitch = nx + 2
  int delta, wall, lwall;  // wall = nx + 1 - delta (src/prtls.rs:153), lwall = delta + 1
  int periodic_x, checked, left_edge;
  int lds_ok;              // dt * c < 1: a particle moves less than one cell per step
  int direct;              // particles are kept tile-sorted and interleaved: the particle
                           // pass reads them in place (no LDS staging or lane permutation)
  int chunk;               // particles per workgroup of the particle pass (host-side choice)
  int abs_pos32;           // precision study: keep positions to absolute-FP32 resolution
  T dt, c, csqinv;         // csqinv = 1 / (c * c) in Float (src/prtls.rs:265)
  T half_dt;               // 0.5 * dt in Float (src/prtls.rs:103)
};

template <class T> struct SpeciesDev {
  int64_t n;
  int32_t* cell;
  T *dx, *dy, *px, *py, *pz;
  int32_t* tag;
  T charge, alpha, beta;
};

struct GridSet {  // device pointers of one precision, indexed by sp_grid
  void* g[11];
  // one species' density, zero between particle passes: a pass deposits sum w
  // here and its launcher then adds q * it to rho and it to dens (and clears it)
  void* dsp;
};

// Workgroups are dealt to the MI355X's 8 XCDs round-robin (workgroup b runs on
// XCD b % 8), each XCD with its own L2.  This maps b to a logical index such
// that every XCD walks one contiguous run of logical indices (a bijection of
// [0, n) for any n), so neighbouring work shares an L2.
__device__ inline int xcd_remap(int b, int n) {
  const int x = b & 7, q = b >> 3, full = n >> 3, rem = n & 7;
  return x * full + (x < rem ? x : rem) + q;
}

__host__ __device__ inline int cell_ix(int32_t c) { return c & 0xFFFF; }
__host__ __device__ inline int cell_iy(int32_t c) { return c >> 16; }
__host__ __device__ inline int32_t make_cell(int ix, int iy) { return (iy << 16) | (ix & 0xFFFF); }

// ---- launchers (all asynchronous on `st`) ----
template <class T> void fill(T* p, T v, int64_t n, hipStream_t st);
// fill of up to 8 arrays of n elements in one launch
template <class T> void fill_many(T* const* ptrs, int count, T v, int64_t n, hipStream_t st);
// The particle pass of one species; its density goes to gs.dsp, which
// `fold` then adds to rho (times the charge) and dens and clears.
template <class T> void launch_move_deposit(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags,
                                            hipStream_t st, bool fold = true);
template <class T> void launch_push(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags,
                                   hipStream_t st);
template <class T> void launch_push_move_deposit(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs,
                                                 int* flags, hipStream_t st, bool fold = true);
// rho += q_k dsp_k, dens += dsp_k, dsp_k = 0 for k = 0 .. n - 1 in order, over
// the padded grids (the folds of n passes run with fold = false, one launch
// per 4 species).
template <class T> void fold_species_densities(const GridSet& gs, const Geo<T>& g, void* const* dsp,
                                               const double* q, int n, hipStream_t st);
// Cell-tile sort keeping each particle chunk spatially compact (see
// k_particles): tiles of 2^twl x 2^thl cells, tile rows in serpentine order.
// `cursor` holds one int per tile, `count` one per tile and then one per
// block of SCAN_BLK tiles (tile_count_size).
constexpr int TILE_LOG_MIN = 1, TILE_LOG_MAX = 5, SCAN_BLK = 1024;
inline size_t tile_count_size(size_t ntiles) { return ntiles + (ntiles + SCAN_BLK - 1) / SCAN_BLK; }
template <class T> void launch_tile_sort(const SpeciesDev<T>& src, const SpeciesDev<T>& dst, int nx, int ny,
                                         int twl, int thl, int* count, int* cursor, hipStream_t st);
template <class T> void launch_init_species(const SpeciesDev<T>& s, const Geo<T>& g, double vth, double gamma_inj,
                                            uint64_t seed, int sid, hipStream_t st);
template <class T> void launch_fold_ghosts(T* const* grids, int ngrids, int nx, int ny, hipStream_t st);
template <class T> void launch_update_ghosts(T* const* grids, int ngrids, int nx, int ny, hipStream_t st);
// `passes` binomial passes with ghost refresh; swaps grids[i] <-> work[i] so
// that grids[i] holds the result.
template <class T> void launch_binomial(void** grids, void** work, int ngrids, int nx, int ny, int passes,
                                        hipStream_t st);
// Columns nx/2 + 1 .. nx - 1 of the spectrum of a real field as the conjugate
// mirrors of the half plane (nx even).
template <class T> void launch_mirror_fill(void* spec, int nx, int ny, hipStream_t st);
// `gauss` (may be null) receives the FP64 Gauss-law sums of this solve:
// gauss[0..2] = sum |k.E + i rho|^2 before / after the correction and
// sum |rho|^2; gauss[3] keeps the largest relative post-correction residual
// (never cleared here); gauss[GAUSS_PARTIALS..] holds 3 partials per block of
// nx*ny/256 blocks.
constexpr int GAUSS_PARTIALS = 8;
// Species per context, and the layout of the diagnostic sums buffer:
// DIAG_FIXED field sums, then 4 per species.
constexpr int MAX_SPECIES = 64;
constexpr int DIAG_FIXED = 4;
constexpr int DIAG_DOUBLES = DIAG_FIXED + 4 * MAX_SPECIES;
inline size_t gauss_doubles(int nx, int ny) { return GAUSS_PARTIALS + 3 * (((size_t)nx * ny + 255) / 256); }
// Sums the three partials of nblk blocks (gauss + GAUSS_PARTIALS) into
// gauss[0..2] and keeps the largest relative post-correction residual
// sqrt(post / |rho|^2) seen so far in gauss[3]; run by every thread of one
// workgroup (a whole number of waves, at most 1024 threads).
__device__ inline void gauss_reduce_block(double* gauss, int nblk) {
  __shared__ double red[3][16];
  const double* part = gauss + GAUSS_PARTIALS;
  const int nt = (int)blockDim.x;
  double s[3] = {0, 0, 0};
  constexpr int U = 8;  // partials per thread in flight at a time, added in order
  for (int b0 = (int)threadIdx.x; b0 < nblk; b0 += U * nt) {
    double v[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * nt;
      if (b < nblk)
        for (int k = 0; k < 3; ++k) v[u][k] = part[(int64_t)b * 3 + k];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (b0 + u * nt < nblk)
        for (int k = 0; k < 3; ++k) s[k] += v[u][k];
    }
  }
  const int w = (int)threadIdx.x >> 6;
  for (int k = 0; k < 3; ++k) {
    double x = s[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) red[k][w] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[3];
    for (int k = 0; k < 3; ++k) {
      t[k] = 0;
      for (int i = 0; i < nt / 64; ++i) t[k] += red[k][i];
    }
    gauss[0] = t[0];
    gauss[1] = t[1];
    gauss[2] = t[2];
    const double rel = t[2] > 0 ? sqrt(t[1] / t[2]) : sqrt(t[1]);
    gauss[3] = fmax(gauss[3], rel);
  }
}
// `step`: the step's update on the real-field pipeline's spectra (projection of
// the self-mirrored columns, E and B on the half plane; see k_spectral_pair);
// otherwise every mode as given (sp_spectral_update).  Returns the number of
// blocks whose Gauss partials the update wrote; with `reduce` false the caller
// has them summed later (gauss_reduce_block, run by the first block of the
// step's inverse column transforms), else a single-workgroup launch sums them.
template <class T> int launch_spectral_update(void* const* spec, const Geo<T>& g, int dens, double* gauss,
                                              hipStream_t st, bool step, bool reduce = true);
// FP64 diagnostics: out[0] += sum over the interior of a^2 + b^2 + c^2;
// kinetic: out[0] += sum (gamma - 1), out[1..3] += sum p.
// sum32: accumulate in FP32 (precision study), the total still stored as FP64.
template <class T> void launch_sumsq3(const T* a, const T* b, const T* c, int nx, int ny, double* out,
                                      hipStream_t st, bool sum32 = false);
template <class T> void launch_kinetic(const SpeciesDev<T>& s, double csqinv, double* out, hipStream_t st,
                                       bool sum32 = false);

// FFT plan for one dimension: radices and an FP64-generated twiddle table.
struct FftDim {
  int n;
  int per_block;        // rows (x) / columns (y) per workgroup; 0 = the default
  int nrad;
  uint64_t radix_code;  // radix of pass p in bits [4p, 4p+4)
  void* tw;             // n complex Float: exp(-2 pi i k / n)
  // the twiddles of the current plan in the order its passes read them: pass p
  // (Ns_p = R_0 ... R_{p-1}) holds tw[r k n / (Ns_p R_p)] at Ns_p - 1 + (r - 1)
  // Ns_p + k, r = 1 .. R_p - 1, k < Ns_p (n - 1 entries for any plan)
  void* twp;
};
// fp32_gen: generate in FP32 (precision study) instead of FP64; also lays
// out twp for the current plan.
template <class T> void fft_make_twiddles(FftDim& d, hipStream_t st, bool fp32_gen = false);
// Lays out twp from tw for the plan in d (after a plan change).
template <class T> void fft_pass_twiddles(const FftDim& d, hipStream_t st);
// Row (x) and column (y) passes over up to 8 spectra at once.
enum FftLoad { LOAD_COMPLEX = 0, LOAD_REAL_PADDED = 1 };
enum FftStore { STORE_COMPLEX = 0, STORE_REAL_PADDED = 1 };
struct FftBatch {
  int count;
  void* spec[8];     // spectra (Ny x Nx complex)
  void* grid[8];     // padded real grids (for the fused load / store)
  double scale[8];   // multiplier applied at the fused real load
  // column transforms: when set, the first block first sums the spectral
  // update's Gauss partials (gauss_reduce_block) -- one launch less per step
  double* gauss;
  int gauss_nblk;
};
// `store_scale` multiplies every stored value (the 1/(Nx Ny) of an inverse).
template <class T> void launch_fft_rows(const FftBatch& b, const FftDim& dx, int ny, int inverse, int load, int store,
                                        double store_scale, hipStream_t st);
// Column transforms of the first ncols columns (0: all nx); mirror (forward):
// also write column nx - c as the conjugate mirror of column c.
template <class T> void launch_fft_cols(const FftBatch& b, const FftDim& dy, int nx, int inverse, double store_scale,
                                        hipStream_t st, int ncols = 0, bool mirror = false);
// The step's real-field pipeline (even nx and ny): forward rows, two real rows
// per complex transform, into the halves kx = 0 .. nx/2 of the spectrum rows;
// inverse rows from those halves back to the padded real grids.
template <class T> void launch_fft_rows_r2c(const FftBatch& b, const FftDim& dx, int ny, hipStream_t st);
template <class T> void launch_fft_rows_c2r(const FftBatch& b, const FftDim& dx, int ny, double store_scale,
                                            hipStream_t st);
// Largest transform length one block holds in LDS.
template <class T> constexpr int fft_max_len() { return (int)(65536 / (2 * sizeof(T))); }

// Learned current denoiser (denoise.hip): DN_LAYERS 3x3 convolutions,
// DN_IN -> DN_CH -> ... -> DN_CH -> DN_IN channels, GELU between them.
// Parameters, one FP32 vector: per layer l, w_l [cout][cin*9] then b_l
// [cout]; then the DN_IN channel scales (not trained).  w[co][ci*9 + ky*3 + kx].
constexpr int DN_IN = 4, DN_CH = 32, DN_LAYERS = 5;
__host__ __device__ constexpr int dn_cin(int l) { return l == 0 ? DN_IN : DN_CH; }
__host__ __device__ constexpr int dn_cout(int l) { return l == DN_LAYERS - 1 ? DN_IN : DN_CH; }
constexpr int dn_ntrain() {
  int n = 0;
  for (int l = 0; l < DN_LAYERS; ++l) n += dn_cout(l) * dn_cin(l) * 9 + dn_cout(l);
  return n;
}
constexpr int DN_NTRAIN = dn_ntrain();  // 30084
constexpr int DN_NPARAMS = DN_NTRAIN + DN_IN;
struct DnParams {
  float* w[DN_LAYERS];
  float* b[DN_LAYERS];
  float* scale;
};
__host__ __device__ inline DnParams dn_params(float* p) {
  DnParams d;
  for (int l = 0; l < DN_LAYERS; ++l) {
    d.w[l] = p;
    p += dn_cout(l) * dn_cin(l) * 9;
    d.b[l] = p;
    p += dn_cout(l);
  }
  d.scale = p;
  return d;
}
// Activations of B images of H x W: x, t (targets), r, dr (DN_IN channels);
// z[l] the pre-activations of layers 0..DN_LAYERS-2 and dza, dzb their
// gradients (DN_CH); wt the flipped weights of one layer (backward).  The
// simulation's fused path uses r and packed only.
struct DnTensors {
  float *x, *t, *r, *dr;
  float* z[DN_LAYERS - 1];
  float *dza, *dzb, *wt;
  short* packed;  // the fused path's BF16 weights (DN_PACK_SHORTS, 16-byte aligned)
};
// BF16 weights of all layers, each [16 ceil(cout/16)][cin*9 rounded up to 32]
// in tap-major K order, for the fused inference kernel.
constexpr int DN_PACK_SHORTS = 32 * 64 + 3 * 32 * 288 + 16 * 288;
void dn_forward(const DnParams& P, const DnTensors& T, int B, int H, int W, bool periodic_y, hipStream_t st);
void dn_backward(const DnParams& P, DnParams G, const DnTensors& T, const float* lw, int B, int H, int W,
                 double* loss, hipStream_t st);
void dn_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, int t, hipStream_t st);
// All layers of one image [4][H][W] in one launch, y periodic (the
// simulation's inference path); same results as dn_forward(periodic_y).
void dn_forward_fused(const DnParams& P, const DnTensors& T, int H, int W, hipStream_t st);
// Runs the fused network on the 4 input grids (read as grid / dens / s_c:
// the binomial-filtered grids, or the unfiltered deposit when the network
// replaces the filter) and, when the checks pass, writes input + dens s_c r
// into the filtered grids for the channels set in chmask (bit c = grid c of
// Jx Jy Jz rho); otherwise the filtered grids stay.  T.packed and T.r are its
// buffers; sums holds DN_SUMS doubles; stats (DN_STATS ints, below) counts
// applications and fallbacks.  gauss_tol <= 0 and hp_tol <= 0 switch those
// checks off; after a fallback the next `backoff` solves skip the network (0:
// none skips).
constexpr int DN_SUMS = 10;  // |x|^2, |r|^2, then per channel the high-pass energies of output and filter
// device counters of the in-loop denoiser: solves that applied the network,
// that kept the filter's result (skipped ones included), the solves still to
// skip, the verdict of the last check (1 = failed), and the skipped solves
enum DnStat { DN_STAT_APPLIED = 0, DN_STAT_FALLBACK, DN_STAT_SKIP, DN_STAT_LAST, DN_STAT_SKIPPED, DN_STATS };
void dn_apply_sim(const DnParams& P, const DnTensors& T, float* const* inputs, float* const* grids, int nx, int ny,
                  int dens, double* sums, const double* gauss, float tol, float gauss_tol, float hp_tol, int chmask,
                  int backoff, int* stats, hipStream_t st);

}  // namespace sp
