// particles.hip -- particle kernels for gfx950: relativistic Boris push with
// TSC gather, half-step mover with boundaries, TSC current/density deposit,
// and the particle initialiser.
//
// Floating-point contract: every expression keeps the association of the
// reference source (Rust never fuses a*b+c), so this file is compiled with
// contraction off and IEEE division/sqrt; cell indices, offsets and momenta
// come out bit-identical to the reference arithmetic.  Deposits are float
// atomics, so J/rho/dens differ from the reference's serial sums by
// summation order only.
#pragma clang fp contract(off)

#include "sp_internal.h"

namespace sp {

template <class T> __device__ __forceinline__ T rsignum(T x) {
  // Rust f32::signum: +1 for +0.0, -1 for -0.0, NaN for NaN.
  return isnan(x) ? x : copysign(T(1), x);
}

// The nine TSC weights of src/lib.rs:339-347 / src/prtls.rs:321-329.  Each
// reference weight is a left-to-right product chain; the y factor is the
// common prefix of its row, so the chains are evaluated once per prefix and
// the results are bit-identical to the reference's.
template <class T> __device__ __forceinline__ void tsc(T dx, T dy, T (&w)[9]) {
  const T h = T(0.5), q = T(0.75);
  const T a = h - dx, b = h + dx, xm = q - dx * dx;
  const T y0 = (h * (h - dy)) * (h - dy);
  const T y1 = q - dy * dy;
  const T y2 = (h * (h + dy)) * (h + dy);
  T t = y0 * h;
  w[0] = (t * a) * a; w[1] = y0 * xm; w[2] = (t * b) * b;
  t = y1 * h;
  w[3] = (t * a) * a; w[4] = y1 * xm; w[5] = (t * b) * b;
  t = y2 * h;
  w[6] = (t * a) * a; w[7] = y2 * xm; w[8] = (t * b) * b;
}

template <class T> __device__ __forceinline__ T gamma_of(T px, T py, T pz, T csqinv) {
  return sqrt(T(1.0) + (px * px + py * py + pz * pz) * csqinv);  // src/prtls.rs:440
}

// Prtl::update_position then Prtl::apply_bc for one particle
// (src/prtls.rs:85-121, :123-190).  `c1` = (0.5 dt) * (1/psa).
template <class T>
__device__ __forceinline__ void move_half(int& ix, int& iy, T& dx, T& dy, T& px, T py, T c1, const Geo<T>& g,
                                          int& bad) {
  if (ix <= 0 || iy <= 0) bad |= F_CELL;
  dx += c1 * px;
  if (dx >= T(0.5)) { dx -= T(1.0); ix += 1; }
  else if (dx < T(-0.5)) { dx += T(1.0); ix -= 1; }
  dy += c1 * py;
  if (dy >= T(0.5)) { dy -= T(1.0); iy += 1; }
  else if (dy < T(-0.5)) { dy += T(1.0); iy -= 1; }
  if (iy < 1) iy += g.ny;
  else if (iy > g.ny) iy -= g.ny;
  if (g.periodic_x) {
    if (ix < 1) ix += g.nx;
    else if (ix > g.nx) ix -= g.nx;
    return;
  }
  if (g.checked) {
    if (ix > g.wall + 1) bad |= F_WALL;
    if (g.left_edge == SP_LEFT_ASSERT && ix <= g.delta) bad |= F_LEFT;
  }
  if (ix >= g.wall) {
    if (ix == g.wall) {
      const T s = rsignum(dx);
      px *= -s;
      dx *= -s;
    } else {
      dx += T(1.0);
      ix -= 1;
      px *= T(-1.0);
    }
  }
  if (g.left_edge == SP_LEFT_REFLECT && ix <= g.lwall) {
    if (ix == g.lwall) {
      const T s = rsignum(dx);
      px *= s;
      dx *= s;
    } else {
      dx -= T(1.0);
      ix += 1;
      px *= T(-1.0);
    }
  }
}

template <class T> __device__ __forceinline__ bool inside(int ix, int iy, const Geo<T>& g) {
  return ix >= 1 && ix <= g.nx && iy >= 1 && iy <= g.ny;
}

template <class T> __device__ __forceinline__ void atomic_add(T* p, T v) { unsafeAtomicAdd(p, v); }

// Sim::deposit_current (src/lib.rs:266-393): v = q p / gamma, J += w v on the
// 3x3 stencil around (ix, iy), ghosts included.
template <class T>
__device__ __forceinline__ void deposit_current(int ix, int iy, T dx, T dy, T px, T py, T pz, T psa, T q,
                                                const Geo<T>& g, T* jx, T* jy, T* jz, int& bad) {
  if (!inside(ix, iy, g)) { bad |= F_CELL; return; }
  const T inv = T(1) / psa;
  const T vx = q * px * inv, vy = q * py * inv, vz = q * pz * inv;
  T w[9];
  tsc(dx, dy, w);
  const int base = (iy - 1) * g.pitch + ix - 1;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int o = base + a * g.pitch + b;
      atomic_add(jx + o, w[3 * a + b] * vx);
      atomic_add(jy + o, w[3 * a + b] * vy);
      atomic_add(jz + o, w[3 * a + b] * vz);
    }
}

// Sim::calc_density (src/lib.rs:395-476): rho += w q, dens += w.
template <class T>
__device__ __forceinline__ void deposit_density(int ix, int iy, T dx, T dy, T q, const Geo<T>& g, T* rho, T* dens,
                                                int& bad) {
  if (!inside(ix, iy, g)) { bad |= F_CELL; return; }
  T w[9];
  tsc(dx, dy, w);
  const int base = (iy - 1) * g.pitch + ix - 1;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int o = base + a * g.pitch + b;
      atomic_add(rho + o, w[3 * a + b] * q);
      atomic_add(dens + o, w[3 * a + b]);
    }
}

// Prtl::boris_push for one particle (src/prtls.rs:306-440): ordered 9-term
// TSC gather of E and B, then the relativistic Boris rotation.
template <class T>
__device__ __forceinline__ void boris(int ix, int iy, T dx, T dy, T& px, T& py, T& pz, const GridSet& gs,
                                      const Geo<T>& g, T alpha, T beta, int& bad) {
  if (!inside(ix, iy, g)) { bad |= F_CELL; return; }
  T w[9];
  tsc(dx, dy, w);
  const int P = g.pitch;
  const int base = (iy - 1) * P + ix - 1;
  const int off[9] = {0, 1, 2, P, P + 1, P + 2, 2 * P, 2 * P + 1, 2 * P + 2};
  T f[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const T* __restrict__ a = (const T*)gs.g[k] + base;  // Ex Ey Ez Bx By Bz
    T s = w[0] * a[off[0]];
#pragma unroll
    for (int m = 1; m < 9; ++m) s += w[m] * a[off[m]];
    f[k] = s;
  }
  const T ext = f[0] * beta, eyt = f[1] * beta, ezt = f[2] * beta;
  T bxt = f[3] * alpha, byt = f[4] * alpha, bzt = f[5] * alpha;
  const T ux = px + ext, uy = py + eyt, uz = pz + ezt;
  const T gt = T(1) / sqrt(T(1.) + (ux * ux + uy * uy + uz * uz) * g.csqinv);
  bxt *= gt;
  byt *= gt;
  bzt *= gt;
  const T bor = T(2.0) * (T(1) / (T(1.0) + bxt * bxt + byt * byt + bzt * bzt));
  const T uxt = ux + uy * bzt - uz * byt;
  const T uyt = uy + uz * bxt - ux * bzt;
  const T uzt = uz + ux * byt - uy * bxt;
  px = ux + bor * (uyt * bzt - uzt * byt) + ext;
  py = uy + bor * (uzt * bxt - uxt * bzt) + eyt;
  pz = uz + bor * (uxt * byt - uyt * bxt) + ezt;
}

template <class T> __device__ __forceinline__ bool finite3(T a, T b, T c) {
  return isfinite(a) && isfinite(b) && isfinite(c);
}

// Sim::move_and_deposit (src/lib.rs:478-489) for one particle.
template <class T>
__device__ __forceinline__ void move_and_deposit(int& ix, int& iy, T& dx, T& dy, T& px, T py, T pz, T q,
                                                 const Geo<T>& g, const GridSet& gs, int& bad) {
  const T psa = gamma_of(px, py, pz, g.csqinv);
  const T c1 = g.half_dt * (T(1) / psa);
  move_half(ix, iy, dx, dy, px, py, c1, g, bad);
  deposit_current(ix, iy, dx, dy, px, py, pz, psa, q, g, (T*)gs.g[SP_JX], (T*)gs.g[SP_JY], (T*)gs.g[SP_JZ], bad);
  move_half(ix, iy, dx, dy, px, py, c1, g, bad);
  deposit_density(ix, iy, dx, dy, q, g, (T*)gs.g[SP_RHO], (T*)gs.g[SP_DENS], bad);
}

template <class T>
__global__ void __launch_bounds__(256) k_move_deposit(SpeciesDev<T> s, Geo<T> g, GridSet gs, int* flags) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int bad = 0;
  if (i < s.n) {
    const int32_t c = s.cell[i];
    int ix = cell_ix(c), iy = cell_iy(c);
    T dx = s.dx[i], dy = s.dy[i], px = s.px[i];
    const T py = s.py[i], pz = s.pz[i];
    move_and_deposit(ix, iy, dx, dy, px, py, pz, s.charge, g, gs, bad);
    s.cell[i] = make_cell(ix, iy);
    s.dx[i] = dx;
    s.dy[i] = dy;
    s.px[i] = px;
  }
  if (bad) atomicOr(flags, bad);
}

template <class T>
__global__ void __launch_bounds__(256) k_push(SpeciesDev<T> s, Geo<T> g, GridSet gs, int* flags) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int bad = 0;
  if (i < s.n) {
    const int32_t c = s.cell[i];
    T px = s.px[i], py = s.py[i], pz = s.pz[i];
    boris(cell_ix(c), cell_iy(c), s.dx[i], s.dy[i], px, py, pz, gs, g, s.alpha, s.beta, bad);
    if (g.checked && !finite3(px, py, pz)) bad |= F_FINITE;
    s.px[i] = px;
    s.py[i] = py;
    s.pz[i] = pz;
  }
  if (bad) atomicOr(flags, bad);
}

// boris_push of step t followed by move_and_deposit of step t+1: adjacent in
// the reference loop (src/lib.rs:171-173, :193-196) with nothing between them
// that reads or writes particles, so one pass reads and writes each particle
// once per step.
template <class T>
__global__ void __launch_bounds__(256) k_push_move_deposit(SpeciesDev<T> s, Geo<T> g, GridSet gs, int* flags) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int bad = 0;
  if (i < s.n) {
    const int32_t c = s.cell[i];
    int ix = cell_ix(c), iy = cell_iy(c);
    T dx = s.dx[i], dy = s.dy[i];
    T px = s.px[i], py = s.py[i], pz = s.pz[i];
    boris(ix, iy, dx, dy, px, py, pz, gs, g, s.alpha, s.beta, bad);
    if (g.checked && !finite3(px, py, pz)) bad |= F_FINITE;
    move_and_deposit(ix, iy, dx, dy, px, py, pz, s.charge, g, gs, bad);
    s.cell[i] = make_cell(ix, iy);
    s.dx[i] = dx;
    s.dy[i] = dy;
    s.px[i] = px;
    s.py[i] = py;
    s.pz[i] = pz;
  }
  if (bad) atomicOr(flags, bad);
}

// ---------------------------------------------------------------- initialiser
__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double unif(uint64_t key, uint64_t i, int k) {  // [0, 1)
  return (double)(mix64(key ^ mix64(i * 8ull + (uint64_t)k)) >> 11) * 0x1.0p-53;
}

// Prtl::new on the device: initialize_positions (src/prtls.rs:191-221)
// exactly, initialize_velocities (src/prtls.rs:222-250) with counter-based
// draws: three Box-Muller normals and one uniform per particle.
template <class T>
__global__ void __launch_bounds__(256) k_init(SpeciesDev<T> s, Geo<T> g, int dens, T vth, T gamma_inj, uint64_t key) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  const int64_t per_row = (int64_t)(g.nx - 2 * g.delta) * dens;
  const int row = (int)(i / per_row);
  const int64_t rem = i - (int64_t)row * per_row;
  const int j = g.delta + (int)(rem / dens);
  const int k = (int)(rem % dens);
  T r1 = T(1.0) / (T(2.0) * T(dens));
  r1 += T(k) / T(dens);
  const T d = r1 - T(0.5);
  s.cell[i] = make_cell(j + 1, row + 1);
  s.dx[i] = d;
  s.dy[i] = d;
  s.tag[i] = (int32_t)i;
  const double twopi = 6.283185307179586476925286766559;
  const double u1 = 1.0 - unif(key, i, 0), u2 = unif(key, i, 1);
  const double u3 = 1.0 - unif(key, i, 2), u4 = unif(key, i, 3);
  const double r12 = sqrt(-2.0 * log(u1)), r34 = sqrt(-2.0 * log(u3));
  const T c = g.c;
  const T sig = vth * c;
  T px = T(r12 * cos(twopi * u2)) * sig;
  T py = T(r12 * sin(twopi * u2)) * sig;
  T pz = T(r34 * cos(twopi * u4)) * sig;
  T psa = sqrt(T(1.0) + (px * px + py * py + pz * pz) * g.csqinv);
  const T beta_inj = sqrt(T(1.) - T(1) / (gamma_inj * gamma_inj));
  T ux = px / c;
  const T rnd = T(unif(key, i, 4));
  if (-beta_inj * ux > rnd * psa) ux *= T(-1.);
  px = gamma_inj * (ux + beta_inj * psa);
  px *= c;
  s.px[i] = px;
  s.py[i] = py;
  s.pz[i] = pz;
}

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <class T>
void launch_move_deposit(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags, hipStream_t st) {
  if (s.n == 0) return;
  hipLaunchKernelGGL(k_move_deposit<T>, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, s, g, gs, flags);
}
template <class T>
void launch_push(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags, hipStream_t st) {
  if (s.n == 0) return;
  hipLaunchKernelGGL(k_push<T>, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, s, g, gs, flags);
}
template <class T>
void launch_push_move_deposit(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags,
                              hipStream_t st) {
  if (s.n == 0) return;
  hipLaunchKernelGGL(k_push_move_deposit<T>, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, s, g, gs, flags);
}
template <class T>
void launch_init_species(const SpeciesDev<T>& s, const Geo<T>& g, double vth, double gamma_inj, uint64_t seed,
                         int sid, hipStream_t st) {
  if (s.n == 0) return;
  // dens is recovered from the count: n = dens * (nx - 2 delta) * ny
  const int64_t cells = (int64_t)(g.nx - 2 * g.delta) * g.ny;
  const int dens = (int)(s.n / cells);
  uint64_t key = seed * 0x100000001B3ull + (uint64_t)(sid + 1) * 0x9E3779B97F4A7C15ull;
  hipLaunchKernelGGL(k_init<T>, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, s, g, dens, (T)vth, (T)gamma_inj,
                     key);
}

#define INST(T)                                                                                               \
  template void launch_move_deposit<T>(const SpeciesDev<T>&, const Geo<T>&, const GridSet&, int*, hipStream_t); \
  template void launch_push<T>(const SpeciesDev<T>&, const Geo<T>&, const GridSet&, int*, hipStream_t);         \
  template void launch_push_move_deposit<T>(const SpeciesDev<T>&, const Geo<T>&, const GridSet&, int*,          \
                                            hipStream_t);                                                     \
  template void launch_init_species<T>(const SpeciesDev<T>&, const Geo<T>&, double, double, uint64_t, int,      \
                                       hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace sp
