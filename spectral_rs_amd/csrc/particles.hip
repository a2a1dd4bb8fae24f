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

#include <climits>
#include <limits>
#include <type_traits>

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

// IEEE-exact 1/x and sqrt(x) for finite normal x > 0, shorter than the
// compiler's general expansions (which also handle denormals, zero and
// infinity, none of which occur here: every argument is >= 1).
//  * rcp: one Newton step on v_rcp_f32 (error < 1 ulp) with fused
//    multiply-adds, y + y (1 - x y);
//  * sqrt: v_sqrt_f32 (error < 1 ulp), then the neighbour whose residual
//    x - s s' has the right sign, the correctly-rounding test of the
//    compiler's own expansion.
// Both were checked against IEEE division and sqrt on gfx950 for every float
// of six binades ([2^-20, 2^-19), [0.5, 4), [2^20, 2^21), [2^40, 2^41)) with
// no mismatch; the bit-exact push and mover tests cover them in use.
template <class T> __device__ __forceinline__ T rcp_exact(T x) { return T(1) / x; }
template <> __device__ __forceinline__ float rcp_exact<float>(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, y, 1.0f), y, y);
}
template <class T> __device__ __forceinline__ T sqrt_exact(T x) { return sqrt(x); }
template <> __device__ __forceinline__ float sqrt_exact<float>(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const int si = __builtin_bit_cast(int, s);
  const float lo = __builtin_bit_cast(float, si - 1), hi = __builtin_bit_cast(float, si + 1);
  const float rl = __builtin_fmaf(-lo, s, x), rh = __builtin_fmaf(-hi, s, x);
  float r = rl <= 0.0f ? lo : s;
  r = rh > 0.0f ? hi : r;
  return r;
}

template <class T> __device__ __forceinline__ T gamma_of(T px, T py, T pz, T csqinv) {
  return sqrt_exact(T(1.0) + (px * px + py * py + pz * pz) * csqinv);  // src/prtls.rs:440
}

// Prtl::update_position then Prtl::apply_bc for one particle
// (src/prtls.rs:85-121, :123-190).  `c1` = (0.5 dt) * (1/psa).  Q: the
// particle's chunk is quiet (see k_particles): no boundary is within reach this
// step, so apply_bc leaves it alone and only the cell crossing remains.
template <bool Q, class T>
__device__ __forceinline__ void move_half(int& ix, int& iy, T& dx, T& dy, T& px, T py, T c1, const Geo<T>& g,
                                          int& bad) {
  if (!Q && (ix <= 0 || iy <= 0)) bad |= F_CELL;
  // a crossing moves by one cell: d - 1 above 0.5, d - (-1) = d + 1 below
  // -0.5, d - 0 = d otherwise (all exact, as the reference's branches)
  dx += c1 * px;
  dy += c1 * py;
  if (!Q && sizeof(T) == 4 && g.abs_pos32) {
    // A wave-uniform branch taken only by the precision study: the empty
    // volatile asm keeps the compiler from computing the rounding on every
    // position update and selecting it away (16 VALU per particle).
    __asm__ volatile("" ::: "memory");
    // precision study: the position held as one absolute FP32 coordinate
    // (x = ix + dx rounded to float), the offset read back exactly
    dx = (T(ix) + dx) - T(ix);
    dy = (T(iy) + dy) - T(iy);
  }
  const int sx = (dx >= T(0.5)) - (dx < T(-0.5));
  dx -= T(sx);
  ix += sx;
  const int sy = (dy >= T(0.5)) - (dy < T(-0.5));
  dy -= T(sy);
  iy += sy;
  if (Q) return;
  iy += (iy < 1 ? g.ny : 0) - (iy > g.ny ? g.ny : 0);
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

template <class T> using lds_ptr = __attribute__((address_space(3))) T*;
template <class T> using glb_ptr = __attribute__((address_space(1))) T*;

// Raw buffer resources of the six field grids (Ex Ey Ez Bx By Bz), for the
// stand-alone push.  The gather then needs one address register per particle:
// the column offset is an instruction immediate and the row offset a scalar
// register, where global loads would need a 64-bit address add per stencil row
// and field.  Word 3 0x00020000 is the gfx9 raw-buffer format; reads past
// `bytes` return 0 (the callers only read inside the padded grid).
struct FieldBufs {
  __amdgpu_buffer_rsrc_t r[6];
  int row_bytes;
};
__device__ __forceinline__ FieldBufs field_bufs(const GridSet& gs, const Geo<float>& g) {
  FieldBufs b;
  const int bytes = (g.ny + 2) * g.pitch * (int)sizeof(float);
#pragma unroll
  for (int k = 0; k < 6; ++k) b.r[k] = __builtin_amdgcn_make_buffer_rsrc(gs.g[k], (short)0, bytes, 0x00020000);
  b.row_bytes = g.pitch * (int)sizeof(float);
  return b;
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// src/prtls.rs:351-411: each field is the ordered sum w0 a0 + w1 a1 + ... + w8 a8
// over the 3x3 stencil in row-major order around (ix, iy).
template <class T>
__device__ __forceinline__ void gather6(const GridSet& gs, const Geo<T>& g, int ix, int iy, const T (&w)[9],
                                        T (&f)[6]) {
  const int P = g.pitch;
  const int base = __mul24(iy - 1, P) + ix - 1;
  const int off[9] = {0, 1, 2, P, P + 1, P + 2, 2 * P, 2 * P + 1, 2 * P + 2};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const T* __restrict__ a = (const T*)gs.g[k] + base;
    T s = w[0] * a[off[0]];
#pragma unroll
    for (int m = 1; m < 9; ++m) s += w[m] * a[off[m]];
    f[k] = s;
  }
}
__device__ __forceinline__ void gather6(const FieldBufs& fb, const Geo<float>& g, int ix, int iy,
                                        const float (&w)[9], float (&f)[6]) {
  // sp_create caps FP32 grids at 8192 x 8192 (the FFT length limit), so a
  // grid has < 2^29 bytes; the mask states that and tells the compiler that vo + 8
  // cannot wrap, so the column offsets fold into the load's immediate and each
  // stencil row becomes one 3-dword load
  const int vo = ((__mul24(iy - 1, g.pitch) + ix - 1) * (int)sizeof(float)) & 0x1FFFFFFC;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    float a[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) a[3 * r + c] = bload(fb.r[k], vo + 4 * c, r * fb.row_bytes);
    float s = w[0] * a[0];
#pragma unroll
    for (int m = 1; m < 9; ++m) s += w[m] * a[m];
    f[k] = s;
  }
}

// i / w for 0 <= i < 2^22 and w >= 1 from rw = 1/w in FP32: (i + 1/2) / w lies
// at least 1/(2w) from an integer, and (i + 1/2) * rw (rw and the product each
// rounded once) is within 2^-23 (i + 1/2) / w < 1/(2w) of it.
__device__ __forceinline__ int row_of(int i, float rw) { return (int)(((float)i + 0.5f) * rw); }

// The six field grids over one chunk's cell box + 1 cell (every stencil of a
// particle in the box), staged in LDS by the FP32 fused pass.  They sit at
// fixed offsets in the tails of the deposit planes (fields 0-3 in the last
// FWIN words of planes 0-3, fields 4-5 in the FWIN words before those of
// planes 0-1), so the window is used when it has at most FWIN cells and the
// deposit copies fit in the first WMAX - 2 FWIN words of each plane.  Gathers
// from LDS replace 18 three-dword vector loads per particle, whose per-lane
// L1 lookups (lanes of a wave in scattered cells) bound the pass before.
#ifndef SP_FWIN
#define SP_FWIN 400
#endif
struct FieldWin {
  lds_ptr<int> p;  // base of the deposit planes
  int x0, y0, w, h;
  bool use;
};
template <int WMAX> __device__ __forceinline__ constexpr int fwin_off(int k) {
  return k < 4 ? k * WMAX + WMAX - SP_FWIN : (k - 4) * WMAX + WMAX - 2 * SP_FWIN;
}
// The same ordered 9-term sums as gather6, from the window when the stencil
// lies in it, else from the grids with global loads (plain pointers: buffer
// descriptors would hold 24 scalar registers through the particle loop).
template <bool Q, int WMAX>
__device__ __forceinline__ void gather6w(const GridSet& gs, const FieldWin& l, const Geo<float>& g, int ix, int iy,
                                         const float (&w)[9], float (&f)[6]) {
  if (Q || (l.use && ix - 1 >= l.x0 && ix + 1 < l.x0 + l.w && iy - 1 >= l.y0 && iy + 1 < l.y0 + l.h)) {
    const lds_ptr<int> p = l.p + (__mul24(iy - 1 - l.y0, l.w) + (ix - 1 - l.x0));
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float a[9];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[3 * r + c] = __builtin_bit_cast(float, (int)p[fwin_off<WMAX>(k) + r * l.w + c]);
      float s = w[0] * a[0];
#pragma unroll
      for (int m = 1; m < 9; ++m) s += w[m] * a[m];
      f[k] = s;
    }
  } else {
    gather6(gs, g, ix, iy, w, f);
  }
}

// Deposit target of one workgroup: a window [x0, x0+w) x [y0, y0+h) of the
// padded grids held in LDS (planes Jx Jy Jz rho dens, w*h cells each), plus
// the global grids for stencils that leave the window.  use_lds == false sends
// everything to the global grids.
// Pointers with their address space spelled out, so every atomic lowers to
// the native LDS (ds_add_f32/f64) or global (global_atomic_add_f32/f64)
// instruction rather than a flat atomic.
template <class T> __device__ __forceinline__ void lds_add(lds_ptr<T> p, T v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T> __device__ __forceinline__ void glb_add(glb_ptr<T> p, T v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS accumulation.  Float atomics in LDS run ~40x slower than integer ones
// on gfx950, so FP32 windows accumulate in int32 fixed point: a contribution
// x becomes round(x * 2^e), with e chosen per chunk and plane so that no
// window cell can overflow (see k_particles).  Integer sums are exact and
// order-independent; the only rounding is the per-contribution quantisation
// (|error| <= 2^-e, see rint_half_up) and the conversion of the window sum at
// the flush.  FP64 windows (the dprec precision) accumulate with native f64
// atomics.
template <class T> struct Acc;
template <> struct Acc<float> {
  using type = int;
  static constexpr bool fixed = true;
};
template <> struct Acc<double> {
  using type = double;
  static constexpr bool fixed = false;
};

// One workgroup's chunk: NT threads, K particles (default; the particle pass
// is also built for 512 and 1024, chosen per context by its tuning), deposit
// windows of at most WMAX cells per plane (planes at fixed strides of WMAX, so
// a plane offset is an instruction immediate).
#ifndef SP_WINDOW_COPIES
#define SP_WINDOW_COPIES 4
#endif
// 5 waves per SIMD: the FP32 pass fits 92 VGPRs without spills (100 when
// unconstrained, 4 waves) and 5 workgroups' windows fit the LDS; measured 1.6%
// faster than 4 waves and than 6 (with 1600-cell windows) at 4096 x 1024.
#ifndef SP_PARTICLE_WAVES_PER_EU
#define SP_PARTICLE_WAVES_PER_EU 5
#endif
#ifndef SP_WMAX32
#define SP_WMAX32 2000
#endif
#ifndef SP_PARTICLE_NT
#define SP_PARTICLE_NT 256
#endif
template <class T> struct ChunkShape {
  static constexpr int NT = SP_PARTICLE_NT;
  // FP32: 4 planes x 2000 cells = 32 KB, so 5 workgroups (5 waves per SIMD) fit one CU's LDS
  static constexpr int WMAX = sizeof(T) == 4 ? SP_WMAX32 : 1000;
  static constexpr int NCOPY = SP_WINDOW_COPIES;  // most private window copies
};

// floor(x + 0.5) in one instruction (v_cvt_rpi_i32_f32).  The hardware adds
// 0.5 in FP32, so for |x| >= 2^22 the result can be off by one from the
// exact round-half-up; either way |result - x| <= 1.
__device__ __forceinline__ int rint_half_up(float x) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

template <class T> struct Sink {
  lds_ptr<typename Acc<T>::type> lds;
  bool use_lds;
  int x0, y0, w, h, wh;
  int pitch;
  T scale[4];  // fixed point: 2^e per LDS plane (Jx Jy Jz dens)
  T qsj;       // charge * scale[0] (fixed point)
  glb_ptr<T> g[4];  // Jx Jy Jz and the species density (GridSet::dsp)
  __device__ __forceinline__ bool holds(int ix, int iy) const {
    return use_lds && ix - 1 >= x0 && ix + 1 < x0 + w && iy - 1 >= y0 && iy + 1 < y0 + h;
  }
  // window cell of the stencil's corner (ix - 1, iy - 1)
  // 24-bit multiply (full rate; a 32-bit one is quarter rate): |iy - 1 - y0| and
  // w stay below 2^16
  __device__ __forceinline__ int corner(int ix, int iy) const { return __mul24(iy - 1 - y0, w) + (ix - 1 - x0); }
};

// Adds w[k] * v to the 3x3 stencil around (ix, iy) of one plane
// (0..2 = Jx Jy Jz, 3 = the species density, in LDS and globally).
// The FP32 LDS windows use the separable forms below instead.
template <class T>
__device__ __forceinline__ void add9(const Sink<T>& k, int lplane, int gplane, int ix, int iy, bool in_lds,
                                     const T (&w)[9], T v) {
  if (in_lds) {
    auto a = k.lds + lplane * ChunkShape<T>::WMAX + k.corner(ix, iy);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) lds_add<typename Acc<T>::type>(a + r * k.w + c, w[3 * r + c] * v);
  } else {
    glb_ptr<T> a = k.g[gplane] + __mul24(iy - 1, k.pitch) + (ix - 1);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) glb_add<T>(a + r * k.pitch + c, w[3 * r + c] * v);
  }
}

// The TSC weights as separate x and y factors: w[3r + c] = wy[r] * wx[c] up
// to rounding (the reference forms each w as one left-to-right product
// chain, tsc above).  Deposits are compared with a tolerance, so the FP32
// windows use this cheaper form: 9 products per component instead of 18.
__device__ __forceinline__ void tsc_sep(float dx, float dy, float (&wx)[3], float (&wy)[3]) {
  const float h = 0.5f, q = 0.75f;
  const float a = h - dx, b = h + dx, c = h - dy, d = h + dy;
  wx[0] = (h * a) * a;
  wx[1] = q - dx * dx;
  wx[2] = (h * b) * b;
  wy[0] = (h * c) * c;
  wy[1] = q - dy * dy;
  wy[2] = (h * d) * d;
}

__device__ __forceinline__ void lds_addi(lds_ptr<int> p, int v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Sim::deposit_current (src/lib.rs:266-393): v = q p / gamma, J += w v on the
// 3x3 stencil around (ix, iy), ghosts included.  `inv` = 1 / psa.
template <bool Q, class T>
__device__ __forceinline__ void deposit_current(int ix, int iy, T dx, T dy, T px, T py, T pz, T inv, T q,
                                                const Geo<T>& g, const Sink<T>& k, int& bad) {
  // a stencil inside the window is inside the grid (the window lies in the
  // padded grid), so the grid test is only needed off the window; in a quiet
  // chunk every stencil is inside it
  const bool in = Q || k.holds(ix, iy);
  if constexpr (Acc<T>::fixed) {
    if (in) {  // fixed point: contributions (wy v 2^e) wx, rounded
      constexpr int W = ChunkShape<T>::WMAX;
      const T f = k.qsj * inv;
      const T vs[3] = {px * f, py * f, pz * f};
      T wx[3], wy[3];
      tsc_sep(dx, dy, wx, wy);
      lds_ptr<int> a = k.lds + k.corner(ix, iy);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        lds_ptr<int> row = a + r * k.w;
#pragma unroll
        for (int pl = 0; pl < 3; ++pl) {
          const T t = wy[r] * vs[pl];
#pragma unroll
          for (int c = 0; c < 3; ++c) lds_addi(row + pl * W + c, rint_half_up(t * wx[c]));
        }
      }
      return;
    }
  }
  if (!Q && !inside(ix, iy, g)) { bad |= F_CELL; return; }
  const T vx = q * px * inv, vy = q * py * inv, vz = q * pz * inv;
  T w[9];
  tsc(dx, dy, w);
  add9(k, 0, 0, ix, iy, in, w, vx);
  add9(k, 1, 1, ix, iy, in, w, vy);
  add9(k, 2, 2, ix, iy, in, w, vz);
}

// Sim::calc_density (src/lib.rs:395-476): rho += w q, dens += w.  One species
// per pass, so the pass accumulates only its density sum w (window and
// fallback alike, into GridSet::dsp); the launcher then adds q * dsp to rho and
// dsp to dens (fold_species_density).
template <bool Q, class T>
__device__ __forceinline__ void deposit_density(int ix, int iy, T dx, T dy, const Geo<T>& g, const Sink<T>& k,
                                                int& bad) {
  const bool in = Q || k.holds(ix, iy);
  if constexpr (Acc<T>::fixed) {
    if (in) {
      constexpr int W = ChunkShape<T>::WMAX;
      T wx[3], wy[3];
      tsc_sep(dx, dy, wx, wy);
      lds_ptr<int> a = k.lds + 3 * W + k.corner(ix, iy);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const T t = wy[r] * k.scale[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) lds_addi(a + r * k.w + c, rint_half_up(t * wx[c]));
      }
      return;
    }
  }
  if (!Q && !inside(ix, iy, g)) { bad |= F_CELL; return; }
  T w[9];
  tsc(dx, dy, w);
  add9(k, 3, 3, ix, iy, in, w, T(1));
}

// Prtl::boris_push for one particle (src/prtls.rs:306-440): ordered 9-term
// TSC gather of E and B, then the relativistic Boris rotation.
// `Fields` is a GridSet (FP32: with the LDS window `fw`, global loads outside
// it; FP64: global loads) or, for the FP32 stand-alone push, FieldBufs.
template <bool Q, class T, class Fields>
__device__ __forceinline__ void boris(int ix, int iy, T dx, T dy, T& px, T& py, T& pz, const Fields& fs,
                                      const FieldWin& fw, const Geo<T>& g, T alpha, T beta, int& bad) {
  if (!Q && !inside(ix, iy, g)) { bad |= F_CELL; return; }
  T w[9];
  tsc(dx, dy, w);
  T f[6];  // Ex Ey Ez Bx By Bz
  if constexpr (sizeof(T) == 4 && std::is_same<Fields, GridSet>::value)
    gather6w<Q, ChunkShape<float>::WMAX>(fs, fw, g, ix, iy, w, f);
  else
    gather6(fs, g, ix, iy, w, f);
  const T ext = f[0] * beta, eyt = f[1] * beta, ezt = f[2] * beta;
  T bxt = f[3] * alpha, byt = f[4] * alpha, bzt = f[5] * alpha;
  const T ux = px + ext, uy = py + eyt, uz = pz + ezt;
  const T gt = rcp_exact(sqrt_exact(T(1.) + (ux * ux + uy * uy + uz * uz) * g.csqinv));
  bxt *= gt;
  byt *= gt;
  bzt *= gt;
  const T bor = T(2.0) * rcp_exact(T(1.0) + bxt * bxt + byt * byt + bzt * bzt);
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
template <bool Q, class T>
__device__ __forceinline__ void move_and_deposit(int& ix, int& iy, T& dx, T& dy, T& px, T py, T pz, T q,
                                                 const Geo<T>& g, const Sink<T>& k, int& bad) {
  const T psa = gamma_of(px, py, pz, g.csqinv);
  // a non-finite momentum (or one whose square overflows) gives a non-finite
  // gamma: one test instead of three on the momenta
  if (g.checked && !(psa <= std::numeric_limits<T>::max())) bad |= F_FINITE;
  const T inv = rcp_exact(psa);
  const T c1 = g.half_dt * inv;
  move_half<Q>(ix, iy, dx, dy, px, py, c1, g, bad);
  deposit_current<Q>(ix, iy, dx, dy, px, py, pz, inv, q, g, k, bad);
  move_half<Q>(ix, iy, dx, dy, px, py, c1, g, bad);
  deposit_density<Q>(ix, iy, dx, dy, g, k, bad);
}

__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
  return v;
}

// The particle pass.  One workgroup owns a chunk of K consecutive particles:
//  1. the chunk's state is staged in LDS with coalesced loads, and the
//     bounding box of its cells is reduced;
//  2. particles move at most one cell per step (c dt < 1) and deposit on a
//     3x3 stencil, so every deposit of the chunk lands within 2 cells of that
//     box: the box + 2 becomes an LDS window of Jx, Jy, Jz and dens (rho of
//     one species is q * dens).  For the fixed-point FP32 windows, a window
//     cell only receives contributions from particles that start within 2
//     cells of it, each bounded by |q| c for J (|v| < c) and 1 for dens, so
//     with m the largest particle count of any cell of the chunk a cell's
//     |sum| <= 25 m |q| c (resp. 25 m); the scale 2^e keeps that below 2^30.
//     With the push (FP32), the six field grids over the box + 1 are staged
//     in LDS too (FieldWin);
//  3. each thread processes PPT particles: DIRECT (the default) takes chunk
//     slot tid + NT j, whose neighbours the tile sort put in neighbouring
//     cells; the staged variant visits slot (s * 33) mod K for s = tid + NT j,
//     particles ~33 slots apart, i.e. in different cells of a cell-sorted
//     chunk.  Either way LDS atomics rarely collide.  Optionally the Boris push
//     of the previous step runs first (PUSH);
//  4. state goes back with coalesced stores and the window is added to HBM
//     with one atomic per non-zero cell and grid.
// When the box of the whole chunk would not fit, the windows cover the
// particles near the chunk's middle particle.  Particles whose stencil leaves
// a window (periodic wrap, unsorted input) gather with global loads and
// deposit with global atomics, so the result never depends on the particle
// order; only the speed does.
template <class T, bool PUSH, bool DIRECT, int K>
__global__ void __launch_bounds__(SP_PARTICLE_NT) __attribute__((amdgpu_waves_per_eu(SP_PARTICLE_WAVES_PER_EU)))
k_particles(SpeciesDev<T> s, Geo<T> g, GridSet gs, int* flags) {
  constexpr int NT = ChunkShape<T>::NT, WMAX = ChunkShape<T>::WMAX, PPT = K / NT;
  static_assert((K & (K - 1)) == 0 && K >= NT, "the chunk is a power of two of at least NT particles");
  constexpr int NCOPY = ChunkShape<T>::NCOPY;
  constexpr int KS = DIRECT ? 1 : K;  // staging arrays (unused when DIRECT)
  using A = typename Acc<T>::type;
  __shared__ int32_t s_cell[KS];
  __shared__ T s_dx[KS], s_dy[KS], s_px[KS], s_py[KS], s_pz[KS];
  __shared__ A s_acc[4 * WMAX];
  __shared__ int s_box[5];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * K;
  const int cnt = (int)min((int64_t)K, s.n - base);
  if (tid == 0) {
    s_box[0] = INT_MAX;
    s_box[1] = INT_MIN;
    s_box[2] = INT_MAX;
    s_box[3] = INT_MIN;
    s_box[4] = 0;
  }
  __syncthreads();
  // Particle i = tid + NT j of the chunk; DIRECT keeps its cell in a register.
  // The box is reduced on the packed cells (ix, iy) as 16-bit halves with
  // packed min/max (grids are at most 8192 cells a side; a corrupt cell with a
  // half >= 2^15 reads as negative, which only widens the box and keeps the
  // chunk off the quiet path).
  int32_t creg[DIRECT ? PPT : 1];
  typedef short pk16 __attribute__((ext_vector_type(2)));
  pk16 bmn = {SHRT_MAX, SHRT_MAX}, bmx = {SHRT_MIN, SHRT_MIN};
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int i = tid + NT * j;
    if (i >= cnt) break;
    const int32_t c = s.cell[base + i];
    if constexpr (DIRECT) {
      creg[j] = c;
    } else {
      s_cell[i] = c;
      s_dx[i] = s.dx[base + i];
      s_dy[i] = s.dy[base + i];
      s_px[i] = s.px[base + i];
      s_py[i] = s.py[base + i];
      s_pz[i] = s.pz[base + i];
    }
    bmn = __builtin_elementwise_min(bmn, __builtin_bit_cast(pk16, c));
    bmx = __builtin_elementwise_max(bmx, __builtin_bit_cast(pk16, c));
  }
  const auto cell_of = [&](int j) -> int32_t {
    if constexpr (DIRECT) return creg[j];
    else return s_cell[tid + NT * j];
  };
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    bmn = __builtin_elementwise_min(bmn, __builtin_bit_cast(pk16, __shfl_xor(__builtin_bit_cast(int, bmn), off)));
    bmx = __builtin_elementwise_max(bmx, __builtin_bit_cast(pk16, __shfl_xor(__builtin_bit_cast(int, bmx), off)));
  }
  if ((tid & 63) == 0) {
    atomicMin(&s_box[0], (int)bmn.x);
    atomicMax(&s_box[1], (int)bmx.x);
    atomicMin(&s_box[2], (int)bmn.y);
    atomicMax(&s_box[3], (int)bmx.y);
  }
  __syncthreads();
  Sink<T> k;
  k.pitch = g.pitch;
  k.g[0] = (glb_ptr<T>)gs.g[SP_JX];
  k.g[1] = (glb_ptr<T>)gs.g[SP_JY];
  k.g[2] = (glb_ptr<T>)gs.g[SP_JZ];
  k.g[3] = (glb_ptr<T>)gs.dsp;
  k.x0 = max(s_box[0] - 2, 0);
  k.y0 = max(s_box[2] - 2, 0);
  k.w = min(s_box[1] + 2, g.nx + 1) - k.x0 + 1;
  k.h = min(s_box[3] + 2, g.ny + 1) - k.y0 + 1;
  // the windows cover every particle of the chunk, and no cell was clipped
  // (so the unsigned ix of every cell lies in the box as well)
  const bool whole_box = k.w * k.h <= WMAX && s_box[0] >= 0 && s_box[1] <= g.nx + 1 && s_box[2] >= 0 &&
                         s_box[3] <= g.ny + 1;
  if (g.lds_ok && cnt > 0 && !whole_box) {
    // A few particles far from the rest (crossed the periodic y boundary since
    // the last sort, or a chunk that spans the end of the tile order): box
    // only the particles within RBOX cells of the chunk's middle particle;
    // the others miss the window and deposit with global atomics.
    constexpr int RBOX = 16;
    const int32_t mc = DIRECT ? s.cell[base + cnt / 2] : s_cell[DIRECT ? 0 : cnt / 2];
    const int mx = cell_ix(mc), my = cell_iy(mc);
    __syncthreads();  // everyone has read the full box
    if (tid == 0) {
      s_box[0] = INT_MAX;
      s_box[1] = INT_MIN;
      s_box[2] = INT_MAX;
      s_box[3] = INT_MIN;
    }
    __syncthreads();
    int a0 = INT_MAX, a1 = INT_MIN, b0 = INT_MAX, b1 = INT_MIN;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      if (tid + NT * j >= cnt) break;
      const int32_t c = cell_of(j);
      const int ix = cell_ix(c), iy = cell_iy(c);
      if (abs(ix - mx) <= RBOX && abs(iy - my) <= RBOX) {
        a0 = min(a0, ix);
        a1 = max(a1, ix);
        b0 = min(b0, iy);
        b1 = max(b1, iy);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a0 = min(a0, __shfl_xor(a0, off));
      a1 = max(a1, __shfl_xor(a1, off));
      b0 = min(b0, __shfl_xor(b0, off));
      b1 = max(b1, __shfl_xor(b1, off));
    }
    if ((tid & 63) == 0) {
      atomicMin(&s_box[0], a0);
      atomicMax(&s_box[1], a1);
      atomicMin(&s_box[2], b0);
      atomicMax(&s_box[3], b1);
    }
    __syncthreads();
    k.x0 = max(s_box[0] - 2, 0);
    k.y0 = max(s_box[2] - 2, 0);
    k.w = min(s_box[1] + 2, g.nx + 1) - k.x0 + 1;
    k.h = min(s_box[3] + 2, g.ny + 1) - k.y0 + 1;
  }
  // g.lds_ok: c dt < 1, i.e. no particle moves more than one cell per step
  const bool use_lds = g.lds_ok && cnt > 0 && k.w > 0 && k.h > 0 && k.w * k.h <= WMAX;
  k.wh = use_lds ? k.w * k.h : 0;
  k.lds = (lds_ptr<A>)s_acc;
  k.use_lds = use_lds;
  if constexpr (Acc<T>::fixed) {
    if (use_lds) {  // largest particle count of a cell of this chunk
      lds_ptr<int> cntg = (lds_ptr<int>)s_acc;
      for (int i = tid; i < k.wh; i += NT) cntg[i] = 0;
      __syncthreads();
      if (whole_box) {  // every particle's cell is in the window
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
          if (tid + NT * j >= cnt) break;
          const int32_t c = cell_of(j);
          __hip_atomic_fetch_add(cntg + __mul24(cell_iy(c) - k.y0, k.w) + (cell_ix(c) - k.x0), 1, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      } else {
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
          if (tid + NT * j >= cnt) break;
          const int32_t c = cell_of(j);
          const int cx = cell_ix(c) - k.x0, cy = cell_iy(c) - k.y0;
          if (cx >= 0 && cx < k.w && cy >= 0 && cy < k.h)  // particles outside deposit globally
            __hip_atomic_fetch_add(cntg + __mul24(cy, k.w) + cx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      __syncthreads();
      int m = 0;
      for (int i = tid; i < k.wh; i += NT) m = max(m, (int)cntg[i]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
      if ((tid & 63) == 0) atomicMax(&s_box[4], m);
      __syncthreads();
    }
    const float m25 = 25.0f * (float)max(s_box[4], 1);
    const float bj = m25 * fabsf((float)s.charge) * (float)g.c;
    // 2^e with e = 30 - ceil(log2(bound)): sums stay below 2^30
    k.scale[0] = k.scale[1] = k.scale[2] = ldexpf(1.0f, 30 - ilogbf(bj) - 1);
    k.scale[3] = ldexpf(1.0f, 30 - ilogbf(m25) - 1);
    k.qsj = (float)s.charge * k.scale[0];
  }
  // Private copies of the window: lane l of each wave deposits into copy
  // l mod nc, so lanes whose particles share a cell mostly hit different
  // addresses (same-address LDS atomics serialise).  Small windows (dense
  // chunks) get more copies; the copy stride is odd so that one cell of
  // consecutive copies falls into different banks.
  int cs = k.wh | 1, nc = 1;
  FieldWin fw;
  fw.p = (lds_ptr<int>)s_acc;
  fw.use = false;
  if constexpr (PUSH && sizeof(T) == 4) {
    // the field window: the box + 1 clipped to the padded grid
    fw.x0 = max(s_box[0] - 1, 0);
    fw.y0 = max(s_box[2] - 1, 0);
    fw.w = min(s_box[1] + 1, g.nx + 1) - fw.x0 + 1;
    fw.h = min(s_box[3] + 1, g.ny + 1) - fw.y0 + 1;
    fw.use = use_lds && fw.w > 0 && fw.h > 0 && fw.w * fw.h <= SP_FWIN && cs <= WMAX - 2 * SP_FWIN;
  }
  const int room = fw.use ? WMAX - 2 * SP_FWIN : WMAX;
  if (cs <= room) nc = min(NCOPY, room / cs);
  else cs = k.wh;
  k.lds = (lds_ptr<A>)s_acc + ((tid & 63) % nc) * cs;
  for (int i = tid; i < nc * cs; i += NT) {
#pragma unroll
    for (int pl = 0; pl < 4; ++pl) s_acc[pl * WMAX + i] = A(0);
  }
  if constexpr (PUSH && sizeof(T) == 4) {
    if (fw.use) {
      const int n = fw.w * fw.h;
      const float rw = 1.0f / (float)fw.w;
      for (int i = tid; i < n; i += NT) {
        const int r = row_of(i, rw), c = i - r * fw.w;
        const int o = __mul24(fw.y0 + r, g.pitch) + fw.x0 + c;
#pragma unroll
        for (int f = 0; f < 6; ++f)
          fw.p[fwin_off<WMAX>(f) + i] = __builtin_bit_cast(int, ((const float*)gs.g[f])[o]);
      }
    }
  }
  __syncthreads();

  int bad = 0;
  // A quiet chunk: every particle's cell stays in the grid interior this step
  // (one crossing at most, c dt < 1) and away from the wall, the left edge
  // and, for periodic x, the x ends; the box is whole and both windows are in
  // LDS unclipped.  Then no boundary or assert can act, every stencil lies in
  // the windows, and the particles take a path without those tests (the
  // arithmetic is the same).  The precision study's absolute-FP32 positions
  // always take the general path.
  bool quiet = use_lds && whole_box && s_box[2] >= 2 && s_box[3] <= g.ny - 1 && !g.abs_pos32;
  if constexpr (PUSH && sizeof(T) == 4) quiet = quiet && fw.use;
  {
    int xlo = 2, xhi = g.nx - 1;
    if (!g.periodic_x) {
      xhi = min(xhi, g.wall - 2);
      if (g.left_edge == SP_LEFT_REFLECT) xlo = max(xlo, g.lwall + 2);
      else if (g.checked) xlo = max(xlo, g.delta + 2);
    }
    quiet = quiet && s_box[0] >= xlo && s_box[1] <= xhi;
  }
  // slow-path particles, for the adaptive sort schedule (one atomic per
  // chunk that is not quiet)
  if (tid == 0 && !quiet && cnt > 0)
    atomicAdd((unsigned long long*)(flags + SLOW_COUNTER), (unsigned long long)cnt);
  const auto particle = [&](auto q, int j) {
    constexpr bool Q = decltype(q)::value;
    if constexpr (DIRECT) {
      // the sort interleaves each tile's particles (k_tile_pass), so
      // consecutive particles, hence lanes, sit in different cells
      // the chunk's arrays from uniform bases, 32-bit lane offsets
      const unsigned i = (unsigned)(tid + NT * j);
      if ((int)i < cnt) {
        int32_t* const pc = s.cell + base;
        T *const pdx = s.dx + base, *const pdy = s.dy + base, *const ppx = s.px + base;
        T *const ppy = s.py + base, *const ppz = s.pz + base;
        const int32_t c = creg[j];
        int ix = cell_ix(c), iy = cell_iy(c);
        T dx = pdx[i], dy = pdy[i], px = ppx[i], py = ppy[i], pz = ppz[i];
        if (PUSH) {
          boris<Q>(ix, iy, dx, dy, px, py, pz, gs, fw, g, s.alpha, s.beta, bad);
          ppy[i] = py;
          ppz[i] = pz;
        }
        move_and_deposit<Q>(ix, iy, dx, dy, px, py, pz, s.charge, g, k, bad);
        pc[i] = make_cell(ix, iy);
        pdx[i] = dx;
        pdy[i] = dy;
        ppx[i] = px;
      }
    } else {
      const int p = ((tid + NT * j) * 33) & (K - 1);
      if (p < cnt) {
        const int32_t c = s_cell[p];
        int ix = cell_ix(c), iy = cell_iy(c);
        T dx = s_dx[p], dy = s_dy[p], px = s_px[p], py = s_py[p], pz = s_pz[p];
        if (PUSH) {
          boris<Q>(ix, iy, dx, dy, px, py, pz, gs, fw, g, s.alpha, s.beta, bad);
          s_py[p] = py;
          s_pz[p] = pz;
        }
        move_and_deposit<Q>(ix, iy, dx, dy, px, py, pz, s.charge, g, k, bad);
        s_cell[p] = make_cell(ix, iy);
        s_dx[p] = dx;
        s_dy[p] = dy;
        s_px[p] = px;
      }
    }
  };
  if (quiet) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) particle(std::true_type{}, j);
  } else {
#pragma unroll
    for (int j = 0; j < PPT; ++j) particle(std::false_type{}, j);
  }
  __syncthreads();
  if constexpr (!DIRECT) {
    for (int i = tid; i < cnt; i += NT) {
      s.cell[base + i] = s_cell[i];
      s.dx[base + i] = s_dx[i];
      s.dy[base + i] = s_dy[i];
      s.px[base + i] = s_px[i];
      if (PUSH) {
        s.py[base + i] = s_py[i];
        s.pz[base + i] = s_pz[i];
      }
    }
  }
  const float rk = 1.0f / (float)max(k.w, 1);
  for (int i = tid; i < k.wh; i += NT) {
    const int r = row_of(i, rk), cc = i - r * k.w;
    const int64_t o = (int64_t)(k.y0 + r) * g.pitch + k.x0 + cc;
#pragma unroll
    for (int pl = 0; pl < 4; ++pl) {
      A a = s_acc[pl * WMAX + i];
      for (int c = 1; c < nc; ++c) a += s_acc[pl * WMAX + c * cs + i];
      if (a == A(0)) continue;
      T v;
      if constexpr (Acc<T>::fixed) v = (T)a * (T(1) / k.scale[pl]);
      else v = a;
      glb_add<T>(k.g[pl] + o, v);
    }
  }
  bad = wave_or(bad);
  if (bad && (tid & 63) == 0) atomicOr(flags, bad);
}

template <class T>
__global__ void __launch_bounds__(256) k_push(SpeciesDev<T> s, Geo<T> g, GridSet gs, int* flags) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int bad = 0;
  // FP32: the raw-buffer gather of the particle pass (one address register
  // per particle, stencil columns as immediates)
  const auto fields = [&] {
    if constexpr (sizeof(T) == 4) return field_bufs(gs, g);
    else return gs;
  }();
  if (i < s.n) {
    const int32_t c = s.cell[i];
    T px = s.px[i], py = s.py[i], pz = s.pz[i];
    FieldWin none{};
    none.use = false;
    boris<false>(cell_ix(c), cell_iy(c), s.dx[i], s.dy[i], px, py, pz, fields, none, g, s.alpha, s.beta, bad);
    if (g.checked && !finite3(px, py, pz)) bad |= F_FINITE;
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

// ------------------------------------------------------------- tile sort
// Counting sort of a species by tile (TILE_W x TILE_H cells, tile rows in
// serpentine order: even rows left to right, odd rows right to left, so
// consecutive tiles are always neighbours); inside a tile, each workgroup's
// share is interleaved over the tile's cells (k_tile_pass), ties within a cell
// in the order the atomics give.  The key of a particle is a pure function of
// its cell, so the keys are bit-exact; only the order within a cell is not
// reproducible.
__device__ __forceinline__ int tile_of(int32_t c, int ntx, int nty, int twl, int thl) {
  const int tx = min(max((cell_ix(c) - 1) >> twl, 0), ntx - 1);
  const int ty = min(max((cell_iy(c) - 1) >> thl, 0), nty - 1);
  return ty * ntx + ((ty & 1) ? ntx - 1 - tx : tx);
}

// Each workgroup takes SORT_PPT * 256 consecutive particles and counts them
// per tile in an LDS hash table (tile -> count), one LDS atomic per group of
// lanes of a wave that share a tile; then one global atomic per distinct tile
// of the workgroup.  The input is sorted by the previous pass except for the
// particles that changed tile since, so a workgroup holds few distinct tiles
// and the global counters see ~1-3 atomics per workgroup instead of one per
// wave and tile.
#ifndef SP_SORT_HT_LOG
#define SP_SORT_HT_LOG 11
#endif
#ifndef SP_SORT_PPT
#define SP_SORT_PPT 8
#endif
// table: 2^SP_SORT_HT_LOG >= SORT_NT * SORT_PPT slots, so every key of a
// workgroup finds one; the smallest such table leaves LDS for 4 workgroups
// per CU (4096 slots: 2; measured 1.3x faster placement, 2x faster counting
// at 4096 x 1024)
constexpr int SORT_NT = 256, SORT_PPT = SP_SORT_PPT, SORT_HT = 1 << SP_SORT_HT_LOG;
static_assert(SORT_HT >= SORT_NT * SORT_PPT, "one slot per particle of a workgroup at least");

__device__ __forceinline__ int hash_slot(lds_ptr<int> keys, int key) {
  int h = (int)(((unsigned)key * 2654435761u) >> (32 - SP_SORT_HT_LOG));
  for (;;) {
    int prev = -1;  // holds the slot's old key afterwards
    __hip_atomic_compare_exchange_strong(keys + h, &prev, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    if (prev == -1 || prev == key) return h;
    h = (h + 1) & (SORT_HT - 1);
  }
}

// Adds the lanes' keys to the table; returns each active lane's slot and its
// rank among the workgroup's particles of that tile.
__device__ __forceinline__ void table_add(lds_ptr<int> keys, lds_ptr<int> cnt, int key, bool active, int& slot,
                                          int& rank) {
  const int lane = __lane_id();
  uint64_t todo = __ballot(active);
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int lkey = __shfl(key, leader);
    const uint64_t grp = __ballot(active && key == lkey) & todo;
    int s = 0, b = 0;
    if (lane == leader) {
      s = hash_slot(keys, lkey);
      b = __hip_atomic_fetch_add(cnt + s, __popcll(grp), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    s = __shfl(s, leader);
    b = __shfl(b, leader);
    if ((grp >> lane) & 1ull) {
      slot = s;
      rank = b + __popcll(grp & ((1ull << lane) - 1ull));
    }
    todo &= ~grp;
  }
}

// x mod n for 0 <= x < 2^23 and 1 <= n <= 2^12, through FP32: x is exact as
// a float, and the quotient from the rounded reciprocal is off by at most one
// (|x rn - x / n| < x / n 2^-22 < 1), which the two corrections remove.  About
// a tenth of the instructions of the integer division it replaces.
__device__ __forceinline__ int mod_small(int x, int n) {
  const float rn = __builtin_amdgcn_rcpf((float)n);
  int r = x - (int)((float)x * rn) * n;
  r += r < 0 ? n : 0;
  r -= r >= n ? n : 0;
  return r;
}

// Inverse of a modulo n (gcd(a, n) = 1), extended Euclid.
__device__ __forceinline__ int inv_mod(int a, int n) {
  int t = 0, nt = 1, r = n, nr = a % n;
  while (nr != 0) {
    const int q = r / nr;
    int x = t - q * nt;
    t = nt;
    nt = x;
    x = r - q * nr;
    r = nr;
    nr = x;
  }
  return t < 0 ? t + n : t;
}
__device__ __forceinline__ int gcd_int(int a, int b) {
  while (b != 0) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// Counts (HIST) or places (!HIST) one workgroup's particles.  Placement
// interleaves each group of n particles that share a tile: the group is ranked
// by cell (row-major inside the tile), and rank r goes to position
// (r * u^-1) mod n for a stride u ~ n / (cells per tile) coprime with n, so
// position p holds rank (p * u) mod n: consecutive positions, hence the lanes
// of a wave of the particle pass, hold particles in consecutive cells of the
// tile (round robin over its cells).  Lanes then rarely hit the same deposit
// address, and neighbouring lanes read neighbouring field values.
template <class T, bool HIST>
__global__ void __launch_bounds__(SORT_NT) __attribute__((amdgpu_waves_per_eu(5)))
k_tile_pass(SpeciesDev<T> a, SpeciesDev<T> b, int ntx, int nty, int twl, int thl, int* ctr) {
  constexpr int CH = SORT_NT * SORT_PPT;
  constexpr int OUT = HIST ? 2 : CH;
  static_assert(CH <= SORT_HT, "the staging pairs reuse the hash table");
  // hash keys | counts; once the output positions are known, the same LDS
  // stages pairs of 32-bit values by local index
  __shared__ __attribute__((aligned(16))) int s_tab[2 * SORT_HT];
  __shared__ int s_n[HIST ? 1 : SORT_HT];
  // placement: output position by local index.  Until the positions are
  // written, its last two words count the workgroup's particles and groups
  // (separate words would take the LDS 8 bytes past 32 KB, and with it the
  // workgroups per CU from 5 to 4)
  __shared__ int s_gpos[OUT];
  int* const s_total = s_gpos + OUT - 2;
  int* const s_groups = s_gpos + OUT - 1;
  lds_ptr<int> keys = (lds_ptr<int>)s_tab, cnt = (lds_ptr<int>)(s_tab + SORT_HT);
  for (int i = threadIdx.x; i < SORT_HT; i += SORT_NT) {
    keys[i] = -1;
    cnt[i] = 0;
  }
  if (threadIdx.x == 0) {
    *s_total = 0;
    *s_groups = 0;
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * CH;
  // the cells first (all in flight together); placement loads the other
  // arrays only once the output positions are known, so that they do not
  // hold registers through the ranking (with them loaded here the kernel
  // took 138 VGPRs, 3 workgroups per CU)
  int32_t c[SORT_PPT];
  // the workgroup's part of each array from a uniform base, per-lane 32-bit
  // offsets (compile-time steps of SORT_NT elements)
  const int32_t* const ac = a.cell + base;
#pragma unroll
  for (int j = 0; j < SORT_PPT; ++j) {
    const int64_t i = base + threadIdx.x + SORT_NT * j;
    c[j] = i < a.n ? ac[threadIdx.x + SORT_NT * j] : 0;
  }
  int slot[SORT_PPT], rank[SORT_PPT];
#pragma unroll
  for (int j = 0; j < SORT_PPT; ++j) {
    const int64_t i = base + threadIdx.x + SORT_NT * j;
    const bool act = i < a.n;
    slot[j] = 0;
    rank[j] = 0;
    table_add(keys, cnt, act ? tile_of(c[j], ntx, nty, twl, thl) : 0, act, slot[j], rank[j]);
  }
  __syncthreads();
  for (int sl = threadIdx.x; sl < SORT_HT; sl += SORT_NT) {
    const int key = keys[sl];
    if (key < 0) continue;
    const int n = cnt[sl];
    if (HIST) {
      atomicAdd(ctr + key, n);
    } else {
      cnt[sl] = atomicAdd(ctr + key, n);  // the group's first output position
      keys[sl] = atomicAdd(s_total, n);  // the group's first local index
      const int d = atomicAdd(s_groups, 1);
      // stride: the group's mean particle count per cell of the tile (n <=
      // 2048 < 2^12, u^-1 < n, group index d < 512 whenever it is used)
      int u = max(n >> (twl + thl), 1);
      while (gcd_int(u, n) != 1) ++u;
      s_n[sl] = n | (n > 1 ? inv_mod(u % n, n) : 0) << 12 | (d & 511) << 23;
    }
  }
  if (HIST) return;
  __syncthreads();
  const int ngroups = *s_groups;
  __syncthreads();  // every thread has the count before s_gpos is rewritten
  // rank of each particle in its group's cell order (cell index inside the
  // tile, then arrival): per-group cell histograms in s_gpos, exclusive scans,
  // and the arrival rank inside the cell from the histogram's atomics.  With
  // more groups than the histograms fit (input far from sorted) the rank stays
  // the arrival order.
  const int C = 1 << (twl + thl);
  if (ngroups * C <= CH) {
    lds_ptr<int> hist = (lds_ptr<int>)s_gpos;
    for (int i = threadIdx.x; i < ngroups * C; i += SORT_NT) hist[i] = 0;
    __syncthreads();
    // histogram slot of particle j: its group's C counters, its cell's one
    // (recomputed where needed rather than held in registers)
    const auto hbin = [&](int j) {
      const int cl = ((cell_ix(c[j]) - 1) & ((1 << twl) - 1)) | ((cell_iy(c[j]) - 1) & ((1 << thl) - 1)) << twl;
      return (int)((unsigned)s_n[slot[j]] >> 23) * C + cl;
    };
#pragma unroll
    for (int j = 0; j < SORT_PPT; ++j) {
      const int64_t i = base + threadIdx.x + SORT_NT * j;
      if (i >= a.n) continue;
      rank[j] = __hip_atomic_fetch_add(hist + hbin(j), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < ngroups; d += SORT_NT) {
      int sum = 0;
      for (int q = 0; q < C; ++q) {
        const int t = hist[d * C + q];
        hist[d * C + q] = sum;
        sum += t;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SORT_PPT; ++j) {
      const int64_t i = base + threadIdx.x + SORT_NT * j;
      if (i >= a.n) continue;
      rank[j] += hist[hbin(j)];
    }
    __syncthreads();  // s_gpos is rewritten below
  }
  // local index of each particle: groups contiguous, interleaved inside
#pragma unroll
  for (int j = 0; j < SORT_PPT; ++j) {
    const int64_t i = base + threadIdx.x + SORT_NT * j;
    if (i >= a.n) continue;
    const int nu = s_n[slot[j]], n = nu & 0xFFF, inv = (nu >> 12) & 0x7FF;
    const int r = n > 1 ? mod_small(rank[j] * inv, n) : 0;
    const int loc = keys[slot[j]] + r;
    s_gpos[loc] = cnt[slot[j]] + r;
    rank[j] = loc;  // from here on: the local index
  }
  // The value arrays go through LDS two at a time, so that the global writes
  // follow the output order (contiguous within a group): (cell, tag), then for
  // FP32 (dx, dy), (px, py), (pz, -).  Each pair's loads are issued one round
  // ahead, so that they are in flight while the previous pair is staged and
  // only two pairs hold registers (all seven arrays loaded at once, or at the
  // start, took 116-138 VGPRs: 3-4 workgroups per CU instead of 5).
  const int m = (int)min((int64_t)CH, a.n - base);
  const unsigned nl = (unsigned)min((int64_t)CH, a.n - base);  // this workgroup's particles
  int2* s_pair = (int2*)s_tab;
  const auto load32 = [&](const void* arr, int esz, int (&x)[SORT_PPT]) {
#pragma unroll
    for (int j = 0; j < SORT_PPT; ++j) {
      const unsigned o = threadIdx.x + SORT_NT * j;
      if (o < nl) x[j] = ((const int*)((const char*)arr + (int64_t)base * esz))[o];
    }
  };
  const auto stage = [&](const int (&x0)[SORT_PPT], const int (&x1)[SORT_PPT]) {
    __syncthreads();  // the previous round's readers (first round: the hash table's) are done
#pragma unroll
    for (int j = 0; j < SORT_PPT; ++j)
      if (threadIdx.x + SORT_NT * j < nl) s_pair[rank[j]] = int2{x0[j], x1[j]};
    __syncthreads();
  };
  int p0[SORT_PPT], p1[SORT_PPT], q0[SORT_PPT], q1[SORT_PPT];
  load32(a.tag, 4, p1);
  if constexpr (sizeof(T) == 4) {
    load32(a.dx, 4, q0);
    load32(a.dy, 4, q1);
  }
  stage(c, p1);
  for (int l = threadIdx.x; l < m; l += SORT_NT) {
    const int g = s_gpos[l];
    const int2 x = s_pair[l];
    b.cell[g] = x.x;
    b.tag[g] = x.y;
  }
  if constexpr (sizeof(T) == 4) {
    load32(a.px, 4, p0);
    load32(a.py, 4, p1);
    stage(q0, q1);
    for (int l = threadIdx.x; l < m; l += SORT_NT) {
      const int g = s_gpos[l];
      const int2 x = s_pair[l];
      b.dx[g] = __builtin_bit_cast(float, x.x);
      b.dy[g] = __builtin_bit_cast(float, x.y);
    }
    load32(a.pz, 4, q0);
    stage(p0, p1);
    for (int l = threadIdx.x; l < m; l += SORT_NT) {
      const int g = s_gpos[l];
      const int2 x = s_pair[l];
      b.px[g] = __builtin_bit_cast(float, x.x);
      b.py[g] = __builtin_bit_cast(float, x.y);
    }
    stage(q0, q0);
    for (int l = threadIdx.x; l < m; l += SORT_NT) b.pz[s_gpos[l]] = __builtin_bit_cast(float, s_pair[l].x);
  } else {
    // FP64 values do not fit the 32-bit staging: written directly
    const T* const src[5] = {a.dx + base, a.dy + base, a.px + base, a.py + base, a.pz + base};
    T* const dst[5] = {b.dx, b.dy, b.px, b.py, b.pz};
#pragma unroll
    for (int j = 0; j < SORT_PPT; ++j) {
      const unsigned o = threadIdx.x + SORT_NT * j;
      if (o >= nl) continue;
      const int g = s_gpos[rank[j]];
#pragma unroll
      for (int q = 0; q < 5; ++q) dst[q][g] = src[q][o];
    }
  }
}

// Exclusive prefix sum of the tile counters into `cursor`, in blocks of
// SCAN_BLK counters: k_tile_block_sums writes the sum of each block after the
// counters, then k_tile_scan runs a workgroup per block: the block's offset is
// the sum of the block sums before it, and its 256 threads read four counters
// each (one int4), scan them and write their running sums.  (A single
// workgroup reading every counter twice ran at one CU's bandwidth: 58-101 us
// at 131,072 tiles.  Block sums added by the counting pass's atomics
// serialised on 128 addresses: the counting went from 0.17 to 1.28 ms.)
__device__ __forceinline__ int4 load4_tail(const int* p, int i, int e) {
  if (i + 3 < e) return *(const int4*)(p + i);
  return make_int4(i < e ? p[i] : 0, i + 1 < e ? p[i + 1] : 0, i + 2 < e ? p[i + 2] : 0, i + 3 < e ? p[i + 3] : 0);
}
__global__ void __launch_bounds__(256) k_tile_block_sums(int* count, int ntiles) {
  __shared__ int red[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int4 v = load4_tail(count, blockIdx.x * SCAN_BLK + 4 * t, ntiles);
  int s = v.x + v.y + v.z + v.w;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) red[w] = s;
  __syncthreads();
  if (t == 0) count[ntiles + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) k_tile_scan(const int* __restrict__ count, int* cursor, int ntiles) {
  static_assert(SCAN_BLK == 4 * 256, "one int4 per thread");
  __shared__ int red[4], wsum[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, g = blockIdx.x;
  const int* __restrict__ part = count + ntiles;
  int off = 0;
  for (int i = t; i < g; i += 256) off += part[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) off += __shfl_xor(off, o);
  if (lane == 0) red[w] = off;
  const int i = g * SCAN_BLK + 4 * t;
  const int4 v = load4_tail(count, i, ntiles);
  const int ts = v.x + v.y + v.z + v.w;
  int x = ts;  // inclusive scan over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int run = red[0] + red[1] + red[2] + red[3] + x - ts;
  for (int k = 0; k < w; ++k) run += wsum[k];
  const int4 o = make_int4(run, run + v.x, run + v.x + v.y, run + v.x + v.y + v.z);
  if (i + 3 < ntiles) {
    *(int4*)(cursor + i) = o;
  } else {
    if (i < ntiles) cursor[i] = o.x;
    if (i + 1 < ntiles) cursor[i + 1] = o.y;
    if (i + 2 < ntiles) cursor[i + 2] = o.z;
  }
}

template <class T>
void launch_tile_sort(const SpeciesDev<T>& a, const SpeciesDev<T>& b, int nx, int ny, int twl, int thl,
                      int* count, int* cursor, hipStream_t st) {
  if (a.n == 0) return;
  const int tw = 1 << twl, th = 1 << thl;
  const int ntx = (nx + tw - 1) / tw, nty = (ny + th - 1) / th;
  const int ntiles = ntx * nty, nblk = (ntiles + SCAN_BLK - 1) / SCAN_BLK;
  fill<int>(count, 0, ntiles, st);
  const unsigned nb = blocks_for(a.n, SORT_NT * SORT_PPT);
  hipLaunchKernelGGL((k_tile_pass<T, true>), dim3(nb), dim3(SORT_NT), 0, st, a, b, ntx, nty, twl, thl, count);
  hipLaunchKernelGGL(k_tile_block_sums, dim3(nblk), dim3(256), 0, st, count, ntiles);
  hipLaunchKernelGGL(k_tile_scan, dim3(nblk), dim3(256), 0, st, count, cursor, ntiles);
  hipLaunchKernelGGL((k_tile_pass<T, false>), dim3(nb), dim3(SORT_NT), 0, st, a, b, ntx, nty, twl, thl, cursor);
}

// The particle pass for the context's chunk size (FP32: 512, 1024 or 2048
// particles per workgroup; FP64: 1024).
template <class T, bool PUSH, bool DIRECT>
static void launch_particles(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags, hipStream_t st) {
  const dim3 nt(ChunkShape<T>::NT);
  if constexpr (sizeof(T) == 4) {
    if (g.chunk == 512)
      hipLaunchKernelGGL((k_particles<T, PUSH, DIRECT, 512>), dim3(blocks_for(s.n, 512)), nt, 0, st, s, g, gs, flags);
    else if (g.chunk == 1024)
      hipLaunchKernelGGL((k_particles<T, PUSH, DIRECT, 1024>), dim3(blocks_for(s.n, 1024)), nt, 0, st, s, g, gs,
                         flags);
    else
      hipLaunchKernelGGL((k_particles<T, PUSH, DIRECT, 2048>), dim3(blocks_for(s.n, 2048)), nt, 0, st, s, g, gs,
                         flags);
  } else {
    hipLaunchKernelGGL((k_particles<T, PUSH, DIRECT, 1024>), dim3(blocks_for(s.n, 1024)), nt, 0, st, s, g, gs, flags);
  }
}

// rho += q_k dsp_k, dens += dsp_k, dsp_k = 0 over the padded grid for the NS
// species k in order (ghosts included: the fold of the ring comes later).
// 16-byte accesses over the body (the grids are separate allocations, so
// 16-byte aligned), scalar ones over a tail.
template <class T, int NS> struct DspSet {
  T* d[NS];
  T q[NS];
};
template <class T, int NS>
__global__ void __launch_bounds__(256) k_fold_species_density(T* __restrict__ rho, T* __restrict__ dens,
                                                              DspSet<T, NS> ds, int64_t n) {
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec __attribute__((ext_vector_type(V)));
  const int64_t stride = (int64_t)gridDim.x * blockDim.x, t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t nv = n / V;
  for (int64_t i = t; i < nv; i += stride) {
    vec d[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) d[k] = ((const vec*)ds.d[k])[i];
    vec r = ((vec*)rho)[i], e = ((vec*)dens)[i];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
#pragma unroll
      for (int l = 0; l < V; ++l) {
        r[l] += ds.q[k] * d[k][l];
        e[l] += d[k][l];
      }
    }
    ((vec*)rho)[i] = r;
    ((vec*)dens)[i] = e;
#pragma unroll
    for (int k = 0; k < NS; ++k) ((vec*)ds.d[k])[i] = vec{};
  }
  for (int64_t i = nv * V + t; i < n; i += stride) {
    T r = rho[i], e = dens[i];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const T d = ds.d[k][i];
      r += ds.q[k] * d;
      e += d;
      ds.d[k][i] = T(0);
    }
    rho[i] = r;
    dens[i] = e;
  }
}
template <class T, int NS>
static void fold_group(const GridSet& gs, int64_t n, void* const* dsp, const double* q, hipStream_t st) {
  DspSet<T, NS> ds;
  for (int k = 0; k < NS; ++k) {
    ds.d[k] = (T*)dsp[k];
    ds.q[k] = (T)q[k];
  }
  // one 16-byte vector per thread up to 2^22 of them (4096 x 1024 cells: a
  // grid-stride loop waited for each iteration's loads)
  const int64_t blocks = std::min<int64_t>((n / (16 / (int64_t)sizeof(T)) + 255) / 256, 16384) + 1;
  hipLaunchKernelGGL((k_fold_species_density<T, NS>), dim3((unsigned)blocks), dim3(256), 0, st, (T*)gs.g[SP_RHO],
                     (T*)gs.g[SP_DENS], ds, n);
}
template <class T>
void fold_species_densities(const GridSet& gs, const Geo<T>& g, void* const* dsp, const double* q, int n,
                            hipStream_t st) {
  const int64_t cells = (int64_t)(g.ny + 2) * g.pitch;
  for (int k = 0; k < n; k += 4) {
    switch (std::min(4, n - k)) {
      case 1: fold_group<T, 1>(gs, cells, dsp + k, q + k, st); break;
      case 2: fold_group<T, 2>(gs, cells, dsp + k, q + k, st); break;
      case 3: fold_group<T, 3>(gs, cells, dsp + k, q + k, st); break;
      default: fold_group<T, 4>(gs, cells, dsp + k, q + k, st); break;
    }
  }
}
template <class T>
static void fold_species_density(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, hipStream_t st) {
  void* d[1] = {gs.dsp};
  const double q[1] = {(double)s.charge};
  fold_species_densities<T>(gs, g, d, q, 1, st);
}

template <class T>
void launch_move_deposit(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags, hipStream_t st,
                         bool fold) {
  if (s.n == 0) return;
  if (g.direct) launch_particles<T, false, true>(s, g, gs, flags, st);
  else launch_particles<T, false, false>(s, g, gs, flags, st);
  if (fold) fold_species_density<T>(s, g, gs, st);
}
template <class T>
void launch_push(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags, hipStream_t st) {
  if (s.n == 0) return;
  hipLaunchKernelGGL(k_push<T>, dim3(blocks_for(s.n, 256)), dim3(256), 0, st, s, g, gs, flags);
}
template <class T>
void launch_push_move_deposit(const SpeciesDev<T>& s, const Geo<T>& g, const GridSet& gs, int* flags,
                              hipStream_t st, bool fold) {
  if (s.n == 0) return;
  if (g.direct) launch_particles<T, true, true>(s, g, gs, flags, st);
  else launch_particles<T, true, false>(s, g, gs, flags, st);
  if (fold) fold_species_density<T>(s, g, gs, st);
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

#define INST(T)                                                                                                   \
  template void launch_move_deposit<T>(const SpeciesDev<T>&, const Geo<T>&, const GridSet&, int*, hipStream_t, bool); \
  template void launch_push<T>(const SpeciesDev<T>&, const Geo<T>&, const GridSet&, int*, hipStream_t);             \
  template void launch_push_move_deposit<T>(const SpeciesDev<T>&, const Geo<T>&, const GridSet&, int*, hipStream_t, \
                                            bool);                                                                  \
  template void fold_species_densities<T>(const GridSet&, const Geo<T>&, void* const*, const double*, int,         \
                                          hipStream_t);                                                             \
  template void launch_init_species<T>(const SpeciesDev<T>&, const Geo<T>&, double, double, uint64_t, int,          \
                                       hipStream_t);                                                                \
  template void launch_tile_sort<T>(const SpeciesDev<T>&, const SpeciesDev<T>&, int, int, int, int, int*, int*,       \
                                    hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace sp
