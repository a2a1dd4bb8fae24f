// cpu_ref.cpp -- CPU restatement of spectral_rs's per-timestep path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product path
// is checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only to check results (or as the timed CPU
// baseline); the product library never links or calls it.
//
// Every function restates the reference Rust source named beside it
// (paths relative to the reference repository root).  The arithmetic keeps the
// reference's left-to-right association; build with -ffp-contract=off and
// without -ffast-math so that float and double results are the IEEE results a
// Rust build produces (Rust never contracts a*b+c into an FMA).
// Float = float  <=>  default build;  Float = double  <=>  `--features dprec`
// (Cargo.toml:27-31, src/lib.rs:15-19).
//
// What pins each part (checked in tests/test_oracle_*.py):
//   * Field ghost copy/fold, copy_to_spectral/spatial, binomial filter,
//     wave numbers, 1D/2D FFT: the reference's own golden vectors
//     (tests/golden/reference_vectors.json, extracted from the reference test
//     modules), plus numpy and closed forms.
//   * update_position / apply_bc / deposit / calc_density / boris_push /
//     spectral Maxwell update (no reference test touches them): closed forms
//     (TSC partition of unity and centre weights, uniform-E kick, pure-B
//     rotation preserves |u|, E x B drift, charge sum = q N), an independent
//     rotation-matrix Boris in the test, vacuum light-wave dispersion of the
//     pseudo-spectral leapfrog, the Gauss-law projection, the time-centred
//     vacuum B of the half-step average (mod.rs:270-278), and the Nyquist
//     lines zeroed in E and B_wrk only (mod.rs:230-266)
//     (tests/test_oracle_spectral.py).
//   * rustfft's own rounding cannot be reproduced (no Rust toolchain); the FFT
//     here is this file's mixed-radix DIT with twiddles evaluated in double
//     and a table that keeps the exact symmetries of the unit circle, so a
//     real field has an exactly Hermitian transform (sizes 2^a 3^b) in both
//     precisions and the dprec B_wrk stays Hermitian over long runs.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <chrono>
#include <algorithm>
#include <type_traits>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace ref {

template <class T> struct Cx { T re, im; };

// src/flds/field.rs:17-23 -- padded (Ny+2)x(Nx+2) real grid + Ny x Nx spectrum.
template <class T> struct Field {
  std::vector<T> spatial;
  std::vector<Cx<T>> spectral;
};

// src/prtls.rs:8-23 -- SoA particle store (tag/track are not on the path).
template <class T> struct Species {
  std::vector<int64_t> ix, iy;
  std::vector<T> dx, dy, px, py, pz, psa;
  T charge, alpha, beta;
};

enum { EX = 0, EY, EZ, BX, BY, BZ, JX, JY, JZ, DSTY, DENS, NGRID };
enum { BXW = 11, BYW = 12, BZW = 13 };

struct Params {
  int64_t size_x, size_y, delta;
  int32_t dens, n_pass;
  double dt, c, gamma_inj;
  int32_t periodic_x, checked, left_edge, ref_quirks;
};

template <class T> struct Ctx {
  Params p;
  T dt, c, gamma_inj;
  int64_t nx, ny;
  Field<T> f[NGRID];
  Field<T> wrkspace;
  std::vector<Cx<T>> bwrk[3];
  std::vector<T> k_x, k_y, k_norm;
  std::vector<Species<T>> sp;
  std::string err;
  double t_dep = 0, t_solve = 0, t_push = 0;
  int64_t pitch() const { return nx + 2; }
};

struct Fail {
  std::string msg;
};
#define REF_ASSERT(cond, where)                                  \
  do {                                                           \
    if (!(cond)) throw Fail{std::string("assertion failed: ") +  \
                            #cond + " (" where ")"};             \
  } while (0)

// ---------------------------------------------------------------- wave numbers
// src/flds/wave_num.rs:11-42.  PI is f64 PI cast to Float (wave_num.rs:3).
template <class T> void wave_numbers(Ctx<T>& s) {
  const T PI = (T)3.14159265358979323846264338327950288;
  int64_t nx = s.nx, ny = s.ny;
  s.k_x.assign(nx * ny, (T)0);
  s.k_y.assign(nx * ny, (T)0);
  s.k_norm.assign(nx * ny, (T)0);
  for (int64_t i = 0; i < ny; ++i) {
    for (int64_t j = 0; j < nx; ++j) {
      int64_t ind = i * nx + j;
      s.k_x[ind] = (T)j;
      if (j >= nx / 2 + 1) s.k_x[ind] -= (T)nx;
      s.k_x[ind] *= (T)2.0 * PI / (T)nx;
      s.k_y[ind] = (T)i;
      if (i >= ny / 2 + 1) s.k_y[ind] -= (T)ny;
      s.k_y[ind] *= (T)2.0 * PI / (T)ny;
    }
  }
  for (int64_t i = 1; i < nx * ny; ++i)
    s.k_norm[i] = (T)1. / (s.k_x[i] * s.k_x[i] + s.k_y[i] * s.k_y[i]);
}

// ------------------------------------------------------------------- ghosts
// src/flds/field.rs:141-257 (update_ghosts): periodic copies into the ring.
template <class T> void update_ghosts(std::vector<T>& f, int64_t nx, int64_t ny) {
  const int64_t P = nx + 2;
  for (int64_t c = 1; c <= nx; ++c) f[0 * P + c] = f[ny * P + c];
  for (int64_t c = 1; c <= nx; ++c) f[(ny + 1) * P + c] = f[1 * P + c];
  for (int64_t r = 1; r <= ny; ++r) f[r * P + 0] = f[r * P + nx];
  for (int64_t r = 1; r <= ny; ++r) f[r * P + nx + 1] = f[r * P + 1];
  f[0] = f[ny * P + nx];
  f[nx + 1] = f[ny * P + 1];
  f[(ny + 1) * P + 0] = f[1 * P + nx];
  f[(ny + 1) * P + nx + 1] = f[1 * P + 1];
}

// src/flds/field.rs:259-373 (deposit_ghosts): add the ring into the opposite
// interior edge; the ring itself is left as is.
template <class T> void deposit_ghosts(std::vector<T>& f, int64_t nx, int64_t ny) {
  const int64_t P = nx + 2;
  for (int64_t c = 1; c <= nx; ++c) f[ny * P + c] += f[0 * P + c];
  for (int64_t c = 1; c <= nx; ++c) f[1 * P + c] += f[(ny + 1) * P + c];
  for (int64_t r = 1; r <= ny; ++r) f[r * P + nx] += f[r * P + 0];
  for (int64_t r = 1; r <= ny; ++r) f[r * P + 1] += f[r * P + nx + 1];
  f[ny * P + nx] += f[0];
  f[ny * P + 1] += f[nx + 1];
  f[1 * P + nx] += f[(ny + 1) * P + 0];
  f[1 * P + 1] += f[(ny + 1) * P + nx + 1];
}

// src/flds/field.rs:77-107
template <class T> void copy_to_spectral(Field<T>& F, int64_t nx, int64_t ny) {
  for (int64_t iy = 0; iy < ny; ++iy) {
    int64_t ij = iy * nx, ijg = (iy + 1) * (nx + 2);
    for (int64_t ix = 0; ix < nx; ++ix) {
      F.spectral[ij + ix].re = F.spatial[ijg + ix + 1];
      F.spectral[ij + ix].im = (T)0.0;
    }
  }
}

// src/flds/field.rs:109-139
template <class T> void copy_to_spatial(Field<T>& F, int64_t nx, int64_t ny) {
  for (int64_t iy = 0; iy < ny; ++iy) {
    int64_t ij = iy * nx, ijg = (iy + 1) * (nx + 2);
    for (int64_t ix = 0; ix < nx; ++ix) F.spatial[ijg + ix + 1] = F.spectral[ij + ix].re;
  }
  update_ghosts(F.spatial, nx, ny);
}

// src/flds/field.rs:375-426.  One pass of the 3x3 (1 2 1)/16 stencil over the
// interior, reading the input's ghost ring as it stands (the first pass of a
// step therefore sees the deposit spill left by deposit_ghosts), then a
// periodic ghost refresh of the result and a copy back.  The reference loop
// also evaluates padded row Ny+1, reading past the array; that row is
// overwritten by update_ghosts, so it is not computed here.
template <class T> void binomial_pass(std::vector<T>& in, std::vector<T>& buf, int64_t nx, int64_t ny) {
  const int64_t P = nx + 2;
#pragma omp parallel for schedule(static)
  for (int64_t r = 1; r <= ny; ++r) {
    for (int64_t c = 1; c <= nx; ++c) {
      int64_t i = r * P + c, ijm1 = i - 2 - nx, ijp1 = i + 2 + nx;
      T v = (T)0.0625 * in[ijm1 - 1];
      v += (T)0.125 * in[ijm1];
      v += (T)0.0625 * in[ijm1 + 1];
      v += (T)0.125 * in[i - 1];
      v += (T)0.25 * in[i];
      v += (T)0.125 * in[i + 1];
      v += (T)0.0625 * in[ijp1 - 1];
      v += (T)0.125 * in[ijp1];
      v += (T)0.0625 * in[ijp1 + 1];
      buf[i] = v;
    }
  }
  update_ghosts(buf, nx, ny);
  in = buf;
}

// ----------------------------------------------------------------------- FFT
// Stands in for rustfft's 1D plans (src/flds/fft_2d.rs:19-58).  Recursive
// mixed-radix decimation in time; sign -1 = forward (numpy convention),
// +1 = unnormalised inverse.  Twiddles are evaluated in double, then rounded
// to T; the butterflies run in T.
template <class T> struct Fft1 {
  int64_t n = 0;
  std::vector<int64_t> fac;
  std::vector<Cx<T>> tw;  // exp(-2 pi i k / n), k < n

  void plan(int64_t n_) {
    n = n_;
    fac.clear();
    int64_t m = n;
    while (m % 4 == 0) { fac.push_back(4); m /= 4; }
    for (int64_t p = 2; m > 1;) {
      if (m % p == 0) { fac.push_back(p); m /= p; }
      else ++p;
    }
    tw.resize(n);
    const double twopi = 6.283185307179586476925286766559;
    // The table obeys the exact symmetries of the unit circle: only the first
    // octant (n % 4 == 0), quadrant (n % 2 == 0) or half (n odd) is evaluated
    // and the rest is built from it by sign changes and swaps, e.g.
    // tw[n-k] = conj(tw[k]), tw[k+n/2] = -tw[k], tw[k+n/4] = -i tw[k].  With
    // separate cos/sin calls for every k a double table breaks these in the
    // last bit (-2 pi k/n and -2 pi (n-k)/n round differently), the transform
    // of a real field is then not exactly Hermitian, and the persistent B_wrk
    // (src/flds/mod.rs:211-225) amplifies the anti-Hermitian part by up to
    // (c dt |k|)^2 - 1 per step.  A float table rounded from double already
    // had the symmetries (the double's last-bit differences round away).
    auto eval = [&](int64_t k) {
      double a = -twopi * (double)k / (double)n;
      tw[k].re = (T)std::cos(a);
      tw[k].im = (T)std::sin(a);
    };
    tw[0].re = (T)1;
    tw[0].im = (T)0;
    if (n % 4 == 0) {
      const int64_t q = n / 4;
      for (int64_t k = 1; 2 * k <= q; ++k) eval(k);
      if (q % 2 == 0 && q > 0) tw[q / 2].im = -tw[q / 2].re;  // theta = pi/4: cos = sin
      for (int64_t k = q / 2 + 1; k <= q; ++k) {  // theta -> pi/2 - theta
        tw[k].re = -tw[q - k].im;
        tw[k].im = -tw[q - k].re;
      }
      for (int64_t k = q; k < 2 * q; ++k) {  // times -i
        tw[k].re = tw[k - q].im;
        tw[k].im = -tw[k - q].re;
      }
    } else if (n % 2 == 0) {
      const int64_t h = n / 2;
      for (int64_t k = 1; 2 * k <= h; ++k) eval(k);
      for (int64_t k = h / 2 + 1; k < h; ++k) {  // theta -> pi - theta
        tw[k].re = -tw[h - k].re;
        tw[k].im = tw[h - k].im;
      }
    } else {
      for (int64_t k = 1; 2 * k < n; ++k) eval(k);
      for (int64_t k = 1; 2 * k < n; ++k) {
        tw[n - k].re = tw[k].re;
        tw[n - k].im = -tw[k].im;
      }
    }
    if (n % 2 == 0) {
      const int64_t h = n / 2;
      tw[h].re = (T)-1;
      tw[h].im = (T)0;
      for (int64_t k = h + 1; k < n; ++k) {  // half turn
        tw[k].re = -tw[k - h].re;
        tw[k].im = -tw[k - h].im;
      }
    }
  }
  Cx<T> w(int64_t e, int sign) const {  // exp(sign 2 pi i e / n)
    Cx<T> t = tw[e % n];
    if (sign > 0) t.im = -t.im;
    return t;
  }
  void rec(const Cx<T>* x, int64_t xs, Cx<T>* y, int64_t len, size_t lev, int sign,
           std::vector<Cx<T>>& tmp) const {
    if (len == 1) { y[0] = x[0]; return; }
    int64_t p = fac[lev], m = len / p, step = n / len;
    for (int64_t r = 0; r < p; ++r) rec(x + r * xs, xs * p, y + r * m, m, lev + 1, sign, tmp);
    // Butterfly inputs and the p x p roots w((r q mod p) n/p): on the stack
    // for the usual small radices.  The values are the table's own entries,
    // read without the index modulo of w(); the arithmetic is unchanged.
    Cx<T> small[16], rsmall[256];
    std::vector<Cx<T>> big(p > 16 ? p : 0), rbig(p > 16 ? p * p : 0);
    Cx<T>* buf = p > 16 ? big.data() : small;
    Cx<T>* root = p > 16 ? rbig.data() : rsmall;
    for (int64_t q = 0; q < p; ++q)
      for (int64_t r = 0; r < p; ++r) root[q * p + r] = w(r * q * m * step, sign);
    for (int64_t k = 0; k < m; ++k) {
      for (int64_t r = 0; r < p; ++r) {
        // r k step < p m step = n: a direct table index
        Cx<T> a = y[r * m + k], t = tw[r * k * step];
        if (sign > 0) t.im = -t.im;
        buf[r].re = a.re * t.re - a.im * t.im;
        buf[r].im = a.re * t.im + a.im * t.re;
      }
      for (int64_t q = 0; q < p; ++q) {
        T sr = 0, si = 0;
        const Cx<T>* rq = root + q * p;
        for (int64_t r = 0; r < p; ++r) {
          Cx<T> a = buf[r], t = rq[r];
          sr += a.re * t.re - a.im * t.im;
          si += a.re * t.im + a.im * t.re;
        }
        tmp[q * m + k].re = sr;
        tmp[q * m + k].im = si;
      }
    }
    for (int64_t i = 0; i < len; ++i) y[i] = tmp[i];
  }
  void run(const Cx<T>* in, Cx<T>* out, int sign) const {
    std::vector<Cx<T>> tmp(n);
    rec(in, 1, out, n, 0, sign, tmp);
  }
};

// src/flds/fft_2d.rs:60-89
template <class T> void transpose(const std::vector<Cx<T>>& in, std::vector<Cx<T>>& out, int64_t sx, int64_t sy) {
  for (int64_t i = 0; i < sy; ++i)
    for (int64_t j = 0; j < sx; ++j) out[j * sy + i] = in[i * sx + j];
}

// src/flds/fft_2d.rs:91-172: rows along x -> transpose -> rows along y ->
// transpose; the inverse multiplies by (Nx*Ny)^-1 in Float at the end.
template <class T> void fft2d(std::vector<Cx<T>>& a, int64_t nx, int64_t ny, bool inverse) {
  Fft1<T> fx, fy;
  fx.plan(nx);
  fy.plan(ny);
  int sign = inverse ? +1 : -1;
  std::vector<Cx<T>> wrk(nx * ny);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < ny; ++r) fx.run(&a[r * nx], &wrk[r * nx], sign);
  transpose(wrk, a, nx, ny);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nx; ++r) fy.run(&a[r * ny], &wrk[r * ny], sign);
  transpose(wrk, a, ny, nx);
  if (inverse) {
    T norm = (T)1 / (T)(nx * ny);
    for (auto& v : a) { v.re *= norm; v.im *= norm; }
  }
}

// ----------------------------------------------------------------- particles
// src/prtls.rs:85-121 (update_position)
template <class T> void update_position(Ctx<T>& s, Species<T>& P) {
  T dt = s.dt;
  size_t n = P.ix.size();
  for (size_t i = 0; i < n; ++i) {
    if (s.p.checked) {
      REF_ASSERT(P.ix[i] > 0, "src/prtls.rs:100");
      REF_ASSERT(P.iy[i] > 0, "src/prtls.rs:101");
    }
    T c1 = (T)0.5 * dt * ((T)1 / P.psa[i]);
    P.dx[i] += c1 * P.px[i];
    if (P.dx[i] >= (T)0.5) { P.dx[i] -= (T)1.0; P.ix[i] += 1; }
    else if (P.dx[i] < (T)-0.5) { P.dx[i] += (T)1.0; P.ix[i] -= 1; }
    P.dy[i] += c1 * P.py[i];
    if (P.dy[i] >= (T)0.5) { P.dy[i] -= (T)1.0; P.iy[i] += 1; }
    else if (P.dy[i] < (T)-0.5) { P.dy[i] += (T)1.0; P.iy[i] -= 1; }
  }
}

template <class T> static inline T rust_signum(T x) {
  // f32::signum / f64::signum: 1.0 for +0.0 and positive, -1.0 for -0.0 and
  // negative, NaN for NaN.
  if (std::isnan(x)) return x;
  return std::signbit(x) ? (T)-1.0 : (T)1.0;
}

// src/prtls.rs:123-190 (apply_bc).  left_edge = 1 adds a mirror of
// the right wall at cell delta+1 (the reference has no left boundary and
// asserts instead, prtls.rs:163-170).
template <class T> void apply_bc(Ctx<T>& s, Species<T>& P) {
  size_t n = P.ix.size();
  int64_t ny = s.ny, nx = s.nx;
  for (size_t i = 0; i < n; ++i) {
    if (P.iy[i] < 1) P.iy[i] += ny;
    else if (P.iy[i] > ny) P.iy[i] -= ny;
  }
  if (s.p.periodic_x) {
    for (size_t i = 0; i < n; ++i) {
      if (P.ix[i] < 1) P.ix[i] += nx;
      else if (P.ix[i] > nx) P.ix[i] -= nx;
    }
    return;
  }
  int64_t wall = nx + 1 - s.p.delta;
  int64_t lwall = s.p.delta + 1;
  if (s.p.checked && n > 0) {
    int64_t mx = P.ix[0], mn = P.ix[0];
    for (size_t i = 1; i < n; ++i) { mx = std::max(mx, P.ix[i]); mn = std::min(mn, P.ix[i]); }
    REF_ASSERT(mx <= wall + 1, "src/prtls.rs:155-162");
    if (s.p.left_edge == 0) REF_ASSERT(mn > s.p.delta, "src/prtls.rs:163-170");
  }
  for (size_t i = 0; i < n; ++i) {
    if (P.ix[i] < wall) {
    } else if (P.ix[i] == wall) {
      T sgn = rust_signum(P.dx[i]);
      P.px[i] *= -sgn;
      P.dx[i] *= -sgn;
    } else {
      P.dx[i] += (T)1.0;
      P.ix[i] -= 1;
      P.px[i] *= (T)-1.0;
    }
    if (s.p.left_edge == 1) {
      if (P.ix[i] > lwall) {
      } else if (P.ix[i] == lwall) {
        T sgn = rust_signum(P.dx[i]);
        P.px[i] *= sgn;
        P.dx[i] *= sgn;
      } else {
        P.dx[i] -= (T)1.0;
        P.ix[i] += 1;
        P.px[i] *= (T)-1.0;
      }
    }
  }
}

// The nine TSC weights, src/lib.rs:339-347 (same text at lib.rs:435-443 and
// prtls.rs:321-329), evaluated left to right.
template <class T> static inline void tsc(T dx, T dy, T w[9]) {
  const T h = (T)0.5, q = (T)0.75;
  w[0] = h * (h - dy) * (h - dy) * h * (h - dx) * (h - dx);
  w[1] = h * (h - dy) * (h - dy) * (q - dx * dx);
  w[2] = h * (h - dy) * (h - dy) * h * (h + dx) * (h + dx);
  w[3] = (q - dy * dy) * h * (h - dx) * (h - dx);
  w[4] = (q - dy * dy) * (q - dx * dx);
  w[5] = (q - dy * dy) * h * (h + dx) * (h + dx);
  w[6] = h * (h + dy) * (h + dy) * h * (h - dx) * (h - dx);
  w[7] = h * (h + dy) * (h + dy) * (q - dx * dx);
  w[8] = h * (h + dy) * (h + dy) * h * (h + dx) * (h + dx);
}

// src/lib.rs:266-393 (deposit_current), serial in particle order.
template <class T> void deposit_current(Ctx<T>& s, Species<T>& P) {
  auto& jx = s.f[JX].spatial;
  auto& jy = s.f[JY].spatial;
  auto& jz = s.f[JZ].spatial;
  const int64_t W = 2 + s.nx;
  size_t n = P.ix.size();
  for (size_t i = 0; i < n; ++i) {
    int64_t ix = P.ix[i], iy = P.iy[i];
    if (s.p.checked) {
      REF_ASSERT(iy > 0, "src/lib.rs:324");
      REF_ASSERT(ix > 0, "src/lib.rs:325");
    }
    int64_t ij = iy * W, ijm1 = (iy - 1) * W, ijp1 = (iy + 1) * W;
    T psa_inv = (T)1 / P.psa[i];
    T vx = P.charge * P.px[i] * psa_inv;
    T vy = P.charge * P.py[i] * psa_inv;
    T vz = P.charge * P.pz[i] * psa_inv;
    T w[9];
    tsc(P.dx[i], P.dy[i], w);
    if (s.p.checked) REF_ASSERT(ijp1 + ix + 1 < (int64_t)jx.size(), "src/lib.rs:353");
    const int64_t at[9] = {ijm1 + ix - 1, ijm1 + ix, ijm1 + ix + 1, ij + ix - 1, ij + ix,
                           ij + ix + 1,   ijp1 + ix - 1, ijp1 + ix, ijp1 + ix + 1};
    for (int k = 0; k < 9; ++k) jx[at[k]] += w[k] * vx;
    for (int k = 0; k < 9; ++k) jy[at[k]] += w[k] * vy;
    for (int k = 0; k < 9; ++k) jz[at[k]] += w[k] * vz;
  }
}

// src/lib.rs:395-476 (calc_density)
template <class T> void calc_density(Ctx<T>& s, Species<T>& P) {
  auto& dsty = s.f[DSTY].spatial;
  auto& dens = s.f[DENS].spatial;
  const int64_t W = 2 + s.nx;
  size_t n = P.ix.size();
  for (size_t i = 0; i < n; ++i) {
    int64_t ix = P.ix[i], iy = P.iy[i];
    if (s.p.checked) {
      REF_ASSERT(iy > 0, "src/lib.rs:416");
      REF_ASSERT(ix > 0, "src/lib.rs:417");
    }
    int64_t ij = iy * W, ijm1 = (iy - 1) * W, ijp1 = (iy + 1) * W;
    T w[9];
    tsc(P.dx[i], P.dy[i], w);
    if (s.p.checked) REF_ASSERT(ijp1 + ix + 1 < (int64_t)dsty.size(), "src/lib.rs:449");
    const int64_t at[9] = {ijm1 + ix - 1, ijm1 + ix, ijm1 + ix + 1, ij + ix - 1, ij + ix,
                           ij + ix + 1,   ijp1 + ix - 1, ijp1 + ix, ijp1 + ix + 1};
    for (int k = 0; k < 9; ++k) dsty[at[k]] += w[k] * P.charge;
    for (int k = 0; k < 9; ++k) dens[at[k]] += w[k];
  }
}

// src/lib.rs:478-489
template <class T> void move_and_deposit(Ctx<T>& s, Species<T>& P) {
  update_position(s, P);
  apply_bc(s, P);
  deposit_current(s, P);
  update_position(s, P);
  apply_bc(s, P);
  calc_density(s, P);
}

// src/prtls.rs:262-443 (boris_push).  Parallel over particles like the
// reference's rayon chunks; every particle is independent.
template <class T> void boris_push(Ctx<T>& s, Species<T>& P) {
  const T csqinv = (T)1. / (s.c * s.c);
  const T beta = P.beta, alpha = P.alpha;
  const auto& ex = s.f[EX].spatial;
  const auto& ey = s.f[EY].spatial;
  const auto& ez = s.f[EZ].spatial;
  const auto& bx = s.f[BX].spatial;
  const auto& by = s.f[BY].spatial;
  const auto& bz = s.f[BZ].spatial;
  const int64_t W = 2 + s.nx;
  const int64_t n = (int64_t)P.ix.size();
  const int64_t len = (int64_t)ex.size();
  bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
  for (int64_t i = 0; i < n; ++i) {
    int64_t ix = P.ix[i], iy = P.iy[i];
    if (s.p.checked && (iy <= 0 || ix <= 0 || (iy + 1) * W + ix + 1 >= len)) { bad = true; continue; }
    int64_t ij = iy * W, ijm1 = (iy - 1) * W, ijp1 = (iy + 1) * W;
    T w[9];
    tsc(P.dx[i], P.dy[i], w);
    const int64_t at[9] = {ijm1 + ix - 1, ijm1 + ix, ijm1 + ix + 1, ij + ix - 1, ij + ix,
                           ij + ix + 1,   ijp1 + ix - 1, ijp1 + ix, ijp1 + ix + 1};
    T ext = w[0] * ex[at[0]], eyt = w[0] * ey[at[0]], ezt = w[0] * ez[at[0]];
    T bxt = w[0] * bx[at[0]], byt = w[0] * by[at[0]], bzt = w[0] * bz[at[0]];
    for (int k = 1; k < 9; ++k) ext += w[k] * ex[at[k]];
    for (int k = 1; k < 9; ++k) eyt += w[k] * ey[at[k]];
    for (int k = 1; k < 9; ++k) ezt += w[k] * ez[at[k]];
    for (int k = 1; k < 9; ++k) bxt += w[k] * bx[at[k]];
    for (int k = 1; k < 9; ++k) byt += w[k] * by[at[k]];
    for (int k = 1; k < 9; ++k) bzt += w[k] * bz[at[k]];
    ext *= beta; eyt *= beta; ezt *= beta;
    bxt *= alpha; byt *= alpha; bzt *= alpha;
    T ux = P.px[i] + ext, uy = P.py[i] + eyt, uz = P.pz[i] + ezt;
    T gt = (T)1 / std::sqrt((T)1. + (ux * ux + uy * uy + uz * uz) * csqinv);
    bxt *= gt; byt *= gt; bzt *= gt;
    T boris = (T)2.0 * ((T)1 / ((T)1.0 + bxt * bxt + byt * byt + bzt * bzt));
    T uxt = ux + uy * bzt - uz * byt;
    T uyt = uy + uz * bxt - ux * bzt;
    T uzt = uz + ux * byt - uy * bxt;
    T px = ux + boris * (uyt * bzt - uzt * byt) + ext;
    T py = uy + boris * (uzt * bxt - uxt * bzt) + eyt;
    T pz = uz + boris * (uxt * byt - uyt * bxt) + ezt;
    P.px[i] = px; P.py[i] = py; P.pz[i] = pz;
    P.psa[i] = std::sqrt((T)1.0 + (px * px + py * py + pz * pz) * csqinv);
  }
  if (bad) throw Fail{"assertion failed: iy > 0 && ix > 0 && ijp1 + ix + 1 < len (src/prtls.rs:303-335)"};
}

// ------------------------------------------------------------------- fields
// src/flds/mod.rs:66-85
template <class T> void copy_spatial_to_spectral(Ctx<T>& s) {
  const int ids[7] = {JX, JY, JZ, EX, EY, EZ, DSTY};
  for (int id : ids) copy_to_spectral(s.f[id], s.nx, s.ny);
  T norm = (T)1.0 / (T)s.p.dens;
  for (auto& v : s.f[DSTY].spectral) { v.re *= norm; v.im *= norm; }
}

template <class T> static inline Cx<T> cmul(Cx<T> a, Cx<T> b) {  // num_complex Mul
  return Cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> static inline Cx<T> rmul(Cx<T> a, T r) { return Cx<T>{a.re * r, a.im * r}; }
template <class T> static inline Cx<T> cadd(Cx<T> a, Cx<T> b) { return Cx<T>{a.re + b.re, a.im + b.im}; }
template <class T> static inline Cx<T> csub(Cx<T> a, Cx<T> b) { return Cx<T>{a.re - b.re, a.im - b.im}; }

// The spectral half of Flds::update, src/flds/mod.rs:105-278, on spectra that
// already hold the forward transforms (J, rho, E) and the persistent B_wrk.
template <class T> void spectral_update(Ctx<T>& s) {
  const int64_t N = s.nx * s.ny;
  auto& jx = s.f[JX].spectral; auto& jy = s.f[JY].spectral; auto& jz = s.f[JZ].spectral;
  auto& ex = s.f[EX].spectral; auto& ey = s.f[EY].spectral; auto& ez = s.f[EZ].spectral;
  auto& bx = s.f[BX].spectral; auto& by = s.f[BY].spectral; auto& bz = s.f[BZ].spectral;
  auto& rho = s.f[DSTY].spectral;
  const auto &kx = s.k_x, &ky = s.k_y, &kn = s.k_norm;
  // mod.rs:106-108 -- j_z at k=0 is left alone.
  jx[0] = Cx<T>{0, 0};
  jy[0] = Cx<T>{0, 0};
  // mod.rs:113-119
  bx = s.bwrk[0]; by = s.bwrk[1]; bz = s.bwrk[2];
  // mod.rs:127-173
  const T ckc = s.dt / (T)s.p.dens;
  const T cdt = s.dt * s.c;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    ex[i] = cadd(ex[i], csub(cmul(rmul(Cx<T>{(T)0.0, cdt}, ky[i]), bz[i]), rmul(jx[i], ckc)));
    ey[i] = cadd(ey[i], csub(cmul(rmul(Cx<T>{(T)0.0, -cdt}, kx[i]), bz[i]), rmul(jy[i], ckc)));
    ez[i] = cadd(ez[i], cmul(Cx<T>{(T)0.0, cdt}, csub(rmul(by[i], kx[i]), rmul(bx[i], ky[i]))));
    ez[i] = csub(ez[i], rmul(jz[i], ckc));
  }
  // mod.rs:177-207 -- Gauss-law ("Boris") correction, k=0 restored.
  Cx<T> ex0 = ex[0], ey0 = ey[0];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    Cx<T> tmp = rmul(cadd(cadd(rmul(ex[i], kx[i]), rmul(ey[i], ky[i])), cmul(Cx<T>{(T)0., (T)1.}, rho[i])), kn[i]);
    ex[i] = csub(ex[i], rmul(tmp, kx[i]));
    ey[i] = csub(ey[i], rmul(tmp, ky[i]));
  }
  ex[0] = ex0;
  ey[0] = ey0;
  // mod.rs:211-225 -- B_wrk to t+1/2 (one step past B.spectral).
  auto& wx = s.bwrk[0]; auto& wy = s.bwrk[1]; auto& wz = s.bwrk[2];
  for (int64_t i = 0; i < N; ++i) wx[i] = csub(wx[i], cmul(rmul(Cx<T>{(T)0.0, cdt}, ky[i]), ez[i]));
  for (int64_t i = 0; i < N; ++i) wy[i] = cadd(wy[i], cmul(rmul(Cx<T>{(T)0.0, cdt}, kx[i]), ez[i]));
  for (int64_t i = 0; i < N; ++i)
    wz[i] = cadd(wz[i], cmul(Cx<T>{(T)0.0, cdt}, csub(rmul(ex[i], ky[i]), rmul(ey[i], kx[i]))));
  // mod.rs:230-266 -- Nyquist column Nx/2 and row Ny/2 of B_wrk and E.
  std::vector<Cx<T>>* ny_f[6] = {&wx, &wy, &wz, &ex, &ey, &ez};
  for (auto* f : ny_f) {
    for (int64_t i = s.nx / 2; i < N; i += s.nx) (*f)[i] = Cx<T>{0, 0};
    int64_t r0 = s.ny * s.nx / 2;
    for (int64_t i = r0; i < r0 + s.nx; ++i) (*f)[i] = Cx<T>{0, 0};
  }
  // mod.rs:270-278 -- B at t+1 by averaging.
  std::vector<Cx<T>>* bs[3] = {&bx, &by, &bz};
  for (int d = 0; d < 3; ++d) {
    auto& b = *bs[d];
    const auto& w = s.bwrk[d];
    for (int64_t i = 0; i < N; ++i) b[i] = rmul(cadd(b[i], w[i]), (T)0.5);
  }
}

// src/flds/mod.rs:87-309 (Flds::update)
template <class T> void fields_update(Ctx<T>& s) {
  for (int pass = 0; pass < s.p.n_pass; ++pass) {
    const int ids[4] = {JX, JY, JZ, DSTY};
    for (int id : ids) binomial_pass(s.f[id].spatial, s.wrkspace.spatial, s.nx, s.ny);
  }
  copy_spatial_to_spectral(s);
  const int cur[4] = {JX, JY, JZ, DSTY};
  for (int id : cur) fft2d(s.f[id].spectral, s.nx, s.ny, false);
  const int e[3] = {EX, EY, EZ};
  for (int id : e) fft2d(s.f[id].spectral, s.nx, s.ny, false);
  spectral_update(s);
  const int inv[6] = {BX, BY, BZ, EX, EY, EZ};
  for (int id : inv) fft2d(s.f[id].spectral, s.nx, s.ny, true);
  for (int id : inv) copy_to_spatial(s.f[id], s.nx, s.ny);
}

// src/lib.rs:157-167
template <class T> void zero_currents(Ctx<T>& s) {
  const int ids[5] = {JX, JY, JZ, DSTY, DENS};
  for (int id : ids) std::fill(s.f[id].spatial.begin(), s.f[id].spatial.end(), (T)0);
}

template <class T> void fold_all(Ctx<T>& s) {  // src/lib.rs:134-142, 174-182
  const int ids[5] = {JX, JY, JZ, DSTY, DENS};
  for (int id : ids) deposit_ghosts(s.f[id].spatial, s.nx, s.ny);
}

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

// One iteration of the loop body, src/lib.rs:155-199.
template <class T> void step(Ctx<T>& s) {
  zero_currents(s);
  auto t0 = clk::now();
  for (auto& P : s.sp) move_and_deposit(s, P);
  fold_all(s);
  s.t_dep += secs(t0);
  t0 = clk::now();
  fields_update(s);
  s.t_solve += secs(t0);
  t0 = clk::now();
  for (auto& P : s.sp) boris_push(s, P);
  s.t_push += secs(t0);
}

// src/prtls.rs:191-221 (initialize_positions); returns the count written.
template <class T> int64_t init_positions(const Params& p, int64_t* ix, int64_t* iy, T* dx, T* dy) {
  int64_t c1 = 0;
  for (int64_t i = 0; i < p.size_y; ++i) {
    for (int64_t j = p.delta; j < p.size_x - p.delta; ++j) {
      for (int64_t k = 0; k < p.dens; ++k) {
        iy[c1 + k] = i + 1;
        ix[c1 + k] = j + 1;
        T r1 = (T)1.0 / ((T)2.0 * (T)p.dens);
        r1 += (T)k / (T)p.dens;
        dx[c1 + k] = r1 - (T)0.5;
        dy[c1 + k] = r1 - (T)0.5;
      }
      c1 += p.dens;
    }
  }
  return c1;
}

}  // namespace ref

// =================================================================== C ABI
using namespace ref;

struct RefHandle {
  int prec;  // 4 or 8
  void* ctx;
};

template <class T> static Ctx<T>* C(RefHandle* h) { return (Ctx<T>*)h->ctx; }

template <class T> static Ctx<T>* make(const Params& p) {
  auto* s = new Ctx<T>();
  s->p = p;
  s->nx = p.size_x;
  s->ny = p.size_y;
  s->dt = (T)p.dt;
  s->c = (T)p.c;
  s->gamma_inj = (T)p.gamma_inj;
  size_t ns = (size_t)(s->nx + 2) * (s->ny + 2), nk = (size_t)(s->nx * s->ny);
  for (auto& F : s->f) { F.spatial.assign(ns, (T)0); F.spectral.assign(nk, Cx<T>{0, 0}); }
  s->wrkspace.spatial.assign(ns, (T)0);
  s->wrkspace.spectral.assign(nk, Cx<T>{0, 0});
  for (auto& b : s->bwrk) b.assign(nk, Cx<T>{0, 0});
  wave_numbers(*s);
  return s;
}

#define DISPATCH(h, CALL)                                 \
  do {                                                    \
    try {                                                 \
      if ((h)->prec == 8) { auto* s = C<double>(h); CALL; } \
      else { auto* s = C<float>(h); CALL; }               \
    } catch (const Fail& f) {                             \
      ref_set_err(h, f.msg);                              \
      return -3;                                          \
    }                                                     \
  } while (0)

static void ref_set_err(RefHandle* h, const std::string& m) {
  if (h->prec == 8) C<double>(h)->err = m;
  else C<float>(h)->err = m;
}

extern "C" {

void* ref_create(int prec, int64_t size_x, int64_t size_y, int64_t delta, int32_t dens, int32_t n_pass,
                 double dt, double c, double gamma_inj, int32_t periodic_x, int32_t checked,
                 int32_t left_edge, int32_t ref_quirks) {
  Params p{size_x, size_y, delta, dens, n_pass, dt, c, gamma_inj, periodic_x, checked, left_edge, ref_quirks};
  auto* h = new RefHandle{prec == 8 ? 8 : 4, nullptr};
  if (h->prec == 8) h->ctx = make<double>(p);
  else h->ctx = make<float>(p);
  return h;
}

void ref_destroy(void* hv) {
  auto* h = (RefHandle*)hv;
  if (!h) return;
  if (h->prec == 8) delete C<double>(h);
  else delete C<float>(h);
  delete h;
}

const char* ref_last_error(void* hv) {
  auto* h = (RefHandle*)hv;
  return h->prec == 8 ? C<double>(h)->err.c_str() : C<float>(h)->err.c_str();
}

void ref_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

// Species: charge, mass as in Prtl::new (src/prtls.rs:60-62); arrays are
// copied.  psa may be NULL, in which case it is sqrt(1 + |p|^2/c^2) as in
// initialize_velocities (src/prtls.rs:248-249).
int ref_species_add(void* hv, double charge, double mass, int64_t n, const int64_t* ix, const int64_t* iy,
                    const void* dx, const void* dy, const void* px, const void* py, const void* pz,
                    const void* psa) {
  auto* h = (RefHandle*)hv;
  int id = -1;
  DISPATCH(h, {
    using T = std::remove_reference_t<decltype(s->dt)>;
    Species<T> P;
    P.charge = (T)charge;
    T m = (T)mass;
    P.beta = (P.charge / m) * (T)0.5 * s->dt;
    P.alpha = (P.charge / m) * (T)0.5 * s->dt / s->c;
    P.ix.assign(ix, ix + n);
    P.iy.assign(iy, iy + n);
    P.dx.assign((const T*)dx, (const T*)dx + n);
    P.dy.assign((const T*)dy, (const T*)dy + n);
    P.px.assign((const T*)px, (const T*)px + n);
    P.py.assign((const T*)py, (const T*)py + n);
    P.pz.assign((const T*)pz, (const T*)pz + n);
    P.psa.resize(n);
    T csqinv = (T)1. / (s->c * s->c);
    for (int64_t i = 0; i < n; ++i) {
      if (psa) P.psa[i] = ((const T*)psa)[i];
      else {
        T v = (T)1.0 + (P.px[i] * P.px[i] + P.py[i] * P.py[i] + P.pz[i] * P.pz[i]) * csqinv;
        P.psa[i] = std::sqrt(v);
      }
    }
    s->sp.push_back(std::move(P));
    id = (int)s->sp.size() - 1;
  });
  return id;
}

int ref_get_species(void* hv, int id, int64_t* ix, int64_t* iy, void* dx, void* dy, void* px, void* py,
                    void* pz, void* psa) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    using T = std::remove_reference_t<decltype(s->dt)>;
    auto& P = s->sp.at(id);
    size_t n = P.ix.size();
    std::copy(P.ix.begin(), P.ix.end(), ix);
    std::copy(P.iy.begin(), P.iy.end(), iy);
    std::memcpy(dx, P.dx.data(), n * sizeof(T));
    std::memcpy(dy, P.dy.data(), n * sizeof(T));
    std::memcpy(px, P.px.data(), n * sizeof(T));
    std::memcpy(py, P.py.data(), n * sizeof(T));
    std::memcpy(pz, P.pz.data(), n * sizeof(T));
    std::memcpy(psa, P.psa.data(), n * sizeof(T));
  });
  return 0;
}

int64_t ref_species_count(void* hv, int id) {
  auto* h = (RefHandle*)hv;
  return h->prec == 8 ? (int64_t)C<double>(h)->sp.at(id).ix.size() : (int64_t)C<float>(h)->sp.at(id).ix.size();
}

// Grid ids 0..10 = Ex Ey Ez Bx By Bz Jx Jy Jz Rho(dsty) Dens (spatial,
// padded).  Spectrum ids 0..10 = the same fields' spectra, 11..13 = B_wrk.
int ref_get_grid(void* hv, int id, void* out) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    using T = std::remove_reference_t<decltype(s->dt)>;
    auto& v = s->f[id].spatial;
    std::memcpy(out, v.data(), v.size() * sizeof(T));
  });
  return 0;
}
int ref_set_grid(void* hv, int id, const void* in) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    using T = std::remove_reference_t<decltype(s->dt)>;
    auto& v = s->f[id].spatial;
    std::memcpy(v.data(), in, v.size() * sizeof(T));
  });
  return 0;
}
int ref_get_spectrum(void* hv, int id, void* out) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    using T = std::remove_reference_t<decltype(s->dt)>;
    auto& v = id >= 11 ? s->bwrk[id - 11] : s->f[id].spectral;
    std::memcpy(out, v.data(), v.size() * 2 * sizeof(T));
  });
  return 0;
}
int ref_set_spectrum(void* hv, int id, const void* in) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    using T = std::remove_reference_t<decltype(s->dt)>;
    auto& v = id >= 11 ? s->bwrk[id - 11] : s->f[id].spectral;
    std::memcpy(v.data(), in, v.size() * 2 * sizeof(T));
  });
  return 0;
}

int ref_wave_numbers(void* hv, void* kx, void* ky, void* kn) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    using T = std::remove_reference_t<decltype(s->dt)>;
    size_t n = s->k_x.size();
    std::memcpy(kx, s->k_x.data(), n * sizeof(T));
    std::memcpy(ky, s->k_y.data(), n * sizeof(T));
    std::memcpy(kn, s->k_norm.data(), n * sizeof(T));
  });
  return 0;
}

int ref_update_position(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, update_position(*s, s->sp.at(id))); return 0; }
int ref_apply_bc(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, apply_bc(*s, s->sp.at(id))); return 0; }
int ref_deposit_current(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, deposit_current(*s, s->sp.at(id))); return 0; }
int ref_calc_density(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, calc_density(*s, s->sp.at(id))); return 0; }
int ref_move_and_deposit(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, move_and_deposit(*s, s->sp.at(id))); return 0; }
int ref_boris_push(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, boris_push(*s, s->sp.at(id))); return 0; }
int ref_deposit_ghosts(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, deposit_ghosts(s->f[id].spatial, s->nx, s->ny)); return 0; }
int ref_update_ghosts(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, update_ghosts(s->f[id].spatial, s->nx, s->ny)); return 0; }
int ref_fold_all(void* hv) { auto* h = (RefHandle*)hv; DISPATCH(h, fold_all(*s)); return 0; }
int ref_zero_currents(void* hv) { auto* h = (RefHandle*)hv; DISPATCH(h, zero_currents(*s)); return 0; }
int ref_binomial_pass(void* hv, int id) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, binomial_pass(s->f[id].spatial, s->wrkspace.spatial, s->nx, s->ny));
  return 0;
}
int ref_copy_to_spectral(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, copy_to_spectral(s->f[id], s->nx, s->ny)); return 0; }
int ref_copy_to_spatial(void* hv, int id) { auto* h = (RefHandle*)hv; DISPATCH(h, copy_to_spatial(s->f[id], s->nx, s->ny)); return 0; }
int ref_fft2d(void* hv, int id, int inverse) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, fft2d(id >= 11 ? s->bwrk[id - 11] : s->f[id].spectral, s->nx, s->ny, inverse != 0));
  return 0;
}
int ref_spectral_update(void* hv) { auto* h = (RefHandle*)hv; DISPATCH(h, spectral_update(*s)); return 0; }
int ref_fields_update(void* hv) { auto* h = (RefHandle*)hv; DISPATCH(h, fields_update(*s)); return 0; }

// src/lib.rs:130-142: the untimed move&deposit of every species and the ghost
// fold before the loop.
int ref_prestep(void* hv) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    for (auto& P : s->sp) move_and_deposit(*s, P);
    fold_all(*s);
  });
  return 0;
}
int ref_step(void* hv, int n) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    for (int i = 0; i < n; ++i) step(*s);
  });
  return 0;
}
// dep, solve, push seconds accumulated by ref_step (src/lib.rs:169-197).
int ref_timers(void* hv, double* out) {
  auto* h = (RefHandle*)hv;
  DISPATCH(h, {
    out[0] = s->t_dep;
    out[1] = s->t_solve;
    out[2] = s->t_push;
  });
  return 0;
}

// Standalone 1D transform (the rustfft calls of fft_2d.rs:535-541).
int ref_fft1d(int prec, int64_t n, void* data, int inverse) {
  if (prec == 8) {
    Fft1<double> f; f.plan(n);
    std::vector<Cx<double>> out(n);
    f.run((const Cx<double>*)data, out.data(), inverse ? 1 : -1);
    std::memcpy(data, out.data(), n * sizeof(Cx<double>));
  } else {
    Fft1<float> f; f.plan(n);
    std::vector<Cx<float>> out(n);
    f.run((const Cx<float>*)data, out.data(), inverse ? 1 : -1);
    std::memcpy(data, out.data(), n * sizeof(Cx<float>));
  }
  return 0;
}

int64_t ref_init_positions(int prec, int64_t size_x, int64_t size_y, int64_t delta, int32_t dens, int64_t* ix,
                           int64_t* iy, void* dx, void* dy) {
  Params p{size_x, size_y, delta, dens, 0, 0, 0, 0, 0, 0, 0, 0};
  if (prec == 8) return init_positions<double>(p, ix, iy, (double*)dx, (double*)dy);
  return init_positions<float>(p, ix, iy, (float*)dx, (float*)dy);
}

}  // extern "C"
