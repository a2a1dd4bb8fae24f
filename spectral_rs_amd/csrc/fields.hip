// fields.hip -- grid kernels for gfx950: fill, ghost fold (deposit_ghosts),
// periodic ghost refresh (update_ghosts), the binomial filter and the
// pointwise spectral Maxwell update.  All are HBM-streaming kernels; every
// arithmetic expression keeps the reference's association (contraction off),
// so on identical inputs the results are bit-identical to the reference
// arithmetic.
#pragma clang fp contract(off)

#include "sp_internal.h"

namespace sp {

template <class T> __global__ void k_fill(T* p, T v, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <class T> void fill(T* p, T v, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)b), dim3(256), 0, st, p, v, n);
}

template <class T> struct Ptrs8 {
  T* p[8];
};

// Field::deposit_ghosts (src/flds/field.rs:259-373).  One thread per interior
// edge cell; each cell adds its ghost contributions in the reference's order:
// top ghost row -> row Ny, bottom ghost row -> row 1, left ghost column ->
// column Nx, right ghost column -> column 1, then the four corners.  The ring
// is left untouched (it is read again by the first filter pass).
template <class T> __global__ void k_fold(Ptrs8<T> gp, int nx, int ny) {
  T* f = gp.p[blockIdx.y];
  const int P = nx + 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  // edge cells enumerated as: rows 1 and Ny (all columns), then columns 1 and
  // Nx of rows 2..Ny-1.
  int r, c;
  if (t < 2 * nx) {
    r = (t < nx) ? 1 : ny;
    c = (t < nx ? t : t - nx) + 1;
    if (ny == 1 && t >= nx) return;
  } else {
    const int u = t - 2 * nx;
    if (u >= 2 * (ny - 2) || ny < 3) return;
    r = 2 + (u >> 1);
    c = (u & 1) ? nx : 1;
    if (nx == 1 && (u & 1)) return;
  }
  T v = f[r * P + c];
  if (r == ny) v += f[c];                      // field.rs:271-283
  if (r == 1) v += f[(ny + 1) * P + c];        // field.rs:284-296
  if (c == nx) v += f[r * P];                  // field.rs:297-317
  if (c == 1) v += f[r * P + nx + 1];          // field.rs:319-339
  if (r == ny && c == nx) v += f[0];           // field.rs:341-346
  if (r == ny && c == 1) v += f[nx + 1];       // field.rs:348-353
  if (r == 1 && c == nx) v += f[(ny + 1) * P]; // field.rs:355-364
  if (r == 1 && c == 1) v += f[(ny + 1) * P + nx + 1];  // field.rs:366-372
  f[r * P + c] = v;
}

template <class T> void launch_fold_ghosts(T* const* grids, int ngrids, int nx, int ny, hipStream_t st) {
  Ptrs8<T> gp{};
  for (int i = 0; i < ngrids; ++i) gp.p[i] = grids[i];
  const int cells = 2 * nx + 2 * (ny > 2 ? ny - 2 : 0);
  hipLaunchKernelGGL(k_fold<T>, dim3((cells + 255) / 256, ngrids), dim3(256), 0, st, gp, nx, ny);
}

// Field::update_ghosts (src/flds/field.rs:141-257): one thread per ring cell
// copies its periodic image from the interior.
template <class T> __device__ __forceinline__ void ring_copy(T* f, int t, int nx, int ny) {
  const int P = nx + 2;
  if (t < nx) {  // top ghost row <- row Ny
    f[t + 1] = f[ny * P + t + 1];
  } else if (t < 2 * nx) {  // bottom ghost row <- row 1
    const int c = t - nx + 1;
    f[(ny + 1) * P + c] = f[P + c];
  } else if (t < 2 * nx + ny) {  // left ghost column <- column Nx
    const int r = t - 2 * nx + 1;
    f[r * P] = f[r * P + nx];
  } else if (t < 2 * nx + 2 * ny) {  // right ghost column <- column 1
    const int r = t - 2 * nx - ny + 1;
    f[r * P + nx + 1] = f[r * P + 1];
  } else if (t == 2 * nx + 2 * ny) {
    f[0] = f[ny * P + nx];
  } else if (t == 2 * nx + 2 * ny + 1) {
    f[nx + 1] = f[ny * P + 1];
  } else if (t == 2 * nx + 2 * ny + 2) {
    f[(ny + 1) * P] = f[P + nx];
  } else if (t == 2 * nx + 2 * ny + 3) {
    f[(ny + 1) * P + nx + 1] = f[P + 1];
  }
}

template <class T> __global__ void k_update_ghosts(Ptrs8<T> gp, int nx, int ny) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  ring_copy(gp.p[blockIdx.y], t, nx, ny);
}

template <class T> void launch_update_ghosts(T* const* grids, int ngrids, int nx, int ny, hipStream_t st) {
  Ptrs8<T> gp{};
  for (int i = 0; i < ngrids; ++i) gp.p[i] = grids[i];
  const int cells = 2 * nx + 2 * ny + 4;
  hipLaunchKernelGGL(k_update_ghosts<T>, dim3((cells + 255) / 256, ngrids), dim3(256), 0, st, gp, nx, ny);
}

// One pass of Field::binomial_filter_2_d (src/flds/field.rs:385-420) over the
// interior: 9 taps summed in the reference's order, reading the input's ring
// as it stands.  Tiles of 64 x 4 outputs stage their 66 x 6 input in LDS.
constexpr int BTX = 64, BTY = 4;
template <class T> __global__ void __launch_bounds__(BTX * BTY) k_binomial(Ptrs8<T> in, Ptrs8<T> out, int nx, int ny) {
  __shared__ T tile[BTY + 2][BTX + 2];
  const T* __restrict__ src = in.p[blockIdx.z];
  T* __restrict__ dst = out.p[blockIdx.z];
  const int P = nx + 2;
  const int c0 = blockIdx.x * BTX, r0 = blockIdx.y * BTY;  // interior offsets (0-based)
  for (int i = threadIdx.y * BTX + threadIdx.x; i < (BTY + 2) * (BTX + 2); i += BTX * BTY) {
    const int lr = i / (BTX + 2), lc = i % (BTX + 2);
    const int gr = r0 + lr, gc = c0 + lc;  // padded coordinates of the tile corner (r0, c0)
    tile[lr][lc] = (gr <= ny + 1 && gc <= nx + 1) ? src[gr * P + gc] : T(0);
  }
  __syncthreads();
  const int r = r0 + threadIdx.y + 1, c = c0 + threadIdx.x + 1;
  if (r > ny || c > nx) return;
  const int y = threadIdx.y, x = threadIdx.x;
  T v = T(0.0625) * tile[y][x];
  v += T(0.125) * tile[y][x + 1];
  v += T(0.0625) * tile[y][x + 2];
  v += T(0.125) * tile[y + 1][x];
  v += T(0.25) * tile[y + 1][x + 1];
  v += T(0.125) * tile[y + 1][x + 2];
  v += T(0.0625) * tile[y + 2][x];
  v += T(0.125) * tile[y + 2][x + 1];
  v += T(0.0625) * tile[y + 2][x + 2];
  dst[r * P + c] = v;
}

// n passes on up to 8 grids.  Each pass filters grid -> work, refreshes the
// work ring (field.rs:422) and swaps roles; an odd count ends with a copy
// back so the result always lands in the caller's grid.
template <class T> __global__ void k_copy(Ptrs8<T> from, Ptrs8<T> to, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) to.p[blockIdx.y][i] = from.p[blockIdx.y][i];
}

template <class T>
void launch_binomial(T* const* grids, T* const* work, int ngrids, int nx, int ny, int passes, hipStream_t st) {
  if (passes <= 0 || ngrids <= 0) return;
  Ptrs8<T> a{}, b{};
  for (int i = 0; i < ngrids; ++i) { a.p[i] = grids[i]; b.p[i] = work[i]; }
  dim3 grid((nx + BTX - 1) / BTX, (ny + BTY - 1) / BTY, ngrids);
  for (int k = 0; k < passes; ++k) {
    hipLaunchKernelGGL(k_binomial<T>, grid, dim3(BTX, BTY), 0, st, a, b, nx, ny);
    launch_update_ghosts<T>(b.p, ngrids, nx, ny, st);
    std::swap(a, b);
  }
  if (passes & 1) {
    const int64_t n = (int64_t)(nx + 2) * (ny + 2);
    hipLaunchKernelGGL(k_copy<T>, dim3((unsigned)((n + 255) / 256), ngrids), dim3(256), 0, st, a, b, n);
  }
}

// ------------------------------------------------------- spectral Maxwell
template <class T> struct Cx {
  T re, im;
};
template <class T> __device__ __forceinline__ Cx<T> ld(const T* p, int64_t i) { return Cx<T>{p[2 * i], p[2 * i + 1]}; }
template <class T> __device__ __forceinline__ void st(T* p, int64_t i, Cx<T> v) {
  p[2 * i] = v.re;
  p[2 * i + 1] = v.im;
}
// num_complex arithmetic: (a+bi)(c+di) = (ac - bd) + (ad + bc)i ; scalar
// products act on both parts.
template <class T> __device__ __forceinline__ Cx<T> mul(Cx<T> a, Cx<T> b) {
  return Cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> __device__ __forceinline__ Cx<T> scl(T r, Cx<T> a) { return Cx<T>{r * a.re, r * a.im}; }
template <class T> __device__ __forceinline__ Cx<T> add(Cx<T> a, Cx<T> b) { return Cx<T>{a.re + b.re, a.im + b.im}; }
template <class T> __device__ __forceinline__ Cx<T> sub(Cx<T> a, Cx<T> b) { return Cx<T>{a.re - b.re, a.im - b.im}; }

struct SpecPtrs {
  void* p[14];  // indexed by sp_grid (spectra) and SP_BXW..SP_BZW
};

// Flds::update between the forward and the inverse transforms
// (src/flds/mod.rs:105-278), one thread per mode.  The wave numbers are
// rebuilt from the mode index with the reference's operations
// (src/flds/wave_num.rs:22-37), so they equal its tables bit for bit.  Every
// reference loop in this range is pointwise in the mode, except the saves of
// Ex[0], Ey[0] around the Gauss correction, which thread 0 handles locally.
// It also accumulates, in FP64, the Gauss-law residual |k.E + i rho|^2 before
// and after the correction and |rho|^2 over the modes that are neither k=0 nor
// on a Nyquist line (gauss[0..2]); the reference has no such check.
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <class T>
__global__ void __launch_bounds__(256) k_spectral(SpecPtrs sp, int nx, int ny, T cdt, T ckc, T twopi_x, T twopi_y,
                                                  double* gauss) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double g_pre = 0, g_post = 0, g_rho = 0;
  if (i < (int64_t)nx * ny) {
  const int row = (int)(i / nx), col = (int)(i - (int64_t)row * nx);
  const bool nyq = col == nx / 2 || row == ny / 2;
  T kx = T(col);
  if (col >= nx / 2 + 1) kx -= T(nx);
  kx *= twopi_x;
  T ky = T(row);
  if (row >= ny / 2 + 1) ky -= T(ny);
  ky *= twopi_y;
  const T kn = (i == 0) ? T(0) : T(1.) / (kx * kx + ky * ky);

  T* jxp = (T*)sp.p[SP_JX];
  T* jyp = (T*)sp.p[SP_JY];
  Cx<T> jx = ld(jxp, i), jy = ld(jyp, i);
  const Cx<T> jz = ld((const T*)sp.p[SP_JZ], i);
  const Cx<T> rho = ld((const T*)sp.p[SP_RHO], i);
  if (i == 0) {  // mod.rs:106-108 (j_z at k=0 kept)
    jx = Cx<T>{0, 0};
    jy = Cx<T>{0, 0};
    st(jxp, i, jx);
    st(jyp, i, jy);
  }
  // mod.rs:113-119: B.spectral <- B_wrk
  T* wxp = (T*)sp.p[SP_BXW];
  T* wyp = (T*)sp.p[SP_BYW];
  T* wzp = (T*)sp.p[SP_BZW];
  const Cx<T> bx = ld(wxp, i), by = ld(wyp, i), bz = ld(wzp, i);
  T* exp_ = (T*)sp.p[SP_EX];
  T* eyp = (T*)sp.p[SP_EY];
  T* ezp = (T*)sp.p[SP_EZ];
  Cx<T> ex = ld(exp_, i), ey = ld(eyp, i), ez = ld(ezp, i);
  const Cx<T> icdt{T(0.0), cdt}, micdt{T(0.0), -cdt};
  // mod.rs:130-173
  ex = add(ex, sub(mul(Cx<T>{icdt.re * ky, icdt.im * ky}, bz), scl(ckc, jx)));
  ey = add(ey, sub(mul(Cx<T>{micdt.re * kx, micdt.im * kx}, bz), scl(ckc, jy)));
  ez = add(ez, mul(icdt, sub(scl(kx, by), scl(ky, bx))));
  ez = sub(ez, scl(ckc, jz));
  // mod.rs:177-207: Gauss correction; k=0 keeps its uncorrected value.
  if (i != 0) {
    const Cx<T> irho = mul(Cx<T>{T(0.), T(1.)}, rho);
    const Cx<T> d = add(add(scl(kx, ex), scl(ky, ey)), irho);
    const Cx<T> tmp{d.re * kn, d.im * kn};
    ex = sub(ex, Cx<T>{tmp.re * kx, tmp.im * kx});
    ey = sub(ey, Cx<T>{tmp.re * ky, tmp.im * ky});
    if (!nyq) {
      const double ar = (double)kx * ex.re + (double)ky * ey.re - (double)rho.im;
      const double ai = (double)kx * ex.im + (double)ky * ey.im + (double)rho.re;
      g_pre = (double)d.re * d.re + (double)d.im * d.im;
      g_post = ar * ar + ai * ai;
      g_rho = (double)rho.re * rho.re + (double)rho.im * rho.im;
    }
  }
  // mod.rs:211-225: B_wrk to the next half step.
  Cx<T> wx = sub(bx, mul(Cx<T>{icdt.re * ky, icdt.im * ky}, ez));
  Cx<T> wy = add(by, mul(Cx<T>{icdt.re * kx, icdt.im * kx}, ez));
  Cx<T> wz = add(bz, mul(icdt, sub(scl(ky, ex), scl(kx, ey))));
  // mod.rs:230-266: Nyquist column Nx/2 and row Ny/2 of B_wrk and E.
  if (nyq) {
    const Cx<T> z{T(0), T(0)};
    wx = wy = wz = z;
    ex = ey = ez = z;
  }
  st(wxp, i, wx);
  st(wyp, i, wy);
  st(wzp, i, wz);
  st(exp_, i, ex);
  st(eyp, i, ey);
  st(ezp, i, ez);
  // mod.rs:270-278: B at t+1 = (B_wrk(old) + B_wrk(new)) / 2
  st((T*)sp.p[SP_BX], i, scl(T(0.5), add(bx, wx)));
  st((T*)sp.p[SP_BY], i, scl(T(0.5), add(by, wy)));
  st((T*)sp.p[SP_BZ], i, scl(T(0.5), add(bz, wz)));
  }
  if (gauss) {  // block partials in a fixed order (reproducible sums)
    __shared__ double red[3][4];
    g_pre = wave_sum_d(g_pre);
    g_post = wave_sum_d(g_post);
    g_rho = wave_sum_d(g_rho);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[0][w] = g_pre;
      red[1][w] = g_post;
      red[2][w] = g_rho;
    }
    __syncthreads();
    if (threadIdx.x < 3)
      gauss[(int64_t)blockIdx.x * 3 + threadIdx.x] =
          ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
  }
}

// Sums the block partials of k_spectral (one block, fixed order) into
// out[0..2] = (pre, post, |rho|^2) of this solve, and keeps in out[3] the
// largest relative post-correction residual sqrt(post / |rho|^2) seen so far.
__global__ void __launch_bounds__(256) k_gauss_reduce(const double* part, int nblk, double* out) {
  __shared__ double red[3][4];
  double s[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int k = 0; k < 3; ++k) s[k] += part[(int64_t)b * 3 + k];
  const int w = threadIdx.x >> 6;
  for (int k = 0; k < 3; ++k) {
    s[k] = wave_sum_d(s[k]);
    if ((threadIdx.x & 63) == 0) red[k][w] = s[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[3];
    for (int k = 0; k < 3; ++k) t[k] = ((red[k][0] + red[k][1]) + red[k][2]) + red[k][3];
    out[0] = t[0];
    out[1] = t[1];
    out[2] = t[2];
    const double rel = t[2] > 0 ? sqrt(t[1] / t[2]) : sqrt(t[1]);
    out[3] = fmax(out[3], rel);
  }
}

// Hermitian projection X[k] <- (X[k] + conj(X[-k])) / 2, X[-k] <- conj(X[k])
// of spectra that are transforms of real fields.  Exact in exact arithmetic;
// in floating point it removes the anti-Hermitian rounding component.  That
// component matters because B_wrk persists in spectral space: its
// anti-Hermitian part evolves as B_a <- (1 - (c dt |k|)^2) B_a each step (the
// matching E part is dropped by the real-part extraction), which grows for
// c dt |k| > sqrt(2).  With Hermitian inputs the pointwise update of
// k_spectral maps mirrored modes to exactly conjugate outputs, so B_wrk stays
// exactly Hermitian.  One thread per mode; the lower index of each mirrored
// pair writes both.
template <class T> __global__ void __launch_bounds__(256) k_hermitian(SpecPtrs sp, int count, int nx, int ny) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)nx * ny) return;
  const int row = (int)(i / nx), col = (int)(i - (int64_t)row * nx);
  const int64_t m = (int64_t)((ny - row) % ny) * nx + (nx - col) % nx;
  if (m < i) return;
  for (int f = 0; f < count; ++f) {
    T* p = (T*)sp.p[f];
    const Cx<T> a = ld(p, i), b = ld(p, m);
    const Cx<T> s{T(0.5) * (a.re + b.re), T(0.5) * (a.im - b.im)};
    st(p, i, s);
    if (m != i) st(p, m, Cx<T>{s.re, -s.im});
  }
}

template <class T> void launch_hermitian(void* const* spec, int count, int nx, int ny, hipStream_t st) {
  SpecPtrs sp{};
  for (int i = 0; i < count && i < 14; ++i) sp.p[i] = spec[i];
  const int64_t n = (int64_t)nx * ny;
  hipLaunchKernelGGL(k_hermitian<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sp, count, nx, ny);
}

template <class T> void launch_spectral_update(void* const* spec, const Geo<T>& g, int dens, double* gauss,
                                               hipStream_t st) {
  SpecPtrs sp{};
  for (int i = 0; i < 14; ++i) sp.p[i] = spec[i];
  const T PI = (T)3.14159265358979323846264338327950288;
  const T twopi_x = T(2.0) * PI / T(g.nx);
  const T twopi_y = T(2.0) * PI / T(g.ny);
  const T ckc = g.dt / T(dens);  // mod.rs:127
  const T cdt = g.dt * g.c;      // mod.rs:128
  const int64_t n = (int64_t)g.nx * g.ny;
  const int nblk = (int)((n + 255) / 256);
  hipLaunchKernelGGL(k_spectral<T>, dim3((unsigned)nblk), dim3(256), 0, st, sp, g.nx, g.ny, cdt, ckc, twopi_x,
                     twopi_y, gauss ? gauss + GAUSS_PARTIALS : nullptr);
  if (gauss) hipLaunchKernelGGL(k_gauss_reduce, dim3(1), dim3(256), 0, st, gauss + GAUSS_PARTIALS, nblk, gauss);
}

#define INST(T)                                                                                 \
  template void fill<T>(T*, T, int64_t, hipStream_t);                                           \
  template void launch_fold_ghosts<T>(T* const*, int, int, int, hipStream_t);                   \
  template void launch_update_ghosts<T>(T* const*, int, int, int, hipStream_t);                 \
  template void launch_binomial<T>(T* const*, T* const*, int, int, int, int, hipStream_t);      \
  template void launch_spectral_update<T>(void* const*, const Geo<T>&, int, double*, hipStream_t);            \
  template void launch_hermitian<T>(void* const*, int, int, int, hipStream_t);
INST(float)
INST(double)
#undef INST
template void fill<int>(int*, int, int64_t, hipStream_t);

}  // namespace sp
