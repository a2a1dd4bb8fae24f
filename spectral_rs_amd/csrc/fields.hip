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

// v into n elements of each of `count` arrays (blockIdx.y), one launch, 16-byte
// stores over the 16-byte-aligned body of each array.
template <class T> __global__ void __launch_bounds__(256) k_fill_many(Ptrs8<T> gp, T v, int64_t n) {
  constexpr int V = 16 / (int)sizeof(T);
  typedef T vec __attribute__((ext_vector_type(V)));
  T* p = gp.p[blockIdx.y];
  const int64_t head = min((int64_t)((16 - ((uintptr_t)p & 15)) & 15) / (int64_t)sizeof(T), n);
  const int64_t nv = (n - head) / V;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  vec vv;
#pragma unroll
  for (int k = 0; k < V; ++k) vv[k] = v;
  vec* body = (vec*)(p + head);
  for (int64_t i = t; i < nv; i += stride) body[i] = vv;
  if (t < head) p[t] = v;
  for (int64_t i = head + nv * V + t; i < n; i += stride) p[i] = v;
}

template <class T> void fill_many(T* const* ptrs, int count, T v, int64_t n, hipStream_t st) {
  if (n <= 0 || count <= 0) return;
  Ptrs8<T> gp{};
  for (int i = 0; i < count && i < 8; ++i) gp.p[i] = ptrs[i];
  int64_t b = (n / (16 / (int64_t)sizeof(T)) + 255) / 256;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  hipLaunchKernelGGL(k_fill_many<T>, dim3((unsigned)b, (unsigned)(count < 8 ? count : 8)), dim3(256), 0, st, gp, v,
                     n);
}

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

// Field::binomial_filter_2_d (src/flds/field.rs:385-420) applied NP times,
// each pass followed by the ghost refresh of field.rs:422 -- in one launch.
// A workgroup owns FTY x FTX output cells.  Two LDS buffers share one cell
// frame, (EH + 2) x (EW + 2) with EH = FTY + 2H, EW = FTX + 2H, H = NP - 1:
// frame cell (r, c) is padded cell (r0 + r, c0 + c) of the input, and pass k
// leaves its values at frame rows and columns k .. E + 1 - k in buffer k & 1.
//  * Pass 1 of a workgroup whose frame lies inside the padded grid (all but
//    the edge workgroups) first copies the frame from global memory into
//    buffer 0 and evaluates from LDS.  The frame includes the ghost ring where
//    it touches it, stale after the deposit fold, as the reference reads it.
//  * Pass 1 of an edge workgroup reads each extended cell's 3x3 padded
//    neighbourhood from global memory at its periodic image, so cells next to
//    the domain edge see the input's ghost ring as it stands and cells beyond
//    the edge are the periodic images of interior cells, which is what the
//    refreshed ghosts of passes 2.. hold.
//  * Passes 2..NP run in LDS; the last one writes the interior from registers
//    and, for edge cells, their periodic copies in the output's ghost ring.
// A thread evaluates a strip of SR cells of one column: 3 (SR + 2) LDS reads
// for SR cells instead of 9 per cell, lanes on consecutive columns.  Each
// cell sums its 9 taps in the reference's order, so the result is
// bit-identical to NP separate passes.
constexpr int FTX = 64, FMAX = 8, FSR = 4;
template <class T> constexpr int filter_tile_rows() { return sizeof(T) == 4 ? 32 : 16; }

// The SR cells of frame column c, rows r .. r + SR - 1, from `f` (pitch LW).
template <class T, int LW>
__device__ __forceinline__ void binomial_strip(const T* f, int r, int c, T (&o)[FSR]) {
  const T q = T(0.0625), e = T(0.125), m = T(0.25);
  T a[FSR + 2][3];
  const T* p = f + (r - 1) * LW + (c - 1);
#pragma unroll
  for (int j = 0; j < FSR + 2; ++j) {
    a[j][0] = p[j * LW];
    a[j][1] = p[j * LW + 1];
    a[j][2] = p[j * LW + 2];
  }
#pragma unroll
  for (int s = 0; s < FSR; ++s) {
    T v = q * a[s][0];
    v += e * a[s][1];
    v += q * a[s][2];
    v += e * a[s + 1][0];
    v += m * a[s + 1][1];
    v += e * a[s + 1][2];
    v += q * a[s + 2][0];
    v += e * a[s + 2][1];
    v += q * a[s + 2][2];
    o[s] = v;
  }
}

template <class T, int NP>
__global__ void __launch_bounds__(256) k_binomial_fused(Ptrs8<T> in, Ptrs8<T> out, int nx, int ny) {
  constexpr int FTY = filter_tile_rows<T>();
  constexpr int H = NP - 1, EH = FTY + 2 * H, EW = FTX + 2 * H, LW = EW + 2;
  // frame rows plus the rows the last strip of a pass reads past its end
  constexpr int ROWS = EH + 2 + FSR;
  __shared__ T buf[2][ROWS * LW];
  const T* __restrict__ src = in.p[blockIdx.z];
  T* __restrict__ dst = out.p[blockIdx.z];
  const int P = nx + 2;
  const int r0 = blockIdx.y * FTY - H, c0 = blockIdx.x * FTX - H;  // 0-based interior coords of extended cell (0, 0)
  const bool inner = r0 >= 0 && c0 >= 0 && r0 + EH <= ny && c0 + EW <= nx;
  if (inner) {
    // all of a thread's frame loads in flight before the first LDS store
    constexpr int FR = (EH + 2) * (EW + 2), NL = (FR + 255) / 256;
    T x[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      const int i = threadIdx.x + 256 * j;
      const int r = i / (EW + 2), c = i - r * (EW + 2);
      if (i < FR) x[j] = src[(int64_t)(r0 + r) * P + c0 + c];
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      const int i = threadIdx.x + 256 * j;
      const int r = i / (EW + 2), c = i - r * (EW + 2);
      if (i < FR) buf[0][r * LW + c] = x[j];
    }
  } else {
    // U extended cells per thread at a time, all 9 U loads in flight before
    // the sums (one cell per iteration waited for its loads each time; the
    // small grids of an ensemble have only edge workgroups)
    constexpr int U = 3;
    const T q = T(0.0625), e = T(0.125), m = T(0.25);
    for (int i0 = threadIdx.x; i0 < EH * EW; i0 += U * 256) {
      T a9[U][9];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * 256;
        if (i >= EH * EW) break;
        const int lr = i / EW, lc = i - lr * EW;
        int ir = r0 + lr, ic = c0 + lc;
        if (ir < 0 || ir >= ny) ir = (ir % ny + ny) % ny;
        if (ic < 0 || ic >= nx) ic = (ic % nx + nx) % nx;
        const T* a = src + (int64_t)ir * P + ic;  // padded (ir+1-1, ic+1-1)
#pragma unroll
        for (int t = 0; t < 9; ++t) a9[u][t] = a[(t / 3) * P + t % 3];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = i0 + u * 256;
        if (i >= EH * EW) break;
        const int lr = i / EW, lc = i - lr * EW;
        const T* a = a9[u];
        T v = q * a[0];
        v += e * a[1];
        v += q * a[2];
        v += e * a[3];
        v += m * a[4];
        v += e * a[5];
        v += q * a[6];
        v += e * a[7];
        v += q * a[8];
        buf[1][(lr + 1) * LW + lc + 1] = v;
      }
    }
  }
  // passes 1 (inner workgroups) or 2 (edge workgroups) .. NP - 1 in LDS
#pragma unroll
  for (int k = 1; k < NP; ++k) {
    if (k == 1 && !inner) continue;
    __syncthreads();
    const T* f = buf[(k - 1) & 1];
    T* o = buf[k & 1];
    const int w = EW + 2 - 2 * k, h = EH + 2 - 2 * k, ns = (h + FSR - 1) / FSR;
    for (int i = threadIdx.x; i < w * ns; i += 256) {
      const int sr = i / w, cc = i - sr * w;
      const int r = k + sr * FSR, c = k + cc;
      T v[FSR];
      binomial_strip<T, LW>(f, r, c, v);
#pragma unroll
      for (int s = 0; s < FSR; ++s)
        if (r + s < k + h) o[(r + s) * LW + c] = v[s];
    }
  }
  __syncthreads();
  // pass NP: frame rows and columns NP .. NP + FT - 1 (H + 1 = NP)
  for (int i = threadIdx.x; i < FTX * (FTY / FSR); i += 256) {
    const int sr = i / FTX, cc = i - sr * FTX;
    const int r = NP + sr * FSR, c = NP + cc;
    T v[FSR];
    if (NP == 1 && !inner) {
#pragma unroll
      for (int s = 0; s < FSR; ++s) v[s] = buf[1][(r + s) * LW + c];
    } else {
      binomial_strip<T, LW>(buf[(NP - 1) & 1], r, c, v);
    }
    const int ic = c0 + c - 1;
    if (ic >= nx) continue;
#pragma unroll
    for (int s = 0; s < FSR; ++s) {
      const int ir = r0 + r + s - 1;
      if (ir >= ny) break;
      const int pr = ir + 1, pc = ic + 1;
      dst[(int64_t)pr * P + pc] = v[s];
      // ghost ring = periodic copies (field.rs:141-257)
      const bool top = pr == 1, bot = pr == ny, lft = pc == 1, rgt = pc == nx;
      if (top) dst[(int64_t)(ny + 1) * P + pc] = v[s];
      if (bot) dst[pc] = v[s];
      if (lft) dst[(int64_t)pr * P + nx + 1] = v[s];
      if (rgt) dst[(int64_t)pr * P] = v[s];
      if (top && lft) dst[(int64_t)(ny + 1) * P + nx + 1] = v[s];
      if (top && rgt) dst[(int64_t)(ny + 1) * P] = v[s];
      if (bot && lft) dst[nx + 1] = v[s];
      if (bot && rgt) dst[0] = v[s];
    }
  }
}

template <class T, int NP>
static void launch_fused(const Ptrs8<T>& a, const Ptrs8<T>& b, int ngrids, int nx, int ny, hipStream_t st) {
  constexpr int FTY = filter_tile_rows<T>();
  dim3 grid((nx + FTX - 1) / FTX, (ny + FTY - 1) / FTY, ngrids);
  hipLaunchKernelGGL((k_binomial_fused<T, NP>), grid, dim3(256), 0, st, a, b, nx, ny);
}

// `passes` filter passes on up to 8 grids, FMAX per launch; each launch
// writes the work grid and the two pointers are swapped, so grids[i] holds the
// result afterwards (and work[i] the previous contents).
template <class T>
void launch_binomial(void** grids, void** work, int ngrids, int nx, int ny, int passes, hipStream_t st) {
  for (int done = 0; done < passes && ngrids > 0;) {
    const int n = std::min(FMAX, passes - done);
    Ptrs8<T> a{}, b{};
    for (int i = 0; i < ngrids; ++i) {
      a.p[i] = (T*)grids[i];
      b.p[i] = (T*)work[i];
    }
    switch (n) {
      case 1: launch_fused<T, 1>(a, b, ngrids, nx, ny, st); break;
      case 2: launch_fused<T, 2>(a, b, ngrids, nx, ny, st); break;
      case 3: launch_fused<T, 3>(a, b, ngrids, nx, ny, st); break;
      case 4: launch_fused<T, 4>(a, b, ngrids, nx, ny, st); break;
      case 5: launch_fused<T, 5>(a, b, ngrids, nx, ny, st); break;
      case 6: launch_fused<T, 6>(a, b, ngrids, nx, ny, st); break;
      case 7: launch_fused<T, 7>(a, b, ngrids, nx, ny, st); break;
      default: launch_fused<T, 8>(a, b, ngrids, nx, ny, st); break;
    }
    for (int i = 0; i < ngrids; ++i) std::swap(grids[i], work[i]);
    done += n;
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

// The update of mode i = (row, col) from its spectra jx ... ez; writes E,
// B_wrk and B of the mode (and the zeroed jx, jy at k = 0) and adds the
// mode's Gauss sums.
template <class T>
__device__ __forceinline__ void spectral_mode(const SpecPtrs& sp, int64_t i, int row, int col, int nx, int ny, T cdt,
                                              T ckc, T twopi_x, T twopi_y, Cx<T> jx, Cx<T> jy, Cx<T> jz, Cx<T> rho,
                                              Cx<T> ex, Cx<T> ey, Cx<T> ez, double& g_pre, double& g_post,
                                              double& g_rho, bool eb = true) {
  const bool nyq = col == nx / 2 || row == ny / 2;
  T kx = T(col);
  if (col >= nx / 2 + 1) kx -= T(nx);
  kx *= twopi_x;
  T ky = T(row);
  if (row >= ny / 2 + 1) ky -= T(ny);
  ky *= twopi_y;
  const T kn = (i == 0) ? T(0) : T(1.) / (kx * kx + ky * ky);

  if (i == 0) {  // mod.rs:106-108 (j_z at k=0 kept)
    jx = Cx<T>{0, 0};
    jy = Cx<T>{0, 0};
    st((T*)sp.p[SP_JX], i, jx);
    st((T*)sp.p[SP_JY], i, jy);
  }
  // mod.rs:113-119: B.spectral <- B_wrk
  T* wxp = (T*)sp.p[SP_BXW];
  T* wyp = (T*)sp.p[SP_BYW];
  T* wzp = (T*)sp.p[SP_BZW];
  const Cx<T> bx = ld(wxp, i), by = ld(wyp, i), bz = ld(wzp, i);
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
      g_pre += (double)d.re * d.re + (double)d.im * d.im;
      g_post += ar * ar + ai * ai;
      g_rho += (double)rho.re * rho.re + (double)rho.im * rho.im;
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
  if (!eb) return;  // E and B of this mode are not read back (real-field pipeline)
  st((T*)sp.p[SP_EX], i, ex);
  st((T*)sp.p[SP_EY], i, ey);
  st((T*)sp.p[SP_EZ], i, ez);
  // mod.rs:270-278: B at t+1 = (B_wrk(old) + B_wrk(new)) / 2
  st((T*)sp.p[SP_BX], i, scl(T(0.5), add(bx, wx)));
  st((T*)sp.p[SP_BY], i, scl(T(0.5), add(by, wy)));
  st((T*)sp.p[SP_BZ], i, scl(T(0.5), add(bz, wz)));
}

// Block partials of the Gauss sums in a fixed order (reproducible sums).
__device__ __forceinline__ void gauss_partials(double g_pre, double g_post, double g_rho, double* gauss) {
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

template <class T>
__global__ void __launch_bounds__(256) k_spectral(SpecPtrs sp, int nx, int ny, T cdt, T ckc, T twopi_x, T twopi_y,
                                                  double* gauss) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double g_pre = 0, g_post = 0, g_rho = 0;
  if (i < (int64_t)nx * ny) {
    const int row = (int)(i / nx), col = (int)(i - (int64_t)row * nx);
    const auto in = [&](int id) { return ld((const T*)sp.p[id], i); };
    spectral_mode<T>(sp, i, row, col, nx, ny, cdt, ckc, twopi_x, twopi_y, in(SP_JX), in(SP_JY), in(SP_JZ),
                     in(SP_RHO), in(SP_EX), in(SP_EY), in(SP_EZ), g_pre, g_post, g_rho);
  }
  if (gauss) gauss_partials(g_pre, g_post, g_rho, gauss);
}

// The step's update (the real-field pipeline, even nx and ny): one thread per
// mirrored pair (i, m = -i mod (Ny, Nx)) with i on the half plane kx = 0 ..
// nx/2 (in the self-mirrored columns 0 and nx/2, m >= i).  The seven forward
// spectra (Jx Jy Jz rho Ex Ey Ez) are transforms of real fields, of which the
// forward pipeline wrote the half plane only; mode m of a pair in the other
// columns is the conjugate of mode i and is not read.  The modes of the
// self-mirrored columns are Hermitian-projected here, X[i] <- (X[i] +
// conj(X[m])) / 2, X[m] <- conj(X[i]) (J and rho stored back).  Both modes are
// updated by spectral_mode.  Exact Hermitian inputs matter because B_wrk
// persists in spectral space: its anti-Hermitian part evolves as B_a <- (1 -
// (c dt |k|)^2) B_a each step (the matching E part is dropped by the real-part
// extraction), which grows for c dt |k| > sqrt(2).  With Hermitian inputs
// spectral_mode maps mirrored modes to exactly conjugate outputs, so B_wrk
// stays exactly Hermitian.  E and B are stored on the half plane only, which is
// all the inverse pipeline reads; B_wrk on every mode.  (The context fills the
// other half of the stored J and rho spectra as conjugate mirrors when they
// are read through the API, k_mirror_fill.)
template <class T>
__global__ void __launch_bounds__(256) k_spectral_pair(SpecPtrs sp, int nx, int ny, T cdt, T ckc, T twopi_x,
                                                       T twopi_y, double* gauss) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int H = nx / 2 + 1;
  double g_pre = 0, g_post = 0, g_rho = 0;
  if (t < (int64_t)ny * H) {
    const int row = (int)(t / H), col = (int)(t - (int64_t)row * H);
    const int mrow = (ny - row) % ny, mcol = (nx - col) % nx;
    const int64_t i = (int64_t)row * nx + col, m = (int64_t)mrow * nx + mcol;
    const bool proj = col == 0 || 2 * col == nx;
    if (!proj || m >= i) {
      const int ids[7] = {SP_JX, SP_JY, SP_JZ, SP_RHO, SP_EX, SP_EY, SP_EZ};
      Cx<T> a[7];
      if (proj) {
#pragma unroll
        for (int f = 0; f < 7; ++f) {
          const T* p = (const T*)sp.p[ids[f]];
          const Cx<T> u = ld(p, i), v = ld(p, m);
          a[f] = Cx<T>{T(0.5) * (u.re + v.re), T(0.5) * (u.im - v.im)};
        }
#pragma unroll
        for (int f = 0; f < 4; ++f) {  // E is rewritten by the update
          st((T*)sp.p[ids[f]], i, a[f]);
          if (m != i) st((T*)sp.p[ids[f]], m, Cx<T>{a[f].re, -a[f].im});
        }
      } else {
#pragma unroll
        for (int f = 0; f < 7; ++f) a[f] = ld((const T*)sp.p[ids[f]], i);
      }
      spectral_mode<T>(sp, i, row, col, nx, ny, cdt, ckc, twopi_x, twopi_y, a[0], a[1], a[2], a[3], a[4], a[5],
                       a[6], g_pre, g_post, g_rho, true);
      if (m != i) {
        Cx<T> b[7];
#pragma unroll
        for (int f = 0; f < 7; ++f) b[f] = Cx<T>{a[f].re, -a[f].im};
        spectral_mode<T>(sp, m, mrow, mcol, nx, ny, cdt, ckc, twopi_x, twopi_y, b[0], b[1], b[2], b[3], b[4], b[5],
                         b[6], g_pre, g_post, g_rho, 2 * mcol <= nx);
      }
    }
  }
  if (gauss) gauss_partials(g_pre, g_post, g_rho, gauss);
}

// Sums the block partials of the spectral update (one workgroup, fixed order):
// gauss[0..2] = (pre, post, |rho|^2) of this solve, gauss[3] the largest
// relative post-correction residual so far (gauss_reduce_block).
__global__ void __launch_bounds__(1024) k_gauss_reduce(double* gauss, int nblk) { gauss_reduce_block(gauss, nblk); }

// Columns nx/2 + 1 .. nx - 1 of a spectrum of a real field from the half
// plane: X[r][c] = conj(X[-r][nx - c]) (nx even; the sources lie in columns
// 1 .. nx/2 - 1, which are not written).
template <class T> __global__ void k_mirror_fill(T* spec, int nx, int ny) {
  const int W = nx / 2 - 1;
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (W <= 0 || e >= (int64_t)ny * W) return;
  const int row = (int)(e / W), col = nx / 2 + 1 + (int)(e - (int64_t)row * W);
  const int64_t src = (int64_t)((ny - row) % ny) * nx + (nx - col), dst = (int64_t)row * nx + col;
  spec[2 * dst] = spec[2 * src];
  spec[2 * dst + 1] = -spec[2 * src + 1];
}
template <class T> void launch_mirror_fill(void* spec, int nx, int ny, hipStream_t st) {
  const int64_t n = (int64_t)ny * (nx / 2 - 1);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mirror_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (T*)spec, nx, ny);
}

template <class T> int launch_spectral_update(void* const* spec, const Geo<T>& g, int dens, double* gauss,
                                              hipStream_t st, bool step, bool reduce) {
  SpecPtrs sp{};
  for (int i = 0; i < 14; ++i) sp.p[i] = spec[i];
  const T PI = (T)3.14159265358979323846264338327950288;
  const T twopi_x = T(2.0) * PI / T(g.nx);
  const T twopi_y = T(2.0) * PI / T(g.ny);
  const T ckc = g.dt / T(dens);  // mod.rs:127
  const T cdt = g.dt * g.c;      // mod.rs:128
  const int64_t n = step ? (int64_t)g.ny * (g.nx / 2 + 1) : (int64_t)g.nx * g.ny;
  const int nblk = (int)((n + 255) / 256);
  if (step)
    hipLaunchKernelGGL(k_spectral_pair<T>, dim3((unsigned)nblk), dim3(256), 0, st, sp, g.nx, g.ny, cdt, ckc, twopi_x,
                       twopi_y, gauss ? gauss + GAUSS_PARTIALS : nullptr);
  else
    hipLaunchKernelGGL(k_spectral<T>, dim3((unsigned)nblk), dim3(256), 0, st, sp, g.nx, g.ny, cdt, ckc, twopi_x,
                       twopi_y, gauss ? gauss + GAUSS_PARTIALS : nullptr);
  if (gauss && reduce) hipLaunchKernelGGL(k_gauss_reduce, dim3(1), dim3(1024), 0, st, gauss, nblk);
  return nblk;
}

#define INST(T)                                                                                 \
  template void fill<T>(T*, T, int64_t, hipStream_t);                                           \
  template void fill_many<T>(T* const*, int, T, int64_t, hipStream_t);                          \
  template void launch_fold_ghosts<T>(T* const*, int, int, int, hipStream_t);                   \
  template void launch_update_ghosts<T>(T* const*, int, int, int, hipStream_t);                 \
  template void launch_binomial<T>(void**, void**, int, int, int, int, hipStream_t);              \
  template void launch_mirror_fill<T>(void*, int, int, hipStream_t);                            \
  template int launch_spectral_update<T>(void* const*, const Geo<T>&, int, double*, hipStream_t, bool, bool);
INST(float)
INST(double)
#undef INST
template void fill<int>(int*, int, int64_t, hipStream_t);

}  // namespace sp
