// diag.hip -- FP64 diagnostics (no reference counterpart): field energies,
// particle kinetic energy and momentum (FP32 accumulation on request, for the
// precision study).  The Gauss-law residuals are reduced
// inside k_spectral (fields.hip) every step.
//
// Units: with the reference's normalisation (J and rho divided by dens,
// src/flds/mod.rs:80-84, :127) the field energy W = sum (E^2 + B^2)/2 over the
// interior cells and the particle energy sum m c^2 (gamma - 1) / dens change
// by opposite amounts (dW/dt = -sum E.J/dens, dK/dt = sum q v.E), so their sum
// is the conserved total.
#include "sp_internal.h"

namespace sp {

template <class A> __device__ __forceinline__ A wave_sum(A v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// out[0] += sum over the interior of g0^2 + g1^2 + g2^2, accumulated in A
// (FP64; FP32 per thread and per wave for the precision study).
template <class T, class A> __global__ void __launch_bounds__(256) k_sumsq3(const T* g0, const T* g1, const T* g2,
                                                                             int nx, int ny, double* out) {
  const int P = nx + 2;
  A acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)nx * ny;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(i / nx) + 1, c = (int)(i % nx) + 1;
    const int64_t o = (int64_t)r * P + c;
    const A a = g0[o], b = g1[o], d = g2[o];
    acc += a * a + b * b + d * d;
  }
  const A tot = wave_sum(acc);  // per thread and per wave in A, across waves in FP64
  if ((threadIdx.x & 63) == 0) atomicAdd(out, (double)tot);
}

// out[0] += sum (gamma - 1), out[1..3] += sum p, gamma from p, accumulated in A.
template <class T, class A> __global__ void __launch_bounds__(256) k_kinetic(SpeciesDev<T> s, A csqinv, double* out) {
  A e = 0, x = 0, y = 0, z = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < s.n; i += (int64_t)gridDim.x * blockDim.x) {
    const A px = s.px[i], py = s.py[i], pz = s.pz[i];
    const A u2 = (px * px + py * py + pz * pz) * csqinv;
    e += u2 / (sqrt(A(1) + u2) + A(1));  // gamma - 1 without cancellation
    x += px;
    y += py;
    z += pz;
  }
  e = wave_sum(e);  // per thread and per wave in A, across waves in FP64
  x = wave_sum(x);
  y = wave_sum(y);
  z = wave_sum(z);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(out + 0, (double)e);
    atomicAdd(out + 1, (double)x);
    atomicAdd(out + 2, (double)y);
    atomicAdd(out + 3, (double)z);
  }
}

static unsigned grid_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

template <class T>
void launch_sumsq3(const T* a, const T* b, const T* c, int nx, int ny, double* out, hipStream_t st, bool sum32) {
  if (sum32)
    hipLaunchKernelGGL((k_sumsq3<T, float>), dim3(grid_blocks((int64_t)nx * ny)), dim3(256), 0, st, a, b, c, nx, ny,
                       out);
  else
    hipLaunchKernelGGL((k_sumsq3<T, double>), dim3(grid_blocks((int64_t)nx * ny)), dim3(256), 0, st, a, b, c, nx, ny,
                       out);
}
template <class T> void launch_kinetic(const SpeciesDev<T>& s, double csqinv, double* out, hipStream_t st, bool sum32) {
  if (s.n == 0) return;
  if (sum32)
    hipLaunchKernelGGL((k_kinetic<T, float>), dim3(grid_blocks(s.n)), dim3(256), 0, st, s, (float)csqinv, out);
  else
    hipLaunchKernelGGL((k_kinetic<T, double>), dim3(grid_blocks(s.n)), dim3(256), 0, st, s, csqinv, out);
}

#define INST(T)                                                                                 \
  template void launch_sumsq3<T>(const T*, const T*, const T*, int, int, double*, hipStream_t, bool); \
  template void launch_kinetic<T>(const SpeciesDev<T>&, double, double*, hipStream_t, bool);
INST(float)
INST(double)
#undef INST

}  // namespace sp
