// fft.hip -- hand-written 2D complex FFT for gfx950, standing in for the
// rustfft row passes of Fft2D (src/flds/fft_2d.rs:91-172).
//
// Each 1D transform runs entirely in LDS as a Stockham autosort sequence of
// radix-8/4/2/3 passes (in place: every thread loads its butterflies into
// registers, the block synchronises, then writes).  Rows (x, contiguous) are
// transformed in blocks of whole rows; columns (y, stride Nx) in blocks of C
// adjacent columns so each global access touches C consecutive elements.
// The spatial->spectral copy (real interior of the padded grid, times an
// optional scale: 1/dens for rho, src/flds/mod.rs:79-84) is fused into the
// forward row load, and the 1/(Nx Ny) normalisation, real-part extraction and
// periodic ghost refresh of copy_to_spatial (src/flds/field.rs:109-139) into
// the inverse row store.  Twiddles exp(-2 pi i k/N) are generated in FP64 on
// the device and rounded to Float once.
#pragma clang fp contract(off)

#include <mutex>

#include "sp_internal.h"

#ifndef SP_FFT_XCD_REMAP
#define SP_FFT_XCD_REMAP 1
#endif

namespace sp {

template <class T> struct V2 {
  T x, y;
};
template <class T> __device__ __forceinline__ V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> __device__ __forceinline__ V2<T> operator-(V2<T> a, V2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> __device__ __forceinline__ V2<T> cm(V2<T> a, V2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by -i (forward) or +i (inverse)
template <class T, bool INV> __device__ __forceinline__ V2<T> rot90(V2<T> a) {
  return INV ? V2<T>{-a.y, a.x} : V2<T>{a.y, -a.x};
}

template <class T, bool INV> __device__ __forceinline__ void dft2(V2<T>* v) {
  const V2<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}
template <class T, bool INV> __device__ __forceinline__ void dft3(V2<T>* v) {
  const T s = T(0.86602540378443864676372317075293618);  // sin(pi/3)
  const V2<T> a = v[0], b = v[1], c = v[2];
  const V2<T> t1 = b + c;
  const V2<T> t2 = {a.x - T(0.5) * t1.x, a.y - T(0.5) * t1.y};
  const V2<T> d = b - c;
  const V2<T> t3 = rot90<T, INV>(V2<T>{s * d.x, s * d.y});
  v[0] = a + t1;
  v[1] = t2 + t3;
  v[2] = t2 - t3;
}
template <class T, bool INV> __device__ __forceinline__ void dft4(V2<T>* v) {
  const V2<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3];
  const V2<T> t3 = rot90<T, INV>(v[1] - v[3]);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}
template <class T, bool INV> __device__ __forceinline__ void dft8(V2<T>* v) {
  V2<T> e[4] = {v[0], v[2], v[4], v[6]};
  V2<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, INV>(e);
  dft4<T, INV>(o);
  const T h = T(0.70710678118654752440084436210484904);
  // o[k] *= w8^k, w8 = exp(-+ 2 pi i / 8)
  const V2<T> o1 = INV ? V2<T>{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                       : V2<T>{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
  const V2<T> o2 = rot90<T, INV>(o[2]);
  const V2<T> o3 = INV ? V2<T>{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                       : V2<T>{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
  v[0] = e[0] + o[0];
  v[4] = e[0] - o[0];
  v[1] = e[1] + o1;
  v[5] = e[1] - o1;
  v[2] = e[2] + o2;
  v[6] = e[2] - o2;
  v[3] = e[3] + o3;
  v[7] = e[3] - o3;
}
template <class T, int R, bool INV> __device__ __forceinline__ void dftR(V2<T>* v) {
  if constexpr (R == 2) dft2<T, INV>(v);
  else if constexpr (R == 3) dft3<T, INV>(v);
  else if constexpr (R == 4) dft4<T, INV>(v);
  else dft8<T, INV>(v);
}

// Elements per thread per pass: columns 16; rows SP_FFT_EPT_ROWS (8 gives a
// 4096-point row 512 threads).  The workgroup has at least E / EPT threads.
#ifndef SP_FFT_EPT_ROWS
#define SP_FFT_EPT_ROWS 8
#endif
#ifndef SP_FFT_EPT_COLS
#define SP_FFT_EPT_COLS 8
#endif
constexpr int EPT_ROWS = SP_FFT_EPT_ROWS, EPT_COLS = SP_FFT_EPT_COLS;
constexpr int ROWS_NT_MAX = EPT_ROWS < 16 ? 1024 : 512, COLS_NT_MAX = EPT_COLS < 16 ? 1024 : 512;

// LDS position of logical element i.  Row transforms (PAD) leave one pad slot
// after every 8 elements, so the first pass's butterfly writes (stride R
// elements, 64 bytes at radix 8) fall into distinct banks: it cuts the row
// kernels' LDS wait cycles ~5x and their time ~10% (rocprofv3, 4096-point
// rows).  Column blocks keep the plain layout, which measured faster for them.
template <bool PAD> __device__ __forceinline__ int lp(int i) { return PAD ? i + (i >> 3) : i; }
__host__ __device__ constexpr int lds_slots(int e) { return e + (e >> 3) + 1; }

// Division by a runtime constant d: q = hi32(x * m) with m = ceil(2^32 / d)
// is exact whenever x * d < 2^32 (the error x (m - 2^32/d) / 2^32 < x / 2^32
// stays below the gap 1/d between x/d and the next integer); every use below
// has x * d < 2^27.  A power of two d = 2^sh divides by a shift instead: the
// kernels are built twice, with every divisor of the launch a power of two
// (POW2, radix 2/4/8 plans: shifts only) or not (the multiply; on the
// 4-cycle-per-lane integer multiplier, so the cheaper path matters).
// (d = 1 would need m = 2^32, which does not fit: it is handled apart.)
struct Div {
  uint32_t d, m;
  int sh;  // d = 2^sh, or -1
};
__host__ __device__ inline Div make_div(uint32_t d) {
  Div v;
  v.d = d;
  v.m = d > 1 ? (uint32_t)((0xFFFFFFFFull + d) / d) : 0u;
  int sh = -1;
  if (d != 0 && (d & (d - 1)) == 0) {
    sh = 0;
    while ((1u << sh) != d) ++sh;
  }
  v.sh = sh;
  return v;
}
template <bool POW2> __device__ __forceinline__ uint32_t qdiv(const Div& v, uint32_t x) {
  if constexpr (POW2) return x >> v.sh;
  else return v.d > 1 ? __umulhi(x, v.m) : x;
}
// x * d, a shift under POW2
template <bool POW2> __device__ __forceinline__ int qmul(const Div& v, int x) {
  if constexpr (POW2) return x << v.sh;
  else return x * (int)v.d;
}

// Per-pass constants of one transform length, built on the host.
struct FftPlan {
  const void* twp;  // the plan's twiddles in pass order (FftDim::twp)
  int n, nrad;
  int R[16];          // radix of pass p
  Div nb[16], ns[16];  // n / R_p and the stride Ns_p = R_0 ... R_{p-1}
  Div nsr[16];         // Ns_p R_p
};

// Twiddle k: a 32-bit byte offset from the table's (uniform) base, so the
// load takes the scalar-base form with no 64-bit address arithmetic.  The
// pass-order table makes a pass's loads unit-stride in k: the lanes of a wave
// read consecutive entries (rows) or a few (columns), where the strided reads
// of tw[r k n / (Ns R)] touched up to one cache line per lane and made the
// twiddles about half of the row transforms' time (4096-point rows: 87 us
// with the passes, 42 without).
template <class T> __device__ __forceinline__ V2<T> tw_load(const V2<T>* __restrict__ tw, int k) {
  const uint32_t off = (uint32_t)k * (uint32_t)sizeof(V2<T>);
  return *(const V2<T>*)((const char*)tw + off);
}

// One Stockham pass of radix R over `seqs` sequences of length N held in LDS.
// Sequence s, element n lives at lds[s * sstride + n * estride].  SEQ_FAST
// picks which index varies fastest across threads (columns: the sequence,
// rows: the butterfly) so that LDS accesses stay unit-stride.  The twiddle
// loads of all of a thread's butterflies are issued first, so their latency
// overlaps the LDS reads.
template <class T, int R, bool INV, bool SEQ_FAST, bool POW2, int EPTV>
__device__ __forceinline__ void stockham_pass(V2<T>* lds, int N, int Ns, const Div& fnb, const Div& fns,
                                              const Div& fnsr, const Div& fseq, int seqs, int sstride, int estride,
                                              const V2<T>* __restrict__ twp, int nthreads) {
  constexpr int BPT = (EPTV + R - 1) / R;
  const int nb = N / R;
  const int total = nb * seqs;
  const Div& fq = SEQ_FAST ? fseq : fnb;
  V2<T> v[BPT][R];
  V2<T> w[BPT][R];
  int src[BPT], dst[BPT];
#pragma unroll
  for (int m = 0; m < BPT; ++m) {
    const int b = threadIdx.x + m * nthreads;
    dst[m] = -1;
    if (b < total) {
      const int q = (int)qdiv<POW2>(fq, (uint32_t)b), rm = b - qmul<POW2>(fq, q);
      const int s = SEQ_FAST ? rm : q;
      const int j = SEQ_FAST ? q : rm;
      const int jq = (int)qdiv<POW2>(fns, (uint32_t)j);
      const int k = j - qmul<POW2>(fns, jq);
      if (Ns > 1) {
        const V2<T>* twb = twp + (Ns - 1);  // this pass's entries
#pragma unroll
        for (int r = 1; r < R; ++r) w[m][r] = tw_load<T>(twb, (r - 1) * Ns + k);
      }
      src[m] = s * sstride + j * estride;
      dst[m] = s * sstride + (qmul<POW2>(fnsr, jq) + k) * estride;
    }
  }
#pragma unroll
  for (int m = 0; m < BPT; ++m) {
    if (dst[m] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[m][r] = lds[lp<!SEQ_FAST>(src[m] + r * nb * estride)];
    }
  }
#pragma unroll
  for (int m = 0; m < BPT; ++m) {
    if (dst[m] >= 0) {
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          V2<T> t = w[m][r];
          if (INV) t.y = -t.y;
          v[m][r] = cm(v[m][r], t);
        }
      }
      dftR<T, R, INV>(v[m]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < BPT; ++m) {
    if (dst[m] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) lds[lp<!SEQ_FAST>(dst[m] + r * Ns * estride)] = v[m][r];
    }
  }
  __syncthreads();
}

template <class T, bool INV, bool SEQ_FAST, bool POW2, int EPTV>
__device__ void fft_lds(V2<T>* lds, const FftPlan& pl, const Div& fseq, int seqs, int sstride, int estride,
                        int nthreads) {
  const V2<T>* twp = (const V2<T>*)pl.twp;
  int Ns = 1;
  for (int p = 0; p < pl.nrad; ++p) {
    const int R = pl.R[p];
    const Div fnb = pl.nb[p], fns = pl.ns[p], fnsr = pl.nsr[p];
#define SP_PASS(RR, P2) \
  stockham_pass<T, RR, INV, SEQ_FAST, P2, EPTV>(lds, pl.n, Ns, fnb, fns, fnsr, fseq, seqs, sstride, estride, twp, \
                                                 nthreads)
    if (R == 8) SP_PASS(8, POW2);
    else if (R == 4) SP_PASS(4, POW2);
    else if (R == 2) SP_PASS(2, POW2);
    else if constexpr (!POW2) SP_PASS(3, false);
#undef SP_PASS
    Ns *= R;
  }
}

// Copies element i = threadIdx.x + k nt < E through `ld(i)` and `st(i, v)`,
// eight elements at a time with all their loads issued before the first
// store: a plain strided loop waits for each load before the next (the
// compiler does not hoist loads across iterations of a runtime-length loop),
// one memory round trip per element.
template <class V, class LD, class ST>
__device__ __forceinline__ void staged_copy(int E, int nt, LD ld, ST st) {
  constexpr int B = 8;
  for (int i0 = threadIdx.x; i0 < E; i0 += B * nt) {
    V v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int i = i0 + u * nt;
      if (i < E) v[u] = ld(i);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int i = i0 + u * nt;
      if (i < E) st(i, v[u]);
    }
  }
}

// Row transforms: block (bx, field) handles rows [bx*rpb, bx*rpb + rpb).
// `dn`, `dP`: division by the row length and by the padded row length.
template <class T, bool INV, bool POW2>
__global__ void __launch_bounds__(ROWS_NT_MAX) k_fft_rows(FftBatch b, FftPlan pl, Div dn, Div dP, int ny, int rpb, int load,
                                                  int store, T scale_store) {
  extern __shared__ unsigned char smem[];
  V2<T>* lds = (V2<T>*)smem;
  const int f = blockIdx.y;
  const int nx = pl.n, P = nx + 2;
  const int r0 = blockIdx.x * rpb;
  const int rows = min(rpb, ny - r0);
  const int E = rows * nx;
  const int nt = blockDim.x;
  V2<T>* spec = (V2<T>*)b.spec[f];
  T* grid = (T*)b.grid[f];
  if (load == LOAD_REAL_PADDED) {
    const T sc = (T)b.scale[f];
    const bool scaled = b.scale[f] != 1.0;
    staged_copy<T>(
        E, nt,
        [&](int i) {
          const int s = (int)qdiv<POW2>(dn, i), n = i - s * nx;
          return grid[(r0 + s + 1) * P + n + 1];
        },
        [&](int i, T v) {
          if (scaled) v = v * sc;
          lds[lp<true>(i)] = V2<T>{v, T(0.0)};
        });
  } else {
    staged_copy<V2<T>>(
        E, nt, [&](int i) { return spec[(int64_t)r0 * nx + i]; }, [&](int i, V2<T> v) { lds[lp<true>(i)] = v; });
  }
  __syncthreads();
  fft_lds<T, INV, false, POW2, EPT_ROWS>(lds, pl, dn, rows, nx, 1, nt);
  if (store == STORE_REAL_PADDED) {
    // copy_to_spatial + update_ghosts: interior, the two ghost columns of this
    // row, and the ghost row that mirrors it (row Ny -> row 0, row 1 -> Ny+1).
    for (int i = threadIdx.x; i < rows * P; i += nt) {
      const int s = (int)qdiv<false>(dP, i), c = i - s * P;
      const int n = (c == 0) ? nx - 1 : (c == P - 1 ? 0 : c - 1);
      const T v = lds[lp<true>(s * nx + n)].x * scale_store;
      const int r = r0 + s + 1;
      grid[r * P + c] = v;
      if (r == ny) grid[c] = v;
      if (r == 1) grid[(ny + 1) * P + c] = v;
    }
  } else {
    const bool scaled = scale_store != T(1);
    staged_copy<V2<T>>(
        E, nt, [&](int i) { return lds[lp<true>(i)]; },
        [&](int i, V2<T> v) {
          if (scaled) v = V2<T>{v.x * scale_store, v.y * scale_store};
          spec[(int64_t)r0 * nx + i] = v;
        });
  }
}

// Real-field rows, forward (the step's pipeline, even nx and ny): complex row
// q of block (bx, field) is z = a + i b for the real interior rows a = 2q,
// b = 2q + 1 of the padded grid (times the load scale).  After its transform
// Z, the transforms of the two real rows are A[k] = (Z[k] + conj(Z[-k])) / 2
// and B[k] = (Z[k] - conj(Z[-k])) / 2i; their halves k = 0 .. nx/2 go to the
// spectrum rows a and b (the other halves are their conjugate mirrors, which
// the column pass writes).
template <class T, bool POW2>
__global__ void __launch_bounds__(ROWS_NT_MAX) k_fft_rows_r2c(FftBatch b, FftPlan pl, Div dn, Div dh, int nq, int rpb) {
  extern __shared__ unsigned char smem[];
  V2<T>* lds = (V2<T>*)smem;
  const int f = blockIdx.y;
  const int nx = pl.n, P = nx + 2, H = nx / 2 + 1;
  const int q0 = blockIdx.x * rpb;
  const int rows = min(rpb, nq - q0);
  const int E = rows * nx;
  const int nt = blockDim.x;
  V2<T>* spec = (V2<T>*)b.spec[f];
  const T* grid = (const T*)b.grid[f];
  const T sc = (T)b.scale[f];
  const bool scaled = b.scale[f] != 1.0;
  staged_copy<V2<T>>(
      E, nt,
      [&](int i) {
        const int s = (int)qdiv<POW2>(dn, i), n = i - s * nx;
        const int64_t o = (int64_t)(2 * (q0 + s) + 1) * P + n + 1;  // padded row 2q + 1 = interior row 2q
        return V2<T>{grid[o], grid[o + P]};
      },
      [&](int i, V2<T> v) {
        if (scaled) v = V2<T>{v.x * sc, v.y * sc};
        lds[lp<true>(i)] = v;
      });
  __syncthreads();
  fft_lds<T, false, false, POW2, EPT_ROWS>(lds, pl, dn, rows, nx, 1, nt);
  const T h = T(0.5);
  for (int i = threadIdx.x; i < rows * H; i += nt) {
    const int s = (int)qdiv<false>(dh, i), k = i - s * H;
    const V2<T> zk = lds[lp<true>(s * nx + k)], zm = lds[lp<true>(s * nx + (k == 0 ? 0 : nx - k))];
    const int64_t o = (int64_t)(2 * (q0 + s)) * nx + k;
    spec[o] = V2<T>{h * (zk.x + zm.x), h * (zk.y - zm.y)};
    spec[o + nx] = V2<T>{h * (zk.y + zm.y), h * (zm.x - zk.x)};
  }
}

// Real-field rows, inverse: the spectrum rows a = 2q and b = 2q + 1 hold the
// halves k = 0 .. nx/2 of the row transforms of two real rows (the column
// pass has run on those columns only); the other halves are their conjugate
// mirrors.  Z = A + i B is inverse-transformed once: Re Z -> row a, Im Z -> row
// b, as copy_to_spatial + update_ghosts (the real part, times the norm, the
// ghost columns and the mirrored ghost rows).  The imaginary parts of A and B
// at k = 0 and nx/2 are dropped: in the reference's complex inverse they only
// reach the imaginary part that copy_to_spatial discards.
template <class T, bool POW2>
__global__ void __launch_bounds__(ROWS_NT_MAX) k_fft_rows_c2r(FftBatch b, FftPlan pl, Div dn, Div dP, int ny, int rpb,
                                                          T scale_store) {
  extern __shared__ unsigned char smem[];
  V2<T>* lds = (V2<T>*)smem;
  const int f = blockIdx.y;
  const int nx = pl.n, P = nx + 2, half = nx / 2, nq = ny / 2;
  const int q0 = blockIdx.x * rpb;
  const int rows = min(rpb, nq - q0);
  const int E = rows * nx;
  const int nt = blockDim.x;
  const V2<T>* spec = (const V2<T>*)b.spec[f];
  T* grid = (T*)b.grid[f];
  struct AC {
    V2<T> a, c;
  };
  staged_copy<AC>(
      E, nt,
      [&](int i) {
        const int s = (int)qdiv<POW2>(dn, i), k = i - s * nx;
        const int64_t o = (int64_t)(2 * (q0 + s)) * nx + (k <= half ? k : nx - k);
        return AC{spec[o], spec[o + nx]};
      },
      [&](int i, AC v) {
        const int s = (int)qdiv<POW2>(dn, i), k = i - s * nx;
        V2<T> a = v.a, c = v.c, z;
        if (k <= half) {
          if (k == 0 || k == half) a.y = c.y = T(0);
          z = V2<T>{a.x - c.y, a.y + c.x};
        } else {  // A[k] = conj(A[nx - k]), B[k] = conj(B[nx - k])
          z = V2<T>{a.x + c.y, c.x - a.y};
        }
        lds[lp<true>(i)] = z;
      });
  __syncthreads();
  fft_lds<T, true, false, POW2, EPT_ROWS>(lds, pl, dn, rows, nx, 1, nt);
  for (int i = threadIdx.x; i < rows * P; i += nt) {
    const int s = (int)qdiv<false>(dP, i), c = i - s * P;
    const int n = (c == 0) ? nx - 1 : (c == P - 1 ? 0 : c - 1);
    const V2<T> z = lds[lp<true>(s * nx + n)];
    const T va = z.x * scale_store, vb = z.y * scale_store;
    const int ra = 2 * (q0 + s) + 1, rb = ra + 1;  // padded rows of the pair
    grid[(int64_t)ra * P + c] = va;
    grid[(int64_t)rb * P + c] = vb;
    if (ra == 1) grid[(int64_t)(ny + 1) * P + c] = va;
    if (rb == ny) grid[c] = vb;
  }
}

// Column transforms: block (bx, field) handles columns [bx*C, bx*C + C) of the
// first ncols columns.  POW2 launches have C a power of two and every block's
// column count a power of two.  MIRROR (forward transforms of real fields, on
// columns 0 .. nx/2): column nx - c is written as the conjugate mirror of
// column c, F[-ky][nx - c] = conj(F[ky][c]), except for the self-mirrored
// columns 0 and nx/2.
template <class T, bool INV, bool POW2, bool MIRROR>
__global__ void __launch_bounds__(COLS_NT_MAX) k_fft_cols(FftBatch b, FftPlan pl, int nx, int ncols, int C,
                                                          T scale_store) {
  extern __shared__ unsigned char smem[];
  V2<T>* lds = (V2<T>*)smem;
  const int f = blockIdx.y;
  const int ny = pl.n;
  // Blocks of C < 16 columns read parts of 128-byte lines that neighbouring
  // blocks also read: each XCD takes a contiguous run of column blocks, so the
  // neighbours share its L2.
  const int bx = SP_FFT_XCD_REMAP ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
  const int c0 = bx * C;
  const int cols = min(C, ncols - c0);
  const int E = cols * ny;
  const int nt = blockDim.x;
  if (b.gauss && blockIdx.x == 0 && blockIdx.y == 0) gauss_reduce_block(b.gauss, b.gauss_nblk);
  V2<T>* spec = (V2<T>*)b.spec[f];
  const Div fc = make_div((uint32_t)cols);
  staged_copy<V2<T>>(
      E, nt,
      [&](int i) {
        const int n = (int)qdiv<POW2>(fc, i), c = i - n * cols;
        return spec[(int64_t)n * nx + c0 + c];
      },
      [&](int i, V2<T> v) { lds[i] = v; });
  __syncthreads();
  fft_lds<T, INV, true, POW2, EPT_COLS>(lds, pl, fc, cols, 1, cols, nt);
  const bool scaled = scale_store != T(1);
  for (int i = threadIdx.x; i < E; i += nt) {
    const int n = (int)qdiv<POW2>(fc, i), c = i - n * cols;
    V2<T> v = lds[i];
    if (scaled) v = V2<T>{v.x * scale_store, v.y * scale_store};
    spec[(int64_t)n * nx + c0 + c] = v;
    if constexpr (MIRROR) {
      const int cc = c0 + c;
      if (cc != 0 && 2 * cc != nx) spec[(int64_t)(n == 0 ? 0 : ny - n) * nx + (nx - cc)] = V2<T>{v.x, -v.y};
    }
  }
}

template <class T> __global__ void k_twiddles(V2<T>* tw, int n, int fp32_gen) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (fp32_gen) {  // precision study: the twiddle computed in FP32
    float s, c;
    sincospif(-2.0f * (float)k / (float)n, &s, &c);
    tw[k] = V2<T>{(T)c, (T)s};
    return;
  }
  double s, c;
  sincospi(-2.0 * (double)k / (double)n, &s, &c);  // FP64 generation
  tw[k] = V2<T>{(T)c, (T)s};
}

// ceil(e / ept) rounded up to whole waves; `cap` >= e / ept for every block
// the launchers form (e <= fft_max_len)
static int round_threads(int e, int ept, int cap) {
  int t = (e + ept - 1) / ept;
  t = ((t + 63) / 64) * 64;
  return t < 64 ? 64 : (t > cap ? cap : t);
}
static_assert(8192 / EPT_COLS <= COLS_NT_MAX && 8192 / EPT_ROWS <= ROWS_NT_MAX, "a block's threads cover its elements");
template <class T> static int max_elems() { return fft_max_len<T>(); }

// The padded LDS of the longest row block exceeds the 64 KB default limit of
// dynamic shared memory; gfx950 has 160 KB per workgroup.
template <class T> static void allow_big_lds() {
  static std::once_flag once;
  std::call_once(once, [] {
    const int bytes = lds_slots(fft_max_len<T>()) * (int)sizeof(V2<T>);
    // a refusal surfaces as the launch's own error
    const void* k[8] = {(const void*)k_fft_rows<T, false, false>, (const void*)k_fft_rows<T, true, false>,
                        (const void*)k_fft_rows<T, false, true>,  (const void*)k_fft_rows<T, true, true>,
                        (const void*)k_fft_rows_r2c<T, false>,    (const void*)k_fft_rows_r2c<T, true>,
                        (const void*)k_fft_rows_c2r<T, false>,    (const void*)k_fft_rows_c2r<T, true>};
    for (const void* f : k) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  });
}

// twp[e] for e < n - 1: the entry's pass p (Ns_p - 1 <= e < Ns_{p+1} - 1),
// then r and k; the values are copies of tw.  twp[n - 1] is unused (1).
template <class T>
__global__ void k_twiddles_pass(V2<T>* __restrict__ twp, const V2<T>* __restrict__ tw, int n, uint64_t code,
                                int nrad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  V2<T> v{T(1), T(0)};
  int ns = 1;
  for (int p = 0; p < nrad && p < 16; ++p) {
    const int R = (int)((code >> (4 * p)) & 15u);
    if (e < ns * R - 1) {
      const int loc = e - (ns - 1), r = loc / ns + 1, k = loc - (r - 1) * ns;
      v = tw[(int64_t)r * k * (n / (ns * R))];
      break;
    }
    ns *= R;
  }
  twp[e] = v;
}

template <class T> void fft_pass_twiddles(const FftDim& d, hipStream_t st) {
  hipLaunchKernelGGL(k_twiddles_pass<T>, dim3((d.n + 255) / 256), dim3(256), 0, st, (V2<T>*)d.twp,
                     (const V2<T>*)d.tw, d.n, d.radix_code, d.nrad);
}

template <class T> void fft_make_twiddles(FftDim& d, hipStream_t st, bool fp32_gen) {
  allow_big_lds<T>();  // at set-up, outside any stream capture
  hipLaunchKernelGGL(k_twiddles<T>, dim3((d.n + 255) / 256), dim3(256), 0, st, (V2<T>*)d.tw, d.n, (int)fp32_gen);
  fft_pass_twiddles<T>(d, st);
}

// The pass constants of a plan; `pow2`: every radix is 2, 4 or 8.
static FftPlan make_plan(const FftDim& d, bool& pow2) {
  FftPlan pl{};
  pl.twp = d.twp;
  pl.n = d.n;
  pl.nrad = d.nrad;
  pow2 = true;
  int Ns = 1;
  for (int p = 0; p < d.nrad && p < 16; ++p) {
    const int R = (int)((d.radix_code >> (4 * p)) & 15u);
    pl.R[p] = R;
    pl.nb[p] = make_div((uint32_t)(d.n / R));
    pl.ns[p] = make_div((uint32_t)Ns);
    pl.nsr[p] = make_div((uint32_t)(Ns * R));
    if (R == 3) pow2 = false;
    Ns *= R;
  }
  return pl;
}

template <class T>
void launch_fft_rows(const FftBatch& b, const FftDim& dx, int ny, int inverse, int load, int store,
                     double store_scale, hipStream_t st) {
  // rows per workgroup: the plan's choice (default 1), within LDS and the grid
  int rpb = dx.per_block > 0 ? dx.per_block : 1;
  if (rpb > max_elems<T>() / dx.n) rpb = max_elems<T>() / dx.n;
  if (rpb < 1) rpb = 1;
  if (rpb > ny) rpb = ny;
  const int E = rpb * dx.n;
  const int nt = round_threads(E, EPT_ROWS, ROWS_NT_MAX);
  const size_t shm = (size_t)lds_slots(E) * sizeof(V2<T>);
  allow_big_lds<T>();
  dim3 grid((ny + rpb - 1) / rpb, b.count);
  const T scale = (T)store_scale;
  bool pow2;
  const FftPlan pl = make_plan(dx, pow2);
  const Div dn = make_div((uint32_t)dx.n), dP = make_div((uint32_t)(dx.n + 2));
  const auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(nt), shm, st, b, pl, dn, dP, ny, rpb, load, store, scale);
  };
  if (pow2) inverse ? go(k_fft_rows<T, true, true>) : go(k_fft_rows<T, false, true>);
  else inverse ? go(k_fft_rows<T, true, false>) : go(k_fft_rows<T, false, false>);
}

// Rows per workgroup of a row launch over `nrows` transforms: the plan's
// choice (default 1), within LDS and the grid.
template <class T> static int rows_per_block(const FftDim& dx, int nrows) {
  int rpb = dx.per_block > 0 ? dx.per_block : 1;
  if (rpb > max_elems<T>() / dx.n) rpb = max_elems<T>() / dx.n;
  if (rpb < 1) rpb = 1;
  if (rpb > nrows) rpb = nrows;
  return rpb;
}

template <class T> void launch_fft_rows_r2c(const FftBatch& b, const FftDim& dx, int ny, hipStream_t st) {
  const int nq = ny / 2, rpb = rows_per_block<T>(dx, nq);
  const int E = rpb * dx.n;
  const int nt = round_threads(E, EPT_ROWS, ROWS_NT_MAX);
  const size_t shm = (size_t)lds_slots(E) * sizeof(V2<T>);
  allow_big_lds<T>();
  dim3 grid((nq + rpb - 1) / rpb, b.count);
  bool pow2;
  const FftPlan pl = make_plan(dx, pow2);
  const Div dn = make_div((uint32_t)dx.n), dh = make_div((uint32_t)(dx.n / 2 + 1));
  if (pow2) hipLaunchKernelGGL((k_fft_rows_r2c<T, true>), grid, dim3(nt), shm, st, b, pl, dn, dh, nq, rpb);
  else hipLaunchKernelGGL((k_fft_rows_r2c<T, false>), grid, dim3(nt), shm, st, b, pl, dn, dh, nq, rpb);
}

template <class T>
void launch_fft_rows_c2r(const FftBatch& b, const FftDim& dx, int ny, double store_scale, hipStream_t st) {
  const int nq = ny / 2, rpb = rows_per_block<T>(dx, nq);
  const int E = rpb * dx.n;
  const int nt = round_threads(E, EPT_ROWS, ROWS_NT_MAX);
  const size_t shm = (size_t)lds_slots(E) * sizeof(V2<T>);
  allow_big_lds<T>();
  dim3 grid((nq + rpb - 1) / rpb, b.count);
  bool pow2;
  const FftPlan pl = make_plan(dx, pow2);
  const Div dn = make_div((uint32_t)dx.n), dP = make_div((uint32_t)(dx.n + 2));
  const T scale = (T)store_scale;
  if (pow2) hipLaunchKernelGGL((k_fft_rows_c2r<T, true>), grid, dim3(nt), shm, st, b, pl, dn, dP, ny, rpb, scale);
  else hipLaunchKernelGGL((k_fft_rows_c2r<T, false>), grid, dim3(nt), shm, st, b, pl, dn, dP, ny, rpb, scale);
}

// ncols: the first ncols columns (nx for a full transform, nx/2 + 1 for the
// real-field pipeline); mirror: write the conjugate mirrors (forward only).
template <class T>
void launch_fft_cols(const FftBatch& b, const FftDim& dy, int nx, int inverse, double store_scale, hipStream_t st,
                     int ncols, bool mirror) {
  if (ncols <= 0) ncols = nx;
  // columns per workgroup: the plan's choice (default 8), within LDS and the grid
  int C = dy.per_block > 0 ? dy.per_block : 8;
  if (C > max_elems<T>() / dy.n) C = max_elems<T>() / dy.n;
  if (C < 1) C = 1;
  if (C > ncols) C = ncols;
  const int E = C * dy.n;
  const int nt = round_threads(E, EPT_COLS, COLS_NT_MAX);
  const size_t shm = (size_t)E * sizeof(V2<T>);
  dim3 grid((ncols + C - 1) / C, b.count);
  const T scale = (T)store_scale;
  bool pow2;
  const FftPlan pl = make_plan(dy, pow2);
  const int rem = ncols % C;
  pow2 = pow2 && (C & (C - 1)) == 0 && (rem & (rem - 1)) == 0;
  const auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(nt), shm, st, b, pl, nx, ncols, C, scale); };
  if (mirror && !inverse) pow2 ? go(k_fft_cols<T, false, true, true>) : go(k_fft_cols<T, false, false, true>);
  else if (pow2) inverse ? go(k_fft_cols<T, true, true, false>) : go(k_fft_cols<T, false, true, false>);
  else inverse ? go(k_fft_cols<T, true, false, false>) : go(k_fft_cols<T, false, false, false>);
}

#define INST(T)                                                                                 \
  template void fft_make_twiddles<T>(FftDim&, hipStream_t, bool);                               \
  template void fft_pass_twiddles<T>(const FftDim&, hipStream_t);                               \
  template void launch_fft_rows<T>(const FftBatch&, const FftDim&, int, int, int, int, double, hipStream_t); \
  template void launch_fft_cols<T>(const FftBatch&, const FftDim&, int, int, double, hipStream_t, int, bool); \
  template void launch_fft_rows_r2c<T>(const FftBatch&, const FftDim&, int, hipStream_t);                    \
  template void launch_fft_rows_c2r<T>(const FftBatch&, const FftDim&, int, double, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace sp
