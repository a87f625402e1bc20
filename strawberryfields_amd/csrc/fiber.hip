// Per-fiber kernels for large states: one thread owns one fiber of a gate.
//
// A fiber is the set of amplitudes a gate mixes for one assignment of the
// other axes: the D amplitudes along a single-mode axis, one photon-number
// anti-diagonal (N, s - N) of a beamsplitter, or one diagonal (p, p + d) of a
// mode's (bra, ket) pair for the loss channel.  The thread reads the fiber
// once into registers, transforms it, and writes it once, so a pass moves
// exactly 2 S bytes (forward) or 3 S (adjoint: state, cotangent in; cotangent
// out) whatever the axis stride.  Consecutive threads take consecutive values
// of the innermost non-gate axes, so every load and store is coalesced when
// that run is long (the dispatcher uses these kernels when it is >= 16
// amplitudes and the element-wise kernels otherwise).  D is compile-time
// (2..20): fibers live in registers and every loop is unrolled with guards.
#include "common.h"
#include "fiber.h"

namespace {

template <typename T>
using C = cpx<T>;

// fiber index f -> base offset for gate axes lo < hi (strides Slo > Shi) of an
// R-axis tensor: f enumerates (a, c, b) = (axes before lo, between, after hi)
struct Geo2 {
  int64_t B, C;        // extents after hi, between lo and hi
  int64_t s_lo, s_hi;  // strides of the two gate axes
};

__device__ __forceinline__ int64_t fiber_base2(int64_t f, const Geo2& g, int D) {
  const int64_t b = f % g.B;
  const int64_t t = f / g.B;
  const int64_t c = t % g.C;
  const int64_t a = t / g.C;
  return a * (D * g.s_lo) + c * (D * g.s_hi) + b;
}

__device__ __forceinline__ void fib_write_partials(double g0, double g1, double* partial, int nwg, int zp1) {
  __shared__ double sh[8];
  const double s0 = block_sum(g0, sh);
  const double s1 = block_sum(g1, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s0;
    partial[nwg + blockIdx.x] = s1;
    if (zp1) {
      partial[2 * nwg + blockIdx.x] = 0.0;
      partial[3 * nwg + blockIdx.x] = 0.0;
    }
  }
}

// ------------------------------------------------------------------ single-mode axis
template <typename T, int D>
__global__ void __launch_bounds__(256)
k_axis_dense_fib(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t NF, int64_t B,
                 const C<T>* __restrict__ M, int conj, int parity) {
  GRID_STRIDE(f, NF) {
    const int64_t a = f / B, b = f - a * B;
    const int64_t base = a * D * B + b;
    C<T> x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = in[base + k * B];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      C<T> acc = mk<T>(0, 0);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (parity && ((i + k) & 1)) continue;
        const C<T> m = M[i * D + k];
        if (conj)
          cmac_conj(acc, m, x[k]);
        else
          cmac(acc, m, x[k]);
      }
      out[base + i * B] = acc;
    }
  }
}

// adjoint: lam_in = M'^H lam_out; g_p += Re sum_i conj(lam_out[i]) (dM'_p psi)[i]
template <typename T, int D>
__global__ void __launch_bounds__(256)
k_axis_dense_adj_fib(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t NF,
                     int64_t B, const C<T>* __restrict__ M, const C<T>* __restrict__ dM0,
                     const C<T>* __restrict__ dM1, int conj, int parity, double* partial, int nwg, int zp1) {
  double g0 = 0.0, g1 = 0.0;
  GRID_STRIDE(f, NF) {
    const int64_t a = f / B, b = f - a * B;
    const int64_t base = a * D * B + b;
    C<T> x[D], l[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      x[k] = psi[base + k * B];
      l[k] = lo_[base + k * B];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      // lam_in[k] = sum_i conj(M'[i][k]) lam[i];  M' = conj(M) if conj
      C<T> acc = mk<T>(0, 0);
#pragma unroll
      for (int i = 0; i < D; ++i) {
        if (parity && ((i + k) & 1)) continue;
        const C<T> m = M[i * D + k];
        if (conj)
          cmac(acc, m, l[i]);
        else
          cmac_conj(acc, m, l[i]);
      }
      li[base + k * B] = acc;
    }
    T s0 = 0, s1 = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      C<T> u0 = mk<T>(0, 0), u1 = mk<T>(0, 0);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (parity && ((i + k) & 1)) continue;
        if (conj) {
          cmac_conj(u0, dM0[i * D + k], x[k]);
          cmac_conj(u1, dM1[i * D + k], x[k]);
        } else {
          cmac(u0, dM0[i * D + k], x[k]);
          cmac(u1, dM1[i * D + k], x[k]);
        }
      }
      s0 += re_conj_mul(l[i], u0);
      s1 += re_conj_mul(l[i], u1);
    }
    g0 += (double)s0;
    g1 += (double)s1;
  }
  fib_write_partials(g0, g1, partial, nwg, zp1);
}

// ------------------------------------------------------------------ beamsplitter
// Work is split into tasks (block s, chunk of kBlock fibers) with s taken from
// the task index, which is uniform over the workgroup: every coefficient
// address is then wave-uniform and the block's matrix is read through scalar
// loads.  Amplitudes of block s are indexed relative to lo = bs_lo(s):
// t = N - lo, at offset t * (sp - sq) from the block's first amplitude.
template <typename T, int D>
__global__ void __launch_bounds__(256)
k_pair_bs_fib(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t NF, Geo2 g, int64_t sp, int64_t sq,
              const C<T>* __restrict__ blk, int conj, int diff) {
  constexpr int NSB = 2 * D - 1;
  const int64_t ntask = (NF + kBlock - 1) / kBlock * NSB;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int s = (int)(task % NSB);
    const int64_t f = (task / NSB) * kBlock + threadIdx.x;
    if (f >= NF) continue;
    const int lo = bs_lo(s, D);
    const int c = bs_hi(s, D) - lo + 1;
    const C<T>* bb = blk + bs_block_off(s, D);
    // photon-number blocks: amplitudes (N, s - N); difference blocks (diff, two-mode
    // squeezing): (N, N - d), d = s - (D - 1)
    const int64_t base = fiber_base2(f, g, D) + lo * sp + (int64_t)(diff ? lo - (s - (D - 1)) : s - lo) * sq;
    const int64_t step = diff ? sp + sq : sp - sq;
    C<T> x[D];
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < c) x[t] = in[base + t * step];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      if (u >= c) continue;
      C<T> acc = mk<T>(0, 0);
#pragma unroll
      for (int t = 0; t < D; ++t) {
        if (t >= c) continue;
        if (conj)
          cmac_conj(acc, bb[u * c + t], x[t]);
        else
          cmac(acc, bb[u * c + t], x[t]);
      }
      out[base + u * step] = acc;
    }
  }
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
k_pair_bs_adj_fib(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t NF,
                  Geo2 g, int64_t sp, int64_t sq, const C<T>* __restrict__ blk, const C<T>* __restrict__ dblk,
                  int conj, int wa, int wb, int diff, double* partial, int nwg, int zp1) {
  constexpr int NSB = 2 * D - 1;
  const T sg = conj ? (T)-1 : (T)1;
  const int64_t ntask = (NF + kBlock - 1) / kBlock * NSB;
  double g0 = 0.0, g1 = 0.0;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int s = (int)(task % NSB);
    const int64_t f = (task / NSB) * kBlock + threadIdx.x;
    if (f >= NF) continue;
    const int lo = bs_lo(s, D);
    const int c = bs_hi(s, D) - lo + 1;
    const C<T>* bb = blk + bs_block_off(s, D);
    const C<T>* db = dblk + bs_block_off(s, D);
    // photon-number blocks: amplitudes (N, s - N); difference blocks (diff, two-mode
    // squeezing): (N, N - d), d = s - (D - 1)
    const int64_t base = fiber_base2(f, g, D) + lo * sp + (int64_t)(diff ? lo - (s - (D - 1)) : s - lo) * sq;
    const int64_t step = diff ? sp + sq : sp - sq;
    C<T> x[D], l[D];
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < c) {
        x[t] = psi[base + t * step];
        l[t] = lo_[base + t * step];
      }
    T s0 = 0, s1 = 0;
#pragma unroll
    for (int n = 0; n < D; ++n) {
      if (n >= c) continue;
      // lam_in[n] = sum_N conj(B'[N][n]) lam[N];  w0 = sum_N conj(lam[N]) dB'[N][n];
      // w1 = sum_N conj(lam[N]) i sg (wa n + wb N) B'[N][n]   (n, N relative to lo)
      C<T> acc = mk<T>(0, 0), w0 = mk<T>(0, 0), w1 = mk<T>(0, 0);
#pragma unroll
      for (int N = 0; N < D; ++N) {
        if (N >= c) continue;
        C<T> b = bb[N * c + n];
        C<T> d = db[N * c + n];
        if (conj) {
          b.im = -b.im;
          d.im = -d.im;
        }
        cmac_conj(acc, b, l[N]);
        cmac_conj(w0, l[N], d);
        const C<T> lb = mk<T>(l[N].re * b.re + l[N].im * b.im, l[N].re * b.im - l[N].im * b.re);
        const T wt = sg * (T)(wa * (lo + n) + wb * (lo + N));
        w1.re += -wt * lb.im;
        w1.im += wt * lb.re;
      }
      li[base + n * step] = acc;
      s0 += w0.re * x[n].re - w0.im * x[n].im;
      s1 += w1.re * x[n].re - w1.im * x[n].im;
    }
    g0 += (double)s0;
    g1 += (double)s1;
  }
  fib_write_partials(g0, g1, partial, nwg, zp1);
}

// ------------------------------------------------------------------ loss
// (bra, ket) axes of one mode: strides sp = D * sq, sq = B.  Tasks are
// (diagonal d = q - p, chunk of kBlock fibers f = (a, b)); d is uniform over
// the workgroup, so the coefficient reads are scalar loads.
template <typename T, int D>
__global__ void __launch_bounds__(256)
k_loss_fib(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t NF, int64_t B, const T* __restrict__ coef) {
  constexpr int NSB = 2 * D - 1;
  const int64_t ntask = (NF + kBlock - 1) / kBlock * NSB;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int d = (int)(task % NSB) - (D - 1);
    const int64_t f = (task / NSB) * kBlock + threadIdx.x;
    if (f >= NF) continue;
    const int64_t a = f / B, b = f - a * B;
    const int p0 = d < 0 ? -d : 0;  // first p with q = p + d >= 0
    const int L = D - (d < 0 ? -d : d);
    const int64_t base = a * D * D * B + b + (int64_t)(p0 * D + p0 + d) * B;
    const int64_t step = (int64_t)(D + 1) * B;
    const T* cf0 = coef + ((int64_t)p0 * D + (p0 + d)) * D;
    C<T> x[D];
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < L) x[t] = in[base + t * step];
#pragma unroll
    for (int t = 0; t < D; ++t) {
      if (t >= L) continue;
      const T* cf = cf0 + t * (D + 1) * D;
      C<T> acc = mk<T>(0, 0);
#pragma unroll
      for (int l = 0; l < D; ++l) {
        if (t + l >= L) continue;
        acc.re += cf[l] * x[t + l].re;
        acc.im += cf[l] * x[t + l].im;
      }
      out[base + t * step] = acc;
    }
  }
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
k_loss_adj_fib(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t NF,
               int64_t B, const T* __restrict__ coef, const T* __restrict__ dcoef, double* partial, int nwg,
               int zp1) {
  constexpr int NSB = 2 * D - 1;
  const int64_t ntask = (NF + kBlock - 1) / kBlock * NSB;
  double g0 = 0.0;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int d = (int)(task % NSB) - (D - 1);
    const int64_t f = (task / NSB) * kBlock + threadIdx.x;
    if (f >= NF) continue;
    const int64_t a = f / B, b = f - a * B;
    const int p0 = d < 0 ? -d : 0;
    const int L = D - (d < 0 ? -d : d);
    const int64_t base = a * D * D * B + b + (int64_t)(p0 * D + p0 + d) * B;
    const int64_t step = (int64_t)(D + 1) * B;
    const int64_t c0 = ((int64_t)p0 * D + (p0 + d)) * D;  // coefficient row of position t = 0
    C<T> x[D], l[D];
#pragma unroll
    for (int t = 0; t < D; ++t)
      if (t < L) {
        x[t] = psi[base + t * step];
        l[t] = lo_[base + t * step];
      }
    T s0 = 0;
#pragma unroll
    for (int t = 0; t < D; ++t) {
      if (t >= L) continue;
      // lam_in at position t: sum over outputs u = t - lp (lp >= 0) of c[u][lp] lam[u]
      C<T> acc = mk<T>(0, 0);
#pragma unroll
      for (int lp = 0; lp < D; ++lp) {
        if (lp > t) continue;
        const int u = t - lp;
        const T cc = coef[c0 + (int64_t)u * (D + 1) * D + lp];
        acc.re += cc * l[u].re;
        acc.im += cc * l[u].im;
      }
      li[base + t * step] = acc;
      const T* dc = dcoef + c0 + (int64_t)t * (D + 1) * D;
      C<T> u = mk<T>(0, 0);
#pragma unroll
      for (int lp = 0; lp < D; ++lp) {
        if (t + lp >= L) continue;
        u.re += dc[lp] * x[t + lp].re;
        u.im += dc[lp] * x[t + lp].im;
      }
      s0 += re_conj_mul(l[t], u);
    }
    g0 += (double)s0;
  }
  fib_write_partials(g0, 0.0, partial, nwg, zp1);
}

// ------------------------------------------------------------------ dispatch
#define SFAMD_FIB_D(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)

Geo2 geo2(int64_t sp, int64_t sq, int64_t N, int D) {
  Geo2 g;
  g.s_lo = sp > sq ? sp : sq;
  g.s_hi = sp > sq ? sq : sp;
  g.B = g.s_hi;
  g.C = g.s_lo / (D * g.s_hi);
  (void)N;
  return g;
}

// one workgroup per (block or diagonal, chunk of kBlock fibers), capped
int task_grid(int64_t NF, int D) {
  int64_t t = (NF + kBlock - 1) / kBlock * (2 * D - 1);
  if (t > kMaxGrid) t = kMaxGrid;
  return (int)t;
}

}  // namespace

namespace sfamd {

bool fiber_supported(int D) { return D >= 2 && D <= kFiberMaxD; }

#define CP(T, p) ((const C<T>*)(p))
#define MP(T, p) ((C<T>*)(p))

int axis_dense_fib(const void* in, void* out, int64_t N, int D, int64_t B, const void* M, int conj, int parity,
                   int dtype, hipStream_t s) {
  const int64_t NF = N / D;
  const int grid = grid_for(NF);
  switch (D) {
#define CASE(d)                                                                                                     \
  case d:                                                                                                           \
    if (dtype == SFAMD_C64)                                                                                         \
      hipLaunchKernelGGL((k_axis_dense_fib<float, d>), dim3(grid), dim3(kBlock), 0, s, CP(float, in),             \
                         MP(float, out), NF, B, CP(float, M), conj, parity);                                        \
    else                                                                                                            \
      hipLaunchKernelGGL((k_axis_dense_fib<double, d>), dim3(grid), dim3(kBlock), 0, s, CP(double, in),           \
                         MP(double, out), NF, B, CP(double, M), conj, parity);                                      \
    break;
    SFAMD_FIB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

int axis_dense_adj_fib(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t B, const void* M,
                       const void* dM0, const void* dM1, int conj, int parity, double* partial, int nwg, int zp1,
                       int dtype, hipStream_t s) {
  const int64_t NF = N / D;
  switch (D) {
#define CASE(d)                                                                                                       \
  case d:                                                                                                             \
    if (dtype == SFAMD_C64)                                                                                           \
      hipLaunchKernelGGL((k_axis_dense_adj_fib<float, d>), dim3(nwg), dim3(kBlock), 0, s, CP(float, psi),           \
                         CP(float, lam_out), MP(float, lam_in), NF, B, CP(float, M), CP(float, dM0), CP(float, dM1), \
                         conj, parity, partial, nwg, zp1);                                                            \
    else                                                                                                              \
      hipLaunchKernelGGL((k_axis_dense_adj_fib<double, d>), dim3(nwg), dim3(kBlock), 0, s, CP(double, psi),         \
                         CP(double, lam_out), MP(double, lam_in), NF, B, CP(double, M), CP(double, dM0),            \
                         CP(double, dM1), conj, parity, partial, nwg, zp1);                                           \
    break;
    SFAMD_FIB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

int pair_bs_fib(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
                int dtype, hipStream_t s, int diff) {
  const int64_t NF = N / ((int64_t)D * D);
  const Geo2 g = geo2(sp, sq, N, D);
  const int grid = task_grid(NF, D);
  switch (D) {
#define CASE(d)                                                                                                    \
  case d:                                                                                                          \
    if (dtype == SFAMD_C64)                                                                                        \
      hipLaunchKernelGGL((k_pair_bs_fib<float, d>), dim3(grid), dim3(kBlock), 0, s, CP(float, in), MP(float, out), \
                         NF, g, sp, sq, CP(float, blocks), conj, diff);                                            \
    else                                                                                                           \
      hipLaunchKernelGGL((k_pair_bs_fib<double, d>), dim3(grid), dim3(kBlock), 0, s, CP(double, in),             \
                         MP(double, out), NF, g, sp, sq, CP(double, blocks), conj, diff);                          \
    break;
    SFAMD_FIB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

int pair_bs_adj_fib(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                    const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial, int nwg,
                    int zp1, int dtype, hipStream_t s, int diff) {
  const int64_t NF = N / ((int64_t)D * D);
  const Geo2 g = geo2(sp, sq, N, D);
  switch (D) {
#define CASE(d)                                                                                                      \
  case d:                                                                                                            \
    if (dtype == SFAMD_C64)                                                                                          \
      hipLaunchKernelGGL((k_pair_bs_adj_fib<float, d>), dim3(nwg), dim3(kBlock), 0, s, CP(float, psi),             \
                         CP(float, lam_out), MP(float, lam_in), NF, g, sp, sq, CP(float, blocks), CP(float, dblocks), \
                         conj, wa, wb, diff, partial, nwg, zp1);                                                     \
    else                                                                                                             \
      hipLaunchKernelGGL((k_pair_bs_adj_fib<double, d>), dim3(nwg), dim3(kBlock), 0, s, CP(double, psi),           \
                         CP(double, lam_out), MP(double, lam_in), NF, g, sp, sq, CP(double, blocks),               \
                         CP(double, dblocks), conj, wa, wb, diff, partial, nwg, zp1);                                \
    break;
    SFAMD_FIB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

int loss_fib(const void* in, void* out, int64_t N, int D, int64_t B, const void* coef, int dtype, hipStream_t s) {
  const int64_t NF = N / ((int64_t)D * D);
  const int grid = task_grid(NF, D);
  switch (D) {
#define CASE(d)                                                                                                  \
  case d:                                                                                                        \
    if (dtype == SFAMD_C64)                                                                                      \
      hipLaunchKernelGGL((k_loss_fib<float, d>), dim3(grid), dim3(kBlock), 0, s, CP(float, in), MP(float, out), \
                         NF, B, (const float*)coef);                                                             \
    else                                                                                                         \
      hipLaunchKernelGGL((k_loss_fib<double, d>), dim3(grid), dim3(kBlock), 0, s, CP(double, in),              \
                         MP(double, out), NF, B, (const double*)coef);                                           \
    break;
    SFAMD_FIB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

int loss_adj_fib(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t B, const void* coef,
                 const void* dcoef, double* partial, int nwg, int zp1, int dtype, hipStream_t s) {
  const int64_t NF = N / ((int64_t)D * D);
  switch (D) {
#define CASE(d)                                                                                                     \
  case d:                                                                                                           \
    if (dtype == SFAMD_C64)                                                                                         \
      hipLaunchKernelGGL((k_loss_adj_fib<float, d>), dim3(nwg), dim3(kBlock), 0, s, CP(float, psi),               \
                         CP(float, lam_out), MP(float, lam_in), NF, B, (const float*)coef, (const float*)dcoef,     \
                         partial, nwg, zp1);                                                                        \
    else                                                                                                            \
      hipLaunchKernelGGL((k_loss_adj_fib<double, d>), dim3(nwg), dim3(kBlock), 0, s, CP(double, psi),             \
                         CP(double, lam_out), MP(double, lam_in), NF, B, (const double*)coef, (const double*)dcoef, \
                         partial, nwg, zp1);                                                                        \
    break;
    SFAMD_FIB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

}  // namespace sfamd
