// Runs of consecutive rotation and Kerr gates applied in one pass.
//
// R(theta) = diag(e^{i theta n}) and K(kappa) = diag(e^{i kappa n^2})
// (fockbackend/ops.py:274-281,309-314) are diagonal in the Fock basis, so a
// run of them on any modes multiplies each amplitude by one phase: on a ket
// prod_g t_g[i_{m_g}], on a DM prod_g t_g[b_{m_g}] conj(t_g[k_{m_g}]) (the
// two-sided update of fockbackend/circuit.py:118-217).  One pass over the
// state replaces one pass per gate.  The adjoint pass returns
// lam_in = conj(phase) lam_out and, for every gate of the run,
//   dL/dp_g = Re sum conj(lam_out) i w_g out = -sum w_g Im(conj(lam_out) out),
// w_g = n or n^2 of the gate's mode (b - k weights on a DM).
//
// Indexing: the state is viewed as rows of R amplitudes, R = D (ket) or D^2
// (DM: the bra and ket axes of the last mode); a lane takes 4 consecutive
// amplitudes of one row, so the other modes' digits are fixed per lane and the
// last mode's digits follow from the position in the row.
#include "common.h"

namespace {

template <typename T>
using C = cpx<T>;

using sfamd::DiagRun;
constexpr int kRun = sfamd::kDiagRunMax;

// phase tables of the run's gates in LDS: tab[g * D + n]
template <typename T>
__device__ __forceinline__ void stage_tables(const C<T>* __restrict__ table, const DiagRun& r, C<T>* tab) {
  for (int e = threadIdx.x; e < r.ngates * r.D; e += blockDim.x) {
    const int g = e / r.D, i = e - g * r.D;
    tab[e] = table[r.toff[g] + i];
  }
  __syncthreads();
}

// digits of every gate's mode for the row `row` (modes other than the last);
// the last mode's digits are filled per amplitude
struct RowDigits {
  int b[kRun], k[kRun];
};

__device__ __forceinline__ void row_digits(const DiagRun& r, int64_t row, RowDigits& d) {
  const int64_t R = r.mixed ? (int64_t)r.D * r.D : r.D;
#pragma unroll
  for (int g = 0; g < kRun; ++g) {
    d.b[g] = d.k[g] = 0;
    if (g < r.ngates && r.mode[g] < r.n - 1) {
      const int64_t p = (row / r.rowdiv[g]) % R;  // digit pair of mode m in the row index
      d.b[g] = r.mixed ? (int)(p / r.D) : (int)p;
      d.k[g] = r.mixed ? (int)(p % r.D) : 0;
    }
  }
}

template <typename T>
__device__ __forceinline__ C<T> run_phase(const DiagRun& r, const C<T>* tab, const RowDigits& d, int lb, int lk) {
  C<T> f = mk<T>((T)1, (T)0);
#pragma unroll
  for (int g = 0; g < kRun; ++g) {
    if (g < r.ngates) {
      const bool last = r.mode[g] == r.n - 1;
      const int b = last ? lb : d.b[g], k = last ? lk : d.k[g];
      C<T> t = tab[g * r.D + b];
      if (r.mixed) t = cmul(t, cconj(tab[g * r.D + k]));
      f = cmul(f, t);
    }
  }
  return f;
}

// four consecutive amplitudes of a row: 16-byte nontemporal accesses for
// complex64 when rows hold whole groups of four (R % 4 == 0), else one by one
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ void load4(const C<T>* __restrict__ p, int valid, bool vec, C<T> (&x)[4]) {
  if constexpr (sizeof(C<T>) == 8) {
    if (vec) {
      const f4v a = __builtin_nontemporal_load((const f4v*)p);
      const f4v b = __builtin_nontemporal_load((const f4v*)p + 1);
      x[0] = mk<T>(a[0], a[1]);
      x[1] = mk<T>(a[2], a[3]);
      x[2] = mk<T>(b[0], b[1]);
      x[3] = mk<T>(b[2], b[3]);
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) x[j] = j < valid ? p[j] : mk<T>((T)0, (T)0);
}

template <typename T>
__device__ __forceinline__ void store4(C<T>* __restrict__ p, int valid, bool vec, const C<T> (&x)[4]) {
  if constexpr (sizeof(C<T>) == 8) {
    if (vec) {
      __builtin_nontemporal_store(f4v{x[0].re, x[0].im, x[1].re, x[1].im}, (f4v*)p);
      __builtin_nontemporal_store(f4v{x[2].re, x[2].im, x[3].re, x[3].im}, (f4v*)p + 1);
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < valid) p[j] = x[j];
}

template <typename T>
__global__ void __launch_bounds__(256)
k_diag_run(const C<T>* __restrict__ in, C<T>* __restrict__ out, const C<T>* __restrict__ table, DiagRun r) {
  __shared__ C<T> tab[kRun * SFAMD_MAX_CUTOFF];
  stage_tables(table, r, tab);
  const int64_t R = r.mixed ? (int64_t)r.D * r.D : r.D;
  const int64_t cpr = (R + 3) / 4, rows = r.N / R;
  const bool vec = R % 4 == 0;
  GRID_STRIDE(w, rows * cpr) {
    const int64_t row = w / cpr;
    const int c0 = (int)(w - row * cpr) * 4;
    const int valid = R - c0 < 4 ? (int)(R - c0) : 4;
    RowDigits d;
    row_digits(r, row, d);
    const int64_t e0 = row * R + c0;
    C<T> x[4];
    load4(in + e0, valid, vec, x);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= valid) continue;
      const int c = c0 + j;
      const int lb = r.mixed ? c / r.D : c, lk = r.mixed ? c % r.D : 0;
      x[j] = cmul(run_phase(r, tab, d, lb, lk), x[j]);
    }
    store4(out + e0, valid, vec, x);
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_diag_run_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li,
               const C<T>* __restrict__ table, DiagRun r, double* partial, int nwg) {
  __shared__ C<T> tab[kRun * SFAMD_MAX_CUTOFF];
  __shared__ double sh[8];
  stage_tables(table, r, tab);
  const int64_t R = r.mixed ? (int64_t)r.D * r.D : r.D;
  const int64_t cpr = (R + 3) / 4, rows = r.N / R;
  double acc[kRun];
#pragma unroll
  for (int g = 0; g < kRun; ++g) acc[g] = 0.0;
  const bool vec = R % 4 == 0;
  GRID_STRIDE(w, rows * cpr) {
    const int64_t row = w / cpr;
    const int c0 = (int)(w - row * cpr) * 4;
    const int valid = R - c0 < 4 ? (int)(R - c0) : 4;
    RowDigits d;
    row_digits(r, row, d);
    const int64_t e0 = row * R + c0;
    C<T> x[4], l[4];
    load4(psi + e0, valid, vec, x);
    load4(lo_ + e0, valid, vec, l);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= valid) continue;  // padding of a short row: no table lookups past the run
      const int c = c0 + j;
      const int lb = r.mixed ? c / r.D : c, lk = r.mixed ? c % r.D : 0;
      const C<T> f = run_phase(r, tab, d, lb, lk);
      const C<T> o = cmul(f, x[j]);
      const double im = (double)l[j].re * o.im - (double)l[j].im * o.re;  // Im(conj(l) o)
#pragma unroll
      for (int g = 0; g < kRun; ++g) {
        if (g < r.ngates) {
          const bool last = r.mode[g] == r.n - 1;
          const int b = last ? lb : d.b[g], k = last ? lk : d.k[g];
          const int wb = r.kind[g] == SFAMD_OP_KERR ? b * b : b;
          const int wk = r.kind[g] == SFAMD_OP_KERR ? k * k : k;
          acc[g] -= (double)(r.mixed ? wb - wk : wb) * im;
        }
      }
      x[j] = cmul(cconj(f), l[j]);
    }
    store4(li + e0, valid, vec, x);
  }
  for (int g = 0; g < r.ngates; ++g) {
    const double s = block_sum(acc[g], sh);
    if (threadIdx.x == 0) {
      // partial layout [((op * 2 + pass) * 2 + slot) * nwg + workgroup]: pass 0 slot 0
      // gets the sum, the other three slots are cleared
      double* base = partial + (int64_t)r.op[g] * 4 * nwg + blockIdx.x;
      base[0] = s;
      base[(int64_t)nwg] = 0.0;
      base[2 * (int64_t)nwg] = 0.0;
      base[3 * (int64_t)nwg] = 0.0;
    }
  }
}

}  // namespace

namespace sfamd {

int diag_run_fwd(const void* in, void* out, const void* table, const DiagRun& r, int dtype, hipStream_t s) {
  const int64_t R = r.mixed ? (int64_t)r.D * r.D : r.D;
  const int grid = grid_for((r.N / R) * ((R + 3) / 4));
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_diag_run<float>, dim3(grid), dim3(kBlock), 0, s, (const C<float>*)in, (C<float>*)out,
                       (const C<float>*)table, r);
  else
    hipLaunchKernelGGL(k_diag_run<double>, dim3(grid), dim3(kBlock), 0, s, (const C<double>*)in, (C<double>*)out,
                       (const C<double>*)table, r);
  return (int)hipGetLastError();
}

int diag_run_adj(const void* psi, const void* lam_out, void* lam_in, const void* table, const DiagRun& r,
                 double* partial, int nwg, int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_diag_run_adj<float>, dim3(nwg), dim3(kBlock), 0, s, (const C<float>*)psi,
                       (const C<float>*)lam_out, (C<float>*)lam_in, (const C<float>*)table, r, partial, nwg);
  else
    hipLaunchKernelGGL(k_diag_run_adj<double>, dim3(nwg), dim3(kBlock), 0, s, (const C<double>*)psi,
                       (const C<double>*)lam_out, (C<double>*)lam_in, (const C<double>*)table, r, partial, nwg);
  return (int)hipGetLastError();
}

}  // namespace sfamd
