// Dense single-mode gates on matrix cores (complex64, gfx950 MFMA).
//
// A dense gate on one axis of a state is a small complex GEMM,
// out[f, :] = U' x[f, :] over the fibers f of that axis (fockbackend's
// apply_gate_BLAS does np.dot(U, fiber) per slice, fockbackend/circuit.py:177,
// 208-210).  Here one wave applies U' to 16 fibers at a time with
// v_mfma_f32_16x16x4_f32 on the real form of U',
//
//   [out_re]   [ Re U'  -Im U' ] [x_re]
//   [out_im] = [ Im U'   Re U' ] [x_im],
//
// whose 2D x 2D real matrix is padded to Mp = 16 * Mb rows and columns
// (Mb = ceil(2D / 16)).  The operand layouts of the 16x16x4 instruction fix
// the register plan: lane l = 16 g + c (group g = 0..3, column c = 0..15)
// supplies A[c][g] and B[g][c] and receives C[4 g + j][c], j = 0..3.  Columns
// c are fibers, so each lane loads and stores amplitudes of one fiber only:
// group g owns amplitudes k = g Q .. g Q + Q - 1 (Q = 2 Mb) for input and
// output alike.  The real rows and columns are ordered so that this holds
// without any data exchange:
//   row   16 mb + c  <->  amplitude (c / 4) Q + 2 mb + (c % 4) / 2, part c % 2
//   k-step s, group g  <->  amplitude g Q + s / 2, part s % 2.
// The gate's real matrix lives in Mb * 4 Mb registers per lane (36 at D = 20)
// for the whole launch.  Any axis stride works: 16 consecutive fibers are 16
// consecutive inner positions (coalesced runs of 16 amplitudes) when the
// stride is >= 16, and 16 whole contiguous fibers when the axis is innermost.
#include "common.h"
#include "mfma.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
using C = cpx<float>;

template <int D>
struct Plan {
  static constexpr int Mb = (2 * D + 15) / 16;  // 16-row blocks
  static constexpr int Q = 2 * Mb;              // amplitudes per lane group
  static constexpr int NS = 4 * Mb;             // k-steps (4 real columns each)
};

// real-form entry of op(M) at row (i, a), column (k, b); a, b: 0 = re, 1 = im.
// op: 0 -> M, 1 -> conj(M), 2 -> M^H, 3 -> M^T
__device__ __forceinline__ float real_entry(const C* __restrict__ M, int D, int op, int i, int a, int k, int b) {
  if (i >= D || k >= D) return 0.f;
  const C m = op >= 2 ? M[k * D + i] : M[i * D + k];
  const float im = (op == 1 || op == 2) ? -m.im : m.im;
  if (a == b) return m.re;
  return a ? im : -im;
}

// this lane's A operands of op(M): A[mb][s]
template <int D>
__device__ __forceinline__ void load_operator(const C* __restrict__ M, int op, float (&A)[Plan<D>::Mb][Plan<D>::NS]) {
  constexpr int Q = Plan<D>::Q;
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < Plan<D>::Mb; ++mb)
#pragma unroll
    for (int s = 0; s < Plan<D>::NS; ++s)
      A[mb][s] = real_entry(M, D, op, (c >> 2) * Q + 2 * mb + ((c & 3) >> 1), c & 1, g * Q + (s >> 1), s & 1);
}

// acc = op(M) x for this lane's 16-fiber block (x in group layout)
template <int D>
__device__ __forceinline__ void apply_operator(const float (&A)[Plan<D>::Mb][Plan<D>::NS], const C (&x)[Plan<D>::Q],
                                               f4 (&acc)[Plan<D>::Mb]) {
#pragma unroll
  for (int mb = 0; mb < Plan<D>::Mb; ++mb) acc[mb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < Plan<D>::NS; ++s) {
    const float b = (s & 1) ? x[s >> 1].im : x[s >> 1].re;
#pragma unroll
    for (int mb = 0; mb < Plan<D>::Mb; ++mb)
      acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mb][s], b, acc[mb], 0, 0, 0);
  }
}

// amplitude q of this lane's group from the accumulators
template <int D>
__device__ __forceinline__ C acc_amp(const f4 (&acc)[Plan<D>::Mb], int q) {
  const f4 v = acc[q >> 1];
  return (q & 1) ? mk<float>(v[2], v[3]) : mk<float>(v[0], v[1]);
}

// fiber f of an axis with stride R (R = 1: innermost) -> offset of amplitude 0
__device__ __forceinline__ int64_t fiber_base(int64_t f, int64_t R, int D) {
  if (R == 1) return f * D;
  const int64_t o = f / R;
  return o * D * R + (f - o * R);
}

template <int D>
__global__ void __launch_bounds__(256)
k_axis_mfma(const C* __restrict__ in, C* __restrict__ out, int64_t NF, int64_t R, const C* __restrict__ M, int op) {
  constexpr int Q = Plan<D>::Q;
  float A[Plan<D>::Mb][Plan<D>::NS];
  load_operator<D>(M, op, A);
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  const int64_t nblk = (NF + 15) / 16;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t blk = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); blk < nblk; blk += nwave) {
    const int64_t f = blk * 16 + c;
    const bool valid = f < NF;
    const int64_t base = fiber_base(valid ? f : 0, R, D) + (int64_t)(g * Q) * R;
    C x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) x[q] = (valid && g * Q + q < D) ? in[base + q * R] : mk<float>(0.f, 0.f);
    f4 acc[Plan<D>::Mb];
    apply_operator<D>(A, x, acc);
    if (valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (g * Q + q < D) out[base + q * R] = acc_amp<D>(acc, q);
    }
  }
}

// adjoint: lam_in = U'^H lam_out;  g_p += Re sum_i conj(lam_out[i]) (dU'_p x)[i]
template <int D>
__global__ void __launch_bounds__(256)
k_axis_mfma_adj(const C* __restrict__ psi, const C* __restrict__ lo, C* __restrict__ li, int64_t NF, int64_t R,
                const C* __restrict__ M, const C* __restrict__ dM0, const C* __restrict__ dM1, int conj,
                double* __restrict__ partial, int nwg, int zp1) {
  constexpr int Q = Plan<D>::Q;
  float AH[Plan<D>::Mb][Plan<D>::NS], A0[Plan<D>::Mb][Plan<D>::NS], A1[Plan<D>::Mb][Plan<D>::NS];
  load_operator<D>(M, conj ? 3 : 2, AH);
  load_operator<D>(dM0, conj ? 1 : 0, A0);
  load_operator<D>(dM1, conj ? 1 : 0, A1);
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  const int64_t nblk = (NF + 15) / 16;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  double g0 = 0.0, g1 = 0.0;
  for (int64_t blk = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); blk < nblk; blk += nwave) {
    const int64_t f = blk * 16 + c;
    const bool valid = f < NF;
    const int64_t base = fiber_base(valid ? f : 0, R, D) + (int64_t)(g * Q) * R;
    C x[Q], l[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const bool in_range = valid && g * Q + q < D;
      x[q] = in_range ? psi[base + q * R] : mk<float>(0.f, 0.f);
      l[q] = in_range ? lo[base + q * R] : mk<float>(0.f, 0.f);
    }
    f4 acc[Plan<D>::Mb];
    apply_operator<D>(AH, l, acc);
    if (valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (g * Q + q < D) li[base + q * R] = acc_amp<D>(acc, q);
    }
    float s0 = 0.f, s1 = 0.f;
    apply_operator<D>(A0, x, acc);
#pragma unroll
    for (int q = 0; q < Q; ++q) s0 += re_conj_mul(l[q], acc_amp<D>(acc, q));
    apply_operator<D>(A1, x, acc);
#pragma unroll
    for (int q = 0; q < Q; ++q) s1 += re_conj_mul(l[q], acc_amp<D>(acc, q));
    g0 += (double)s0;
    g1 += (double)s1;
  }
  __shared__ double sh[8];
  const double t0 = block_sum(g0, sh);
  const double t1 = block_sum(g1, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = t0;
    partial[nwg + blockIdx.x] = t1;
    if (zp1) {
      partial[2 * nwg + blockIdx.x] = 0.0;
      partial[3 * nwg + blockIdx.x] = 0.0;
    }
  }
}

#define SFAMD_MFMA_D(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)

}  // namespace

namespace sfamd {

bool mfma_supported(int D, int dtype) { return dtype == SFAMD_C64 && D >= 2 && D <= kMfmaMaxD; }

int axis_dense_mfma(const void* in, void* out, int64_t N, int D, int64_t R, const void* M, int conj, hipStream_t s) {
  const int64_t NF = N / D;
  const int grid = grid_for((NF + 15) / 16 * 64);
  switch (D) {
#define CASE(d)                                                                                               \
  case d:                                                                                                     \
    hipLaunchKernelGGL((k_axis_mfma<d>), dim3(grid), dim3(kBlock), 0, s, (const C*)in, (C*)out, NF, R,       \
                       (const C*)M, conj ? 1 : 0);                                                            \
    break;
    SFAMD_MFMA_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

int axis_dense_adj_mfma(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t R,
                        const void* M, const void* dM0, const void* dM1, int conj, double* partial, int nwg, int zp1,
                        hipStream_t s) {
  const int64_t NF = N / D;
  switch (D) {
#define CASE(d)                                                                                                    \
  case d:                                                                                                          \
    hipLaunchKernelGGL((k_axis_mfma_adj<d>), dim3(nwg), dim3(kBlock), 0, s, (const C*)psi, (const C*)lam_out,     \
                       (C*)lam_in, NF, R, (const C*)M, (const C*)dM0, (const C*)dM1, conj, partial, nwg, zp1);    \
    break;
    SFAMD_MFMA_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

}  // namespace sfamd
