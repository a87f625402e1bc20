// Matrix-core building blocks shared by the complex64 gate kernels (gfx950).
//
// A complex operator (out = Op x, Op given entry by entry) is applied to 16
// fibers per wave with v_mfma_f32_16x16x4_f32 on its real form
//   [out_re; out_im] = [Re Op, -Im Op; Im Op, Re Op] [x_re; x_im],
// padded to 16 MB rows and columns.  The instruction's operand layouts fix the
// register plan: lane l = 16 g + c supplies A[c][g] and B[g][c] and receives
// C[4 g + j][c].  Columns c are fibers, so a lane touches amplitudes of one
// fiber only.  Each lane group g holds Q = 2 MB slots of input and output;
// for an operator of size n, group g owns amplitudes g P .. g P + P - 1 in
// slots 0 .. P - 1 (P = ceil(n / 4) <= Q), and slots P .. Q - 1 are padding
// (zero rows and columns):
//   row 16 mb + c      <->  slot 2 mb + (c % 4) / 2 of group c / 4, part c % 2
//   k-step s, group g  <->  slot s / 2 of group g, part s % 2,
// slot t of group g  <->  amplitude g P + t (t < P).  Spreading the amplitudes
// evenly over the four groups leaves at most three lanes' slots without an
// amplitude (instead of 4 Q - n), so almost every load and store carries one.
#pragma once

#include "common.h"

namespace sfamd {
namespace mf {

typedef float f4 __attribute__((ext_vector_type(4)));
using C = cpx<float>;


// Register plan for an operator of up to 8 MB complex rows and columns.
template <int MB>
struct Blk {
  static constexpr int Q = 2 * MB;   // slots per lane group
  static constexpr int NS = 4 * MB;  // k-steps (4 real columns each)
};

// amplitudes per lane group for an operator of size n
__host__ __device__ constexpr int group_amps(int n) { return (n + 3) / 4; }

// this lane's A operands of the real form of a complex operator of size n
// given by entry(i, k) (row i = output amplitude, column k = input amplitude;
// zero outside the operator), P = group_amps(n): A[mb][s]
template <int MB, typename F>
__device__ __forceinline__ void load_op(F entry, float (&A)[MB][Blk<MB>::NS], int P) {
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int s = 0; s < Blk<MB>::NS; ++s) {
      const int a = c & 1, b = s & 1;
      const int tr = 2 * mb + ((c & 3) >> 1), tc = s >> 1;
      const C m = entry((c >> 2) * P + tr, g * P + tc);
      const float v = a == b ? m.re : (a ? m.im : -m.im);
      A[mb][s] = (tr < P && tc < P) ? v : 0.f;
    }
}

// entry of op(M) for an n x n row-major matrix; op: 0 -> M, 1 -> conj(M),
// 2 -> M^H, 3 -> M^T; zero outside the matrix.  The load is unconditional
// (indices clamped into the matrix) and the zero a select after it: with a
// branch around each load, the compiler waits for every load of an unrolled
// operator build before issuing the next.
__device__ __forceinline__ C op_entry(const C* __restrict__ M, int n, int op, int i, int k) {
  const int ic = i < n ? i : n - 1, kc = k < n ? k : n - 1;
  C m = op >= 2 ? M[kc * n + ic] : M[ic * n + kc];
  if (op == 1 || op == 2) m.im = -m.im;
  return (i < n && k < n) ? m : mk<float>(0.f, 0.f);
}

// amplitude index a clamped to the last of n (for loads that are always
// issued; the value is replaced by zero when the slot holds no amplitude)
__device__ __forceinline__ int clamp_amp(int a, int n) { return a < n ? a : n - 1; }

// acc = op x for this lane's 16-fiber block (x in group layout).  Only the
// first 2 P k-steps carry amplitudes (slots P .. Q - 1 are zero columns), so
// the product stops there: P = group_amps(n).
template <int MB>
__device__ __forceinline__ void apply_op(const float (&A)[MB][Blk<MB>::NS], const C (&x)[Blk<MB>::Q], f4 (&acc)[MB],
                                         int P) {
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) acc[mb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < Blk<MB>::NS; ++s) {
    if (s >= 2 * P) continue;
    const float b = (s & 1) ? x[s >> 1].im : x[s >> 1].re;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mb][s], b, acc[mb], 0, 0, 0);
  }
}

// Operators shared by every wave of a workgroup (adjoint kernels) live in LDS
// in lane order, dst[(mb * NS + s) * 64 + lane]; the block's threads build them
// together and the A operand of each MFMA is one conflict-free ds_read.  A
// thread computes its entries in batches of kOpBatch (every entry of a batch
// is read, the element index clamped to the last) before storing them, so
// that their loads are in flight together (8 measured slower in the
// beamsplitter adjoint at cutoff 20: 1148 -> 1413 us).
constexpr int kOpBatch = 4;
template <int MB, typename F>
__device__ __forceinline__ void build_op_lds(F entry, float* dst, int P) {
  constexpr int NS = Blk<MB>::NS, NE = MB * NS * 64;
  for (int e0 = 0; e0 < NE; e0 += kOpBatch * (int)blockDim.x) {
    float v[kOpBatch];
#pragma unroll
    for (int u = 0; u < kOpBatch; ++u) {
      const int e = e0 + (int)threadIdx.x + u * (int)blockDim.x;
      const int ec = e < NE ? e : NE - 1;
      const int lane = ec & 63, slot = ec >> 6;
      const int mb = slot / NS, s = slot - mb * NS;
      const int c = lane & 15, g = lane >> 4;
      const int a = c & 1, b = s & 1;
      const int tr = 2 * mb + ((c & 3) >> 1), tc = s >> 1;
      const C m = entry((c >> 2) * P + tr, g * P + tc);
      const float x = a == b ? m.re : (a ? m.im : -m.im);
      v[u] = (tr < P && tc < P) ? x : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kOpBatch; ++u) {
      const int e = e0 + (int)threadIdx.x + u * (int)blockDim.x;
      if (e < NE) dst[e] = v[u];
    }
  }
}

template <int MB>
__device__ __forceinline__ void apply_op_lds(const float* opl, const C (&x)[Blk<MB>::Q], f4 (&acc)[MB], int P) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int mb = 0; mb < MB; ++mb) acc[mb] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < Blk<MB>::NS; ++s) {
    if (s >= 2 * P) continue;
    const float b = (s & 1) ? x[s >> 1].im : x[s >> 1].re;
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
      acc[mb] = __builtin_amdgcn_mfma_f32_16x16x4f32(opl[(mb * Blk<MB>::NS + s) * 64 + lane], b, acc[mb], 0, 0, 0);
  }
}

// slot q of this lane's group from the accumulators
template <int MB>
__device__ __forceinline__ C acc_at(const f4 (&acc)[MB], int q) {
  const f4 v = acc[q >> 1];
  return (q & 1) ? mk<float>(v[2], v[3]) : mk<float>(v[0], v[1]);
}

}  // namespace mf
}  // namespace sfamd
