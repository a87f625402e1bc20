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
// group g owns amplitudes k = g P .. g P + P - 1 (P = ceil(D / 4)) in its
// Q = 2 Mb slots for input and output alike.  The real rows and columns are
// ordered so that this holds without any data exchange (mfma_ops.h):
//   row   16 mb + c  <->  slot 2 mb + (c % 4) / 2 of group c / 4, part c % 2
//   k-step s, group g  <->  slot s / 2 of group g, part s % 2.
// The gate's real matrix lives in Mb * 4 Mb registers per lane (36 at D = 20)
// for the whole launch.  Any axis stride works: 16 consecutive fibers are 16
// consecutive inner positions (coalesced runs of 16 amplitudes) when the
// stride is >= 16, and 16 whole contiguous fibers when the axis is innermost.
#include "common.h"
#include "mfma.h"
#include "mfma_ops.h"

namespace {

using namespace sfamd::mf;

template <int D>
struct Plan {
  static constexpr int Mb = (2 * D + 15) / 16;  // 16-row blocks
  static constexpr int Q = Blk<Mb>::Q;
  static constexpr int NS = Blk<Mb>::NS;
  static constexpr int P = group_amps(D);  // amplitudes per lane group
};

template <int D>
__device__ __forceinline__ void load_operator(const C* __restrict__ M, int op, float (&A)[Plan<D>::Mb][Plan<D>::NS]) {
  load_op<Plan<D>::Mb>([&](int i, int k) { return op_entry(M, D, op, i, k); }, A, Plan<D>::P);
}

template <int D>
__device__ __forceinline__ void apply_operator(const float (&A)[Plan<D>::Mb][Plan<D>::NS], const C (&x)[Plan<D>::Q],
                                               f4 (&acc)[Plan<D>::Mb]) {
  apply_op<Plan<D>::Mb>(A, x, acc, Plan<D>::P);
}

template <int D>
__device__ __forceinline__ C acc_amp(const f4 (&acc)[Plan<D>::Mb], int q) {
  return acc_at<Plan<D>::Mb>(acc, q);
}

// fiber f of an axis with stride R (R = 1: innermost) -> offset of amplitude 0
__device__ __forceinline__ int64_t fiber_base(int64_t f, int64_t R, int D) {
  if (R == 1) return f * D;
  const int64_t o = f / R;
  return o * D * R + (f - o * R);
}

// Innermost axis (INNER: stride 1, a whole number of 16-fiber chains): the
// 16 fibers of a chain are 16 D contiguous amplitudes, but a lane's slots of
// one fiber are D amplitudes away from the next lane's, so each 8-byte load
// instruction would touch 16 cache lines.  Instead each wave copies its chain
// through its own LDS buffer with coalesced 16-byte accesses (64 lanes read
// 1 KB contiguous per instruction) and the lanes take their slots from there;
// results go back the same way.  LDS accesses of one wave complete in order,
// so the buffer needs no barrier, only the compiler's ordering.
template <int D>
__device__ __forceinline__ void chain_in(const C* __restrict__ src, C* buf) {
  constexpr int NV = 8 * D, IT = (NV + 63) / 64;  // 16-byte vectors of 16 D amplitudes
  const int lane = threadIdx.x & 63;
  f4 v[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int j = lane + 64 * i;
    v[i] = reinterpret_cast<const f4*>(src)[j < NV ? j : NV - 1];
  }
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int j = lane + 64 * i;
    if (j < NV) reinterpret_cast<f4*>(buf)[j] = v[i];
  }
  __asm__ volatile("" ::: "memory");
}

template <int D>
__device__ __forceinline__ void chain_out(const C* buf, C* __restrict__ dst) {
  constexpr int NV = 8 * D, IT = (NV + 63) / 64;
  const int lane = threadIdx.x & 63;
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int j = lane + 64 * i;
    if (j < NV) reinterpret_cast<f4*>(dst)[j] = reinterpret_cast<const f4*>(buf)[j];
  }
  __asm__ volatile("" ::: "memory");
}

template <int D, bool INNER>
__global__ void __launch_bounds__(256)
k_axis_mfma(const C* __restrict__ in, C* __restrict__ out, int64_t NF, int64_t R, const C* __restrict__ M, int op) {
  constexpr int Q = Plan<D>::Q, P = Plan<D>::P, CH = 16 * D;
  __shared__ f4 stage_v[INNER ? 4 * CH / 2 : 1];
  C* buf = reinterpret_cast<C*>(stage_v) + (INNER ? (threadIdx.x >> 6) * CH : 0);
  float A[Plan<D>::Mb][Plan<D>::NS];
  load_operator<D>(M, op, A);
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  const int64_t nblk = (NF + 15) / 16;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t blk = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); blk < nblk; blk += nwave) {
    const int64_t f = blk * 16 + c;
    const bool valid = f < NF;
    const int64_t fb = fiber_base(valid ? f : 0, R, D);
    C x[Q];
    if constexpr (INNER) chain_in<D>(in + blk * CH, buf);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      x[q] = mk<float>(0.f, 0.f);
      if (q >= P) continue;
      const int a = g * P + q;
      const C v = INNER ? buf[c * D + clamp_amp(a, D)] : in[fb + (int64_t)clamp_amp(a, D) * R];
      if (valid && a < D) x[q] = v;
    }
    f4 acc[Plan<D>::Mb];
    apply_operator<D>(A, x, acc);
    if constexpr (INNER) {
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (g * P + q < D) buf[c * D + g * P + q] = acc_amp<D>(acc, q);
      chain_out<D>(buf, out + blk * CH);
    } else if (valid) {
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (g * P + q < D) out[fb + (int64_t)(g * P + q) * R] = acc_amp<D>(acc, q);
    }
  }
}

// adjoint: lam_in = U'^H lam_out;  g_p += Re sum_i conj(lam_out[i]) (dU'_p x)[i]
template <int D, bool INNER>
__global__ void __launch_bounds__(256)
k_axis_mfma_adj(const C* __restrict__ psi, const C* __restrict__ lo, C* __restrict__ li, int64_t NF, int64_t R,
                const C* __restrict__ M, const C* __restrict__ dM0, const C* __restrict__ dM1, int conj,
                double* __restrict__ partial, int nwg, int zp1) {
  constexpr int MB = Plan<D>::Mb, Q = Plan<D>::Q, P = Plan<D>::P, OPN = MB * Plan<D>::NS * 64, CH = 16 * D;
  __shared__ float opl[3 * OPN];
  __shared__ f4 stage_v[INNER ? 2 * 4 * CH / 2 : 1];  // per wave: the chain of psi, then of lam_out
  C* bx = reinterpret_cast<C*>(stage_v) + (INNER ? 2 * (threadIdx.x >> 6) * CH : 0);
  C* bl = bx + (INNER ? CH : 0);
  build_op_lds<MB>([&](int i, int k) { return op_entry(M, D, conj ? 3 : 2, i, k); }, opl, P);
  build_op_lds<MB>([&](int i, int k) { return op_entry(dM0, D, conj ? 1 : 0, i, k); }, opl + OPN, P);
  build_op_lds<MB>([&](int i, int k) { return op_entry(dM1, D, conj ? 1 : 0, i, k); }, opl + 2 * OPN, P);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  const int64_t nblk = (NF + 15) / 16;
  const int64_t nwave = (int64_t)gridDim.x * (blockDim.x >> 6);
  double g0 = 0.0, g1 = 0.0;
  for (int64_t blk = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); blk < nblk; blk += nwave) {
    // keep the operators in LDS: without this barrier the compiler hoists
    // their loop-invariant reads into (three operators' worth of) registers
    __asm__ volatile("" ::: "memory");
    const int64_t f = blk * 16 + c;
    const bool valid = f < NF;
    const int64_t fb = fiber_base(valid ? f : 0, R, D);
    C x[Q], l[Q];
    if constexpr (INNER) {
      chain_in<D>(psi + blk * CH, bx);
      chain_in<D>(lo + blk * CH, bl);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      x[q] = l[q] = mk<float>(0.f, 0.f);
      if (q >= P) continue;
      const int a = g * P + q;
      const int64_t o = fb + (int64_t)clamp_amp(a, D) * R;
      const C xv = INNER ? bx[c * D + clamp_amp(a, D)] : psi[o];
      const C lv = INNER ? bl[c * D + clamp_amp(a, D)] : lo[o];
      if (valid && a < D) {
        x[q] = xv;
        l[q] = lv;
      }
    }
    f4 acc[MB];
    apply_op_lds<MB>(opl, l, acc, P);
    if constexpr (INNER) {
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (g * P + q < D) bl[c * D + g * P + q] = acc_at<MB>(acc, q);
      chain_out<D>(bl, li + blk * CH);
    } else if (valid) {
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (g * P + q < D) li[fb + (int64_t)(g * P + q) * R] = acc_at<MB>(acc, q);
    }
    float s0 = 0.f, s1 = 0.f;
    apply_op_lds<MB>(opl + OPN, x, acc, P);
#pragma unroll
    for (int q = 0; q < Q; ++q) s0 += re_conj_mul(l[q], acc_at<MB>(acc, q));
    apply_op_lds<MB>(opl + 2 * OPN, x, acc, P);
#pragma unroll
    for (int q = 0; q < Q; ++q) s1 += re_conj_mul(l[q], acc_at<MB>(acc, q));
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

// ------------------------------------------------------------------ beamsplitter
// The photon-number block s of a beamsplitter (c = c(s) <= D amplitudes,
// fockbackend/circuit.py:306-335) is a c x c complex operator applied to the
// fibers (N, s - N) of the two gate axes.  Work is split into tasks (s, chunk
// of kBsChunk fibers): s is uniform over the workgroup, each wave builds the
// block's real operator once per task (Mb = ceil(2c / 16) in 1..3) and applies
// it to 16 fibers per MFMA chain.  Fibers of a chunk are consecutive inner
// positions of the non-gate axes (coalesced when that run is >= 16).
constexpr int kBsChunk = 1024;  // 512 or 2048: c5 5.59 -> 5.52 grad-evals/s

struct PairGeo {
  int64_t B, C;        // extents after the inner gate axis, between the two gate axes
  int64_t s_lo, s_hi;  // strides of the outer and inner gate axes
};

__device__ __forceinline__ int64_t pair_base(int64_t f, const PairGeo& g, int D) {
  const int64_t b = f % g.B;
  const int64_t t = f / g.B;
  const int64_t c = t % g.C;
  const int64_t a = t / g.C;
  return a * (D * g.s_lo) + c * (D * g.s_hi) + b;
}

__device__ __forceinline__ void mfma_write_partials(double g0, double g1, double* partial, int nwg, int zp1) {
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

// Fibers of block s.  Photon-number blocks (BS, MZ, CK: n1 + n2 = s) hold the
// amplitudes (k, s - k); difference blocks (S2, diff = 1: n1 - n2 = d,
// d = s - (D - 1)) hold (k, k - d), for k = lo .. lo + c - 1.  Offset of the
// first amplitude from the fiber base, and the step between amplitudes:
__device__ __forceinline__ int64_t block_first(int64_t sp, int64_t sq, int lo, int s, int D, int diff) {
  return lo * sp + (int64_t)(diff ? lo - (s - (D - 1)) : s - lo) * sq;
}
__device__ __forceinline__ int64_t block_step(int64_t sp, int64_t sq, int diff) { return diff ? sp + sq : sp - sq; }

// A block of c amplitudes has P = group_amps(c) = Q - 1 or Q amplitudes per
// lane group (Q = 2 MB, MB = ceil(2c / 16)); the tasks are instantiated for
// both, so that the operator products stop after 2 P k-steps without a
// branch per k-step.
template <int MB>
__device__ __forceinline__ bool full_groups(int c) {
  return group_amps(c) == 2 * MB;
}

// one task of a beamsplitter pass: block s (c amplitudes from N = lo) on the
// fibers f0 .. f0 + kBsChunk - 1
template <int MB, int P, int D>
__device__ __forceinline__ void bs_task(const C* __restrict__ in, C* __restrict__ out, int64_t NF, const PairGeo& geo,
                                        int64_t sp, int64_t sq, const C* __restrict__ bb, int c, int lo, int s,
                                        int conj, int diff, int64_t f0) {
  constexpr int Q = Blk<MB>::Q;
  float A[MB][Blk<MB>::NS];
  load_op<MB>([&](int i, int k) { return op_entry(bb, c, conj ? 1 : 0, i, k); }, A, P);
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15, g = lane >> 4;
  const int64_t step = block_step(sp, sq, diff);
  const int64_t first = block_first(sp, sq, lo, s, D, diff);
  const int64_t left = NF - f0;
  const int nf = left < kBsChunk ? (int)left : kBsChunk;
  for (int b16 = threadIdx.x >> 6; b16 * 16 < nf; b16 += blockDim.x >> 6) {
    const bool valid = b16 * 16 + cl < nf;
    const int64_t f = f0 + (valid ? b16 * 16 + cl : 0);
    const int64_t fb = pair_base(f, geo, D) + first;
    C x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int a = g * P + q;
      const C v = in[fb + (int64_t)clamp_amp(a, c) * step];
      x[q] = (valid && q < P && a < c) ? v : mk<float>(0.f, 0.f);
    }
    f4 acc[MB];
    apply_op<MB>(A, x, acc, P);
    if (valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q < P && g * P + q < c) out[fb + (int64_t)(g * P + q) * step] = acc_at<MB>(acc, q);
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256)
k_bs_mfma(const C* __restrict__ in, C* __restrict__ out, int64_t NF, PairGeo geo, int64_t sp, int64_t sq,
          const C* __restrict__ blk, int conj, int diff) {
  constexpr int NSB = 2 * D - 1;
  const int64_t ntask = (NF + kBsChunk - 1) / kBsChunk * NSB;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int s = (int)(task % NSB);
    const int64_t f0 = task / NSB * kBsChunk;
    const int lo = bs_lo(s, D);
    const int c = bs_hi(s, D) - lo + 1;
    const C* bb = blk + bs_block_off(s, D);
    const int mb = (2 * c + 15) / 16;
#define SFAMD_BS_TASK(MB)                                                        \
  if (full_groups<MB>(c))                                                        \
    bs_task<MB, 2 * MB, D>(in, out, NF, geo, sp, sq, bb, c, lo, s, conj, diff, f0); \
  else                                                                           \
    bs_task<MB, 2 * MB - 1, D>(in, out, NF, geo, sp, sq, bb, c, lo, s, conj, diff, f0);
    if (mb == 1) {
      SFAMD_BS_TASK(1)
    } else if constexpr (D > 8) {
      if (mb == 2) {
        SFAMD_BS_TASK(2)
      } else if constexpr (D > 16) {
        SFAMD_BS_TASK(3)
      }
    }
#undef SFAMD_BS_TASK
  }
}

// adjoint of one task: lam_in = B'^H lam_out; gradients through dB'/dp0
// (table) and dB'/dp1 = i sg (wa n + wb N) B'[N][n] (sg = -1 on the conjugated
// side).  opl: the workgroup's LDS for three operators of MB blocks.
template <int MB, int P, int D>
__device__ __forceinline__ void bs_adj_task(const C* __restrict__ psi, const C* __restrict__ lo_, C* __restrict__ li,
                                            int64_t NF, const PairGeo& geo, int64_t sp, int64_t sq,
                                            const C* __restrict__ bb, const C* __restrict__ db, int c, int lo, int s,
                                            int conj, int wa, int wb, int diff, int64_t f0, float* opl, double& g0,
                                            double& g1) {
  constexpr int Q = Blk<MB>::Q, OPN = MB * Blk<MB>::NS * 64;
  __syncthreads();  // the previous task's operators are no longer read
  build_op_lds<MB>([&](int i, int k) { return op_entry(bb, c, conj ? 3 : 2, i, k); }, opl, P);
  build_op_lds<MB>([&](int i, int k) { return op_entry(db, c, conj ? 1 : 0, i, k); }, opl + OPN, P);
  const float sg = conj ? -1.f : 1.f;
  build_op_lds<MB>(
      [&](int i, int k) {
        const C m = op_entry(bb, c, conj ? 1 : 0, i, k);
        const float w = sg * (float)(wa * (lo + k) + wb * (lo + i));
        return mk<float>(-w * m.im, w * m.re);
      },
      opl + 2 * OPN, P);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15, g = lane >> 4;
  const int64_t step = block_step(sp, sq, diff);
  const int64_t first = block_first(sp, sq, lo, s, D, diff);
  const int64_t left = NF - f0;
  const int nf = left < kBsChunk ? (int)left : kBsChunk;
  for (int b16 = threadIdx.x >> 6; b16 * 16 < nf; b16 += blockDim.x >> 6) {
    __asm__ volatile("" ::: "memory");  // operators stay in LDS (see k_axis_mfma_adj)
    const bool valid = b16 * 16 + cl < nf;
    const int64_t f = f0 + (valid ? b16 * 16 + cl : 0);
    const int64_t fb = pair_base(f, geo, D) + first;
    C x[Q], l[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int a = g * P + q;
      const bool in_range = valid && q < P && a < c;
      const int64_t o = fb + (int64_t)clamp_amp(a, c) * step;
      const C xv = psi[o], lv = lo_[o];
      x[q] = in_range ? xv : mk<float>(0.f, 0.f);
      l[q] = in_range ? lv : mk<float>(0.f, 0.f);
    }
    f4 acc[MB];
    apply_op_lds<MB>(opl, l, acc, P);
    if (valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q < P && g * P + q < c) li[fb + (int64_t)(g * P + q) * step] = acc_at<MB>(acc, q);
    }
    float s0 = 0.f, s1 = 0.f;
    apply_op_lds<MB>(opl + OPN, x, acc, P);
#pragma unroll
    for (int q = 0; q < Q; ++q) s0 += re_conj_mul(l[q], acc_at<MB>(acc, q));
    apply_op_lds<MB>(opl + 2 * OPN, x, acc, P);
#pragma unroll
    for (int q = 0; q < Q; ++q) s1 += re_conj_mul(l[q], acc_at<MB>(acc, q));
    g0 += (double)s0;
    g1 += (double)s1;
  }
}

template <int D>
__global__ void __launch_bounds__(256)
k_bs_mfma_adj(const C* __restrict__ psi, const C* __restrict__ lo_, C* __restrict__ li, int64_t NF, PairGeo geo,
              int64_t sp, int64_t sq, const C* __restrict__ blk, const C* __restrict__ dblk, int conj, int wa,
              int wb, int diff, double* __restrict__ partial, int nwg, int zp1) {
  constexpr int NSB = 2 * D - 1, MBX = Plan<D>::Mb;
  __shared__ float opl[3 * MBX * Blk<MBX>::NS * 64];
  const int64_t ntask = (NF + kBsChunk - 1) / kBsChunk * NSB;
  double g0 = 0.0, g1 = 0.0;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int s = (int)(task % NSB);
    const int64_t f0 = task / NSB * kBsChunk;
    const int lo = bs_lo(s, D);
    const int c = bs_hi(s, D) - lo + 1;
    const int64_t off = bs_block_off(s, D);
    const int mb = (2 * c + 15) / 16;
#define SFAMD_BS_ADJ_TASK(MB)                                                                                      \
  if (full_groups<MB>(c))                                                                                          \
    bs_adj_task<MB, 2 * MB, D>(psi, lo_, li, NF, geo, sp, sq, blk + off, dblk + off, c, lo, s, conj, wa, wb, diff, \
                               f0, opl, g0, g1);                                                                   \
  else                                                                                                             \
    bs_adj_task<MB, 2 * MB - 1, D>(psi, lo_, li, NF, geo, sp, sq, blk + off, dblk + off, c, lo, s, conj, wa, wb,   \
                                   diff, f0, opl, g0, g1);
    if (mb == 1) {
      SFAMD_BS_ADJ_TASK(1)
    } else if constexpr (D > 8) {
      if (mb == 2) {
        SFAMD_BS_ADJ_TASK(2)
      } else if constexpr (D > 16) {
        SFAMD_BS_ADJ_TASK(3)
      }
    }
#undef SFAMD_BS_ADJ_TASK
  }
  mfma_write_partials(g0, g1, partial, nwg, zp1);
}

// ------------------------------------------------------------------ two-sided dense gate on a DM
// rho' = U rho U^H on the (bra, ket) axes of one mode (strides sb = D sk, sk)
// in one HBM pass (fockbackend applies U rho_blk U^H per slice,
// fockbackend/circuit.py:179-183,212-214).  A workgroup takes 16 "columns" tau
// of D x D blocks X_tau[i][j] and applies U along i (fibers (tau, j)), keeps
// Y = U X in LDS, then applies conj(U) along j (fibers (tau, i)).
//   sk >= 16: tau = 16 consecutive inner positions; both passes read and write
//             HBM directly in 128-byte runs, Y lives in LDS as [i][j][tau].
//   sk == 1:  tau = 16 consecutive D x D slabs (contiguous); the slabs are
//             staged in LDS ([tau][i][j], slab stride D*D + 1 against bank
//             conflicts) and both passes run in place.
constexpr int kStageCopy = 8;

template <int D, bool STAGED>
__global__ void __launch_bounds__(256)
k_dm_dense_mfma(const C* __restrict__ in, C* __restrict__ out, int64_t N, int64_t sk, const C* __restrict__ M) {
  constexpr int MB = Plan<D>::Mb, Q = Plan<D>::Q, P = Plan<D>::P;
  constexpr int PT = D * D + 1;
  __shared__ C sh[STAGED ? 16 * PT : D * D * 16];
  float AU[MB][Plan<D>::NS], AC[MB][Plan<D>::NS];
  load_operator<D>(M, 0, AU);
  load_operator<D>(M, 1, AC);
  const int lane = threadIdx.x & 63;
  const int c = lane & 15, g = lane >> 4;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int64_t sb = D * sk;
  const int64_t nchunk = STAGED ? 1 : (sk + 15) / 16;
  const int64_t nouter = N / ((int64_t)D * D * sk);
  const int64_t ntile = STAGED ? (nouter + 15) / 16 : nouter * nchunk;
  for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    if (STAGED) {
      const int64_t a0 = tile * 16;
      const int gs = (int)(nouter - a0 < 16 ? nouter - a0 : 16);
      const C* src = in + a0 * D * D;
      const int ne = gs * D * D;
      // batches of kStageCopy loads in flight per thread (index clamped, store guarded)
      for (int e0 = 0; e0 < ne; e0 += kStageCopy * (int)blockDim.x) {
        C v[kStageCopy];
#pragma unroll
        for (int u = 0; u < kStageCopy; ++u) {
          const int e = e0 + (int)threadIdx.x + u * (int)blockDim.x;
          v[u] = src[e < ne ? e : ne - 1];
        }
#pragma unroll
        for (int u = 0; u < kStageCopy; ++u) {
          const int e = e0 + (int)threadIdx.x + u * (int)blockDim.x;
          if (e < ne) sh[(e / (D * D)) * PT + e % (D * D)] = v[u];
        }
      }
      __syncthreads();
      // pass 1: fibers (tau, j), amplitudes i; pass 2: fibers (tau, i), amplitudes j
#pragma unroll
      for (int pass = 0; pass < 2; ++pass) {
        const int nblk = (gs * D + 15) / 16;
        for (int b = wave; b < nblk; b += nw) {
          const int phi = b * 16 + c;
          const bool valid = phi < gs * D;
          const int tau = valid ? phi / D : 0, r = valid ? phi % D : 0;
          const int base = tau * PT + (pass == 0 ? r : r * D);
          const int st = pass == 0 ? D : 1;
          C x[Q];
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            x[q] = mk<float>(0.f, 0.f);
            if (q >= P) continue;
            const C v = sh[base + clamp_amp(g * P + q, D) * st];
            if (valid && g * P + q < D) x[q] = v;
          }
          f4 acc[MB];
          if (pass == 0)
            apply_op<MB>(AU, x, acc, P);
          else
            apply_op<MB>(AC, x, acc, P);
          if (valid) {
#pragma unroll
            for (int q = 0; q < P; ++q)
              if (g * P + q < D) sh[base + (g * P + q) * st] = acc_at<MB>(acc, q);
          }
        }
        __syncthreads();
      }
      C* dst = out + a0 * D * D;
      for (int e = threadIdx.x; e < gs * D * D; e += blockDim.x) dst[e] = sh[(e / (D * D)) * PT + e % (D * D)];
      __syncthreads();
    } else {
      const int64_t a = tile / nchunk;
      const int64_t b0 = (tile - a * nchunk) * 16;
      const bool valid = b0 + c < sk;
      const int64_t col = a * D * sb + b0 + (valid ? c : 0);
      // pass 1: block j, amplitudes i from HBM -> Y[i][j][tau] in LDS
      for (int j = wave; j < D; j += nw) {
        C x[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          x[q] = mk<float>(0.f, 0.f);
          if (q >= P) continue;
          const C v = in[col + (int64_t)clamp_amp(g * P + q, D) * sb + j * sk];
          if (valid && g * P + q < D) x[q] = v;
        }
        f4 acc[MB];
        apply_op<MB>(AU, x, acc, P);
#pragma unroll
        for (int q = 0; q < P; ++q)
          if (g * P + q < D) sh[((g * P + q) * D + j) * 16 + c] = acc_at<MB>(acc, q);
      }
      __syncthreads();
      // pass 2: block i, amplitudes j from LDS -> HBM
      for (int i = wave; i < D; i += nw) {
        C x[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          x[q] = mk<float>(0.f, 0.f);
          if (q >= P) continue;
          const C v = sh[(i * D + clamp_amp(g * P + q, D)) * 16 + c];
          if (g * P + q < D) x[q] = v;
        }
        f4 acc[MB];
        apply_op<MB>(AC, x, acc, P);
        if (valid) {
#pragma unroll
          for (int q = 0; q < P; ++q)
            if (g * P + q < D) out[col + i * sb + (int64_t)(g * P + q) * sk] = acc_at<MB>(acc, q);
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ two-sided two-mode gate on a DM
// rho' = G rho G^H for a beamsplitter-family gate G on modes (m0, m1) of a DM
// in one HBM pass (the reference applies the gate to one index pair, then its
// conjugate to the other, in two passes: fockbackend/circuit.py:262-300,
// 309-335).  G is block diagonal: block s on the unconjugated axis pair
// (amplitudes k = 0 .. c_s - 1 of the block), block t on the conjugated pair
// (k').  Over 16 consecutive positions tau of the innermost non-gate run, the
// (s, t) part of rho is X[k][k'][tau], and rho'_(s,t) = G_s X G_t^H.  A
// workgroup takes one (s, t) pair, loads its two operators into registers
// once, and runs through a group of chunks of 16 positions: each chunk is
// staged in LDS as [k][k'][tau] with batched loads (a fiber loop reading HBM
// directly waits for each fiber's loads in turn), G_s is applied along k in
// place (fibers (tau, k')), then conj(G_t) along k' (fibers (tau, u)) with the
// results written to HBM.
struct DmPairGeo {
  int64_t ap, aq, bp, bq;  // strides of the first / second mode's axes: unconjugated (a), conjugated (b)
  int64_t R, SM, SO;       // innermost non-gate run; strides of the middle and outer non-gate indices
  int64_t M, nch, nchunk;  // middle extent; chunks of 16 per run; chunks in all
};

constexpr int kDmPairGroup = 8;  // chunks of 16 positions per task

template <int MBS, int MBT, int D>
__device__ __forceinline__ void dm_pair_task(const C* __restrict__ in, C* __restrict__ out, const DmPairGeo& geo,
                                             const C* __restrict__ blk, int s, int t, int diff, int64_t ch0,
                                             int64_t ch1, C* sh) {
  constexpr int QS = Blk<MBS>::Q, QT = Blk<MBT>::Q;
  const int lo_s = bs_lo(s, D), cs = bs_hi(s, D) - lo_s + 1;
  const int lo_t = bs_lo(t, D), ct = bs_hi(t, D) - lo_t + 1;
  const int ps = group_amps(cs), pt = group_amps(ct);
  float AS[MBS][Blk<MBS>::NS], AT[MBT][Blk<MBT>::NS];
  load_op<MBS>([&](int i, int k) { return op_entry(blk + bs_block_off(s, D), cs, 0, i, k); }, AS, ps);
  load_op<MBT>([&](int i, int k) { return op_entry(blk + bs_block_off(t, D), ct, 1, i, k); }, AT, pt);
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15, g = lane >> 4;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int64_t sa = block_step(geo.ap, geo.aq, diff), sb = block_step(geo.bp, geo.bq, diff);
  const int64_t first = block_first(geo.ap, geo.aq, lo_s, s, D, diff) + block_first(geo.bp, geo.bq, lo_t, t, D, diff);
  const int ne = cs * ct * 16;
  for (int64_t ch = ch0; ch < ch1; ++ch) {
    const int64_t om = ch / geo.nch;
    const int64_t t0 = (ch - om * geo.nch) * 16;
    const int64_t o = om / geo.M;
    const int64_t base = o * geo.SO + (om - o * geo.M) * geo.SM + t0 + first;
    // stage X[k][k'][tau] (LDS [k][k'][tau]) with kStageCopy loads in flight per
    // thread; positions past the run read position t0 (their results are not stored)
    for (int e0 = 0; e0 < ne; e0 += kStageCopy * (int)blockDim.x) {
      C v[kStageCopy];
#pragma unroll
      for (int u = 0; u < kStageCopy; ++u) {
        const int e = e0 + (int)threadIdx.x + u * (int)blockDim.x;
        const int ec = e < ne ? e : ne - 1;
        const int tau = ec & 15, kk = ec >> 4;
        const int k = kk / ct, kp = kk - k * ct;
        v[u] = in[base + (t0 + tau < geo.R ? tau : 0) + (int64_t)k * sa + (int64_t)kp * sb];
      }
#pragma unroll
      for (int u = 0; u < kStageCopy; ++u) {
        const int e = e0 + (int)threadIdx.x + u * (int)blockDim.x;
        if (e < ne) sh[e] = v[u];
      }
    }
    __syncthreads();
    // in place: Y[u][k'][tau] = sum_k G_s[u][k] X[k][k'][tau]  (fibers (tau, k'))
    for (int kb = wave; kb < ct; kb += nw) {
      C x[QS];
#pragma unroll
      for (int q = 0; q < QS; ++q) {
        const int a = g * ps + q;
        const C v = sh[(clamp_amp(a, cs) * ct + kb) * 16 + cl];
        x[q] = (q < ps && a < cs) ? v : mk<float>(0.f, 0.f);
      }
      f4 acc[MBS];
      apply_op<MBS>(AS, x, acc, QS);
#pragma unroll
      for (int q = 0; q < QS; ++q)
        if (q < ps && g * ps + q < cs) sh[((g * ps + q) * ct + kb) * 16 + cl] = acc_at<MBS>(acc, q);
    }
    __syncthreads();
    // rho'[u][u'][tau] = sum_k' conj(G_t[u'][k']) Y[u][k'][tau]  (fibers (tau, u), to HBM)
    const bool valid = t0 + cl < geo.R;
    for (int u = wave; u < cs; u += nw) {
      C x[QT];
#pragma unroll
      for (int q = 0; q < QT; ++q) {
        const int a = g * pt + q;
        const C v = sh[(u * ct + clamp_amp(a, ct)) * 16 + cl];
        x[q] = (q < pt && a < ct) ? v : mk<float>(0.f, 0.f);
      }
      f4 acc[MBT];
      apply_op<MBT>(AT, x, acc, QT);
      if (valid) {
#pragma unroll
        for (int q = 0; q < QT; ++q)
          if (q < pt && g * pt + q < ct)
            out[base + cl + (int64_t)u * sa + (int64_t)(g * pt + q) * sb] = acc_at<MBT>(acc, q);
      }
    }
    __syncthreads();
  }
}

template <int D>
__global__ void __launch_bounds__(256, 3)
k_dm_pair_mfma(const C* __restrict__ in, C* __restrict__ out, DmPairGeo geo, const C* __restrict__ blk, int diff) {
  constexpr int NSB = 2 * D - 1;
  __shared__ C sh[D * D * 16];
  const int64_t ngrp = (geo.nchunk + kDmPairGroup - 1) / kDmPairGroup;
  const int64_t ntask = (int64_t)NSB * NSB * ngrp;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int st = (int)(task % (NSB * NSB));
    const int64_t ch0 = task / (NSB * NSB) * kDmPairGroup;
    const int64_t ch1 = ch0 + kDmPairGroup < geo.nchunk ? ch0 + kDmPairGroup : geo.nchunk;
    const int s = st / NSB, t = st - s * NSB;
    const int mbs = (2 * (bs_hi(s, D) - bs_lo(s, D) + 1) + 15) / 16;
    const int mbt = (2 * (bs_hi(t, D) - bs_lo(t, D) + 1) + 15) / 16;
    // block sizes c <= D: mb <= Plan<D>::Mb, so only those variants exist
#define SFAMD_DM_PAIR_T(MS)                                                  \
  if (mbt == 1)                                                              \
    dm_pair_task<MS, 1, D>(in, out, geo, blk, s, t, diff, ch0, ch1, sh);     \
  else if constexpr (D > 8) {                                                \
    if (mbt == 2)                                                            \
      dm_pair_task<MS, 2, D>(in, out, geo, blk, s, t, diff, ch0, ch1, sh);   \
    else if constexpr (D > 16)                                               \
      dm_pair_task<MS, 3, D>(in, out, geo, blk, s, t, diff, ch0, ch1, sh);   \
  }
    if (mbs == 1) {
      SFAMD_DM_PAIR_T(1)
    } else if constexpr (D > 8) {
      if (mbs == 2) {
        SFAMD_DM_PAIR_T(2)
      } else if constexpr (D > 16) {
        SFAMD_DM_PAIR_T(3)
      }
    }
#undef SFAMD_DM_PAIR_T
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
  const bool inner = R == 1 && NF % 16 == 0;  // whole 16-fiber chains, contiguous
  switch (D) {
#define CASE(d)                                                                                               \
  case d:                                                                                                     \
    if (inner)                                                                                                \
      hipLaunchKernelGGL((k_axis_mfma<d, true>), dim3(grid), dim3(kBlock), 0, s, (const C*)in, (C*)out, NF, R, \
                         (const C*)M, conj ? 1 : 0);                                                          \
    else                                                                                                      \
      hipLaunchKernelGGL((k_axis_mfma<d, false>), dim3(grid), dim3(kBlock), 0, s, (const C*)in, (C*)out, NF,   \
                         R, (const C*)M, conj ? 1 : 0);                                                       \
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
  const bool inner = R == 1 && NF % 16 == 0;  // whole 16-fiber chains, contiguous
  switch (D) {
#define CASE(d)                                                                                                    \
  case d:                                                                                                          \
    if (inner)                                                                                                     \
      hipLaunchKernelGGL((k_axis_mfma_adj<d, true>), dim3(nwg), dim3(kBlock), 0, s, (const C*)psi,                \
                         (const C*)lam_out, (C*)lam_in, NF, R, (const C*)M, (const C*)dM0, (const C*)dM1, conj,  \
                         partial, nwg, zp1);                                                                     \
    else                                                                                                           \
      hipLaunchKernelGGL((k_axis_mfma_adj<d, false>), dim3(nwg), dim3(kBlock), 0, s, (const C*)psi,               \
                         (const C*)lam_out, (C*)lam_in, NF, R, (const C*)M, (const C*)dM0, (const C*)dM1, conj,  \
                         partial, nwg, zp1);                                                                     \
    break;
    SFAMD_MFMA_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

static PairGeo pair_geo(int64_t sp, int64_t sq, int D) {
  PairGeo g;
  g.s_lo = sp > sq ? sp : sq;
  g.s_hi = sp > sq ? sq : sp;
  g.B = g.s_hi;
  g.C = g.s_lo / (D * g.s_hi);
  return g;
}

int pair_bs_mfma(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
                 hipStream_t s, int diff) {
  const int64_t NF = N / ((int64_t)D * D);
  const PairGeo g = pair_geo(sp, sq, D);
  int64_t ntask = (NF + kBsChunk - 1) / kBsChunk * (2 * D - 1);
  const int grid = (int)(ntask < kMaxGrid ? ntask : kMaxGrid);
  switch (D) {
#define CASE(d)                                                                                                 \
  case d:                                                                                                       \
    hipLaunchKernelGGL((k_bs_mfma<d>), dim3(grid), dim3(kBlock), 0, s, (const C*)in, (C*)out, NF, g, sp, sq,   \
                       (const C*)blocks, conj, diff);                                                           \
    break;
    SFAMD_MFMA_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

int pair_bs_adj_mfma(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                     const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial, int nwg,
                     int zp1, hipStream_t s, int diff) {
  const int64_t NF = N / ((int64_t)D * D);
  const PairGeo g = pair_geo(sp, sq, D);
  switch (D) {
#define CASE(d)                                                                                                     \
  case d:                                                                                                           \
    hipLaunchKernelGGL((k_bs_mfma_adj<d>), dim3(nwg), dim3(kBlock), 0, s, (const C*)psi, (const C*)lam_out,        \
                       (C*)lam_in, NF, g, sp, sq, (const C*)blocks, (const C*)dblocks, conj, wa, wb, diff, partial, \
                       nwg, zp1);                                                                                  \
    break;
    SFAMD_MFMA_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

bool dm_dense_mfma_supported(int D, int dtype, int64_t sk) {
  return mfma_supported(D, dtype) && (sk == 1 || sk >= 16);
}

int dm_dense_mfma(const void* in, void* out, int64_t N, int D, int64_t sk, const void* M, hipStream_t s) {
  if (!(sk == 1 || sk >= 16)) return SFAMD_E_ARG;
  const int64_t nouter = N / ((int64_t)D * D * sk);
  const int64_t ntile = sk == 1 ? (nouter + 15) / 16 : nouter * ((sk + 15) / 16);
  const int grid = (int)(ntile < kMaxGrid ? ntile : kMaxGrid);
  switch (D) {
#define CASE(d)                                                                                              \
  case d:                                                                                                    \
    if (sk == 1)                                                                                             \
      hipLaunchKernelGGL((k_dm_dense_mfma<d, true>), dim3(grid), dim3(kBlock), 0, s, (const C*)in, (C*)out, \
                         N, sk, (const C*)M);                                                                \
    else                                                                                                     \
      hipLaunchKernelGGL((k_dm_dense_mfma<d, false>), dim3(grid), dim3(kBlock), 0, s, (const C*)in,         \
                         (C*)out, N, sk, (const C*)M);                                                       \
    break;
    SFAMD_MFMA_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

bool dm_pair_mfma_supported(int n, int D, int m0, int m1, int dtype) {
  const int q = m0 > m1 ? m0 : m1;
  if (!mfma_supported(D, dtype) || m0 == m1 || q >= n - 1) return false;
  int64_t R = 1;
  for (int i = 0; i < 2 * (n - 1 - q); ++i) R *= D;
  return R >= 16;
}

int dm_pair_mfma(const void* in, void* out, int n, int D, int m0, int m1, const void* blocks, int diff,
                 hipStream_t s) {
  if (!dm_pair_mfma_supported(n, D, m0, m1, SFAMD_C64)) return SFAMD_E_ARG;
  auto pw = [D](int e) {
    int64_t r = 1;
    for (int i = 0; i < e; ++i) r *= D;
    return r;
  };
  const int p = m0 < m1 ? m0 : m1, q = m0 > m1 ? m0 : m1;
  DmPairGeo geo;
  // mode m: unconjugated axis stride D^(2(n-1-m)+1), conjugated D^(2(n-1-m))
  geo.ap = pw(2 * (n - 1 - m0) + 1);
  geo.aq = pw(2 * (n - 1 - m1) + 1);
  geo.bp = pw(2 * (n - 1 - m0));
  geo.bq = pw(2 * (n - 1 - m1));
  geo.R = pw(2 * (n - 1 - q));
  geo.SM = pw(2 * (n - q));
  geo.SO = pw(2 * (n - p));
  geo.M = pw(2 * (q - p - 1));
  geo.nch = (geo.R + 15) / 16;
  geo.nchunk = pw(2 * p) * geo.M * geo.nch;
  const int64_t nsb = 2 * D - 1;
  const int64_t ntask = nsb * nsb * ((geo.nchunk + kDmPairGroup - 1) / kDmPairGroup);
  const int grid = (int)(ntask < kMaxGrid ? ntask : kMaxGrid);
  switch (D) {
#define CASE(d)                                                                                          \
  case d:                                                                                                \
    hipLaunchKernelGGL((k_dm_pair_mfma<d>), dim3(grid), dim3(kBlock), 0, s, (const C*)in, (C*)out, geo, \
                       (const C*)blocks, diff);                                                          \
    break;
    SFAMD_MFMA_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
  return (int)hipGetLastError();
}

}  // namespace sfamd
