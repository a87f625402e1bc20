// Internal helpers shared by the HIP sources of libsfamd (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sfamd.h"

#define SFAMD_VERSION 1

// ---- status helpers ------------------------------------------------------
#define HIP_RET(expr)                                   \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return (int)e_;               \
  } while (0)

#define CHECK_ARG(cond)                                 \
  do {                                                  \
    if (!(cond)) return SFAMD_E_ARG;                    \
  } while (0)

#define RET_IF(st)                                      \
  do {                                                  \
    int s_ = (st);                                      \
    if (s_ != 0) return s_;                             \
  } while (0)

// ---- complex arithmetic (interleaved re, im as torch stores it) ----------
template <typename T>
struct cpx {
  T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T re, T im) {
  cpx<T> z;
  z.re = re;
  z.im = im;
  return z;
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return mk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return mk<T>(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cconj(cpx<T> a) {
  return mk<T>(a.re, -a.im);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) {
  return mk<T>(a.re * s, a.im * s);
}
// acc += a * b
template <typename T>
__host__ __device__ __forceinline__ void cmac(cpx<T>& acc, cpx<T> a, cpx<T> b) {
  acc.re += a.re * b.re - a.im * b.im;
  acc.im += a.re * b.im + a.im * b.re;
}
// acc += conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ void cmac_conj(cpx<T>& acc, cpx<T> a, cpx<T> b) {
  acc.re += a.re * b.re + a.im * b.im;
  acc.im += a.re * b.im - a.im * b.re;
}
// Re(conj(a) * b)
template <typename T>
__host__ __device__ __forceinline__ T re_conj_mul(cpx<T> a, cpx<T> b) {
  return a.re * b.re + a.im * b.im;
}

// ---- sizes ---------------------------------------------------------------
__host__ __device__ __forceinline__ int64_t ipow(int64_t b, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

// Photon-number block s of a two-mode gate at cutoff D: in-range first-mode
// indices lo..hi, and the offset of the block in the compact block layout
// (sum over s' < s of c(s')^2, c(s) = hi - lo + 1).
__host__ __device__ __forceinline__ int bs_lo(int s, int D) { return s - D + 1 > 0 ? s - D + 1 : 0; }
__host__ __device__ __forceinline__ int bs_hi(int s, int D) { return s < D - 1 ? s : D - 1; }
__host__ __device__ __forceinline__ int64_t sum_sq(int64_t x) { return x * (x + 1) * (2 * x + 1) / 6; }
// block offsets in 32-bit unsigned arithmetic: for s <= 2D - 1 <= 79 the
// products stay below 2^32, and the division by 6 becomes a multiply-high on
// the GPU instead of a 64-bit division sequence (it runs per amplitude in the
// element-wise and stage kernels)
__host__ __device__ __forceinline__ unsigned sum_sq_u32(unsigned x) { return x * (x + 1) * (2 * x + 1) / 6u; }
__host__ __device__ __forceinline__ int64_t bs_block_off(int s, int D) {
  if (s <= D) return (int64_t)sum_sq_u32((unsigned)s);
  return (int64_t)(sum_sq_u32((unsigned)D) + sum_sq_u32((unsigned)(D - 1)) - sum_sq_u32((unsigned)(2 * D - s - 1)));
}
__host__ __device__ __forceinline__ int64_t bs_nb(int D) { return bs_block_off(2 * D - 1, D); }

// ---- launch geometry -----------------------------------------------------
constexpr int kBlock = 256;
constexpr int kMaxGrid = 8192;

inline int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return (int)g;
}

#define GRID_STRIDE(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// Block-wide sum of a double (blockDim.x == kBlock, wavefront 64).
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// ---- mode subsets ----------------------------------------------------------
// kept: the listed modes in increasing order; traced: the others, increasing
struct ModeList {
  int32_t n, k;
  int32_t kept[32];
  int32_t traced[32];
};

inline int make_modelist(int n, const int* modes, int k, ModeList* ml) {
  if (n > 32 || k < 0 || k > n) return SFAMD_E_ARG;
  ml->n = n;
  ml->k = k;
  bool used[32] = {false};
  for (int t = 0; t < k; ++t) {
    if (modes[t] < 0 || modes[t] >= n || used[modes[t]]) return SFAMD_E_ARG;
    if (t > 0 && modes[t] <= modes[t - 1]) return SFAMD_E_ARG;
    used[modes[t]] = true;
    ml->kept[t] = modes[t];
  }
  int u = 0;
  for (int m = 0; m < n; ++m)
    if (!used[m]) ml->traced[u++] = m;
  return 0;
}

// ---- internal launchers (implemented in the .hip files) ------------------
namespace sfamd {

// A run of consecutive rotation / Kerr gates applied in one pass (diag_run.hip).
constexpr int kDiagRunMax = 8;
struct DiagRun {
  int32_t n, D, mixed, ngates;
  int32_t mode[kDiagRunMax], kind[kDiagRunMax], op[kDiagRunMax];
  int64_t toff[kDiagRunMax];    // table offset of each gate's D phases (complex elements)
  int64_t rowdiv[kDiagRunMax];  // R^(n-2-mode): the gate's digit pair in a row index
  int64_t N;                    // state elements
};
int diag_run_fwd(const void* in, void* out, const void* table, const DiagRun& r, int dtype, hipStream_t s);
int diag_run_adj(const void* psi, const void* lam_out, void* lam_in, const void* table, const DiagRun& r,
                 double* partial, int nwg, int dtype, hipStream_t s);

int64_t table_size(int kind, int D);

// one operation of a table build: kind and table offset; the operation's
// parameters are params[2*(first + i) + slot].  An operation inside a fused
// stage also writes its regions of the stage's matrix images (the LDS layout
// of the forward and adjoint stage kernels) at img_fwd / img_adj (< 0: none).
struct BuildDesc {
  int32_t kind;
  int32_t off;      // complex elements
  int32_t img_fwd;  // complex elements, or -1
  int32_t img_adj;  // complex elements, or -1
};
constexpr int kMaxBuildPerLaunch = 240;  // kernel-argument space (4 KB)
struct BuildBatch {
  int32_t first;  // program index of d[0]; < 0: single gate with inline p0, p1
  int32_t pad;
  double p0, p1;
  BuildDesc d[kMaxBuildPerLaunch];
};
int launch_build(const BuildBatch& b, int count, int D, const double* params,
                 void* table, int dtype, hipStream_t s);

// forward kernels on a tensor of axes of extent D (N elements)
int axis_dense(const void* in, void* out, int64_t N, int D, int64_t B,
               const void* M, int conj, int dtype, hipStream_t s);
int axis_diag(const void* in, void* out, int64_t N, int D, int64_t B,
              const void* d, int conj, int dtype, hipStream_t s);
int pair_diag(const void* in, void* out, int64_t N, int D, int64_t Bbra,
              int64_t Bket, const void* d, int dtype, hipStream_t s);
int pair_bs(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq,
            const void* blocks, int conj, int dtype, hipStream_t s);
int pair_s2(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq,
            const void* blocks, int conj, int dtype, hipStream_t s);
int pair_s2_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N,
                int D, int64_t sp, int64_t sq, const void* blocks,
                const void* dblocks, int conj, double* partial, int nwg, int zp1,
                int dtype, hipStream_t s);
int loss_fwd(const void* in, void* out, int64_t N, int D, int64_t sp,
             int64_t sq, const void* coef, int dtype, hipStream_t s);
int mix_fwd(const void* ket, void* dm, int n, int D, int dtype, hipStream_t s);
int replace_mode(const void* in, void* out, int64_t N, int D, int64_t sp,
                 int64_t sq, const void* sigma, int dtype, hipStream_t s);
int fill_zero(void* p, int64_t bytes, hipStream_t s);

// adjoint kernels; partial = base of this launch's [slot][nwg] doubles; with
// zp1 set the launch also writes zeros to the next pass block (partial + 2 nwg)
int axis_dense_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N,
                   int D, int64_t B, const void* M, const void* dM0,
                   const void* dM1, int conj, double* partial, int nwg, int zp1,
                   int dtype, hipStream_t s);
int axis_diag_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N,
                  int D, int64_t B, const void* d, int kerr, int conj,
                  double* partial, int nwg, int zp1, int dtype, hipStream_t s);
int pair_diag_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N,
                  int D, int64_t Bbra, int64_t Bket, const void* d, int kerr,
                  double* partial, int nwg, int zp1, int dtype, hipStream_t s);
// beamsplitter-family adjoint: the second parameter's derivative is
// i sigma (wa n + wb N) G[N][n] (BS: wa = 1, wb = -1; MZ: 1, 0; CK: 0, 0)
int pair_bs_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N,
                int D, int64_t sp, int64_t sq, const void* blocks,
                const void* dblocks, int conj, int wa, int wb, double* partial,
                int nwg, int zp1, int dtype, hipStream_t s);
int loss_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N,
             int D, int64_t sp, int64_t sq, const void* coef, const void* dcoef,
             double* partial, int nwg, int zp1, int dtype, hipStream_t s);
int mix_adj(const void* ket, const void* g_dm, void* g_ket, int n, int D,
            int dtype, hipStream_t s);
// grads[2*op + slot] = sum over passes p < npass and columns w < ncol of
// partial[((op*2 + p)*2 + slot)*nwg + w], for the listed ops; cols[i] holds
// ncol (<= nwg) with bit 15 set when only pass 0 was written (fused stages
// write one column per launched workgroup)
constexpr int kOpListMax = 600;  // fits the 4 KB kernel-argument space
constexpr unsigned kOpColsOnePass = 0x8000u;
struct OpList {
  int32_t n;
  int32_t ops[kOpListMax];
  uint16_t cols[kOpListMax];
};
int reduce_grads(const double* partial, double* grads, const OpList& ol, int nwg, hipStream_t s);

}  // namespace sfamd
