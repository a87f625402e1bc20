// Per-fiber kernels for large states (fiber.hip); launchers return
// SFAMD_E_CUTOFF when no kernel is compiled for the cutoff.
#pragma once

#include "common.h"

namespace sfamd {

constexpr int kFiberMaxD = 20;
// use the per-fiber kernels when the innermost non-gate run is at least this
// many amplitudes (coalesced fibers); shorter runs stay on the element-wise
// kernels, whose re-reads then fall in L1
constexpr int64_t kFiberMinRun = 16;

bool fiber_supported(int D);
int axis_dense_fib(const void* in, void* out, int64_t N, int D, int64_t B, const void* M, int conj, int parity,
                   int dtype, hipStream_t s);
int axis_dense_adj_fib(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t B, const void* M,
                       const void* dM0, const void* dM1, int conj, int parity, double* partial, int nwg, int zp1,
                       int dtype, hipStream_t s);
// diff = 1: difference-conserving blocks (two-mode squeezing) instead of
// photon-number blocks
int pair_bs_fib(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
                int dtype, hipStream_t s, int diff = 0);
int pair_bs_adj_fib(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                    const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial, int nwg,
                    int zp1, int dtype, hipStream_t s, int diff = 0);
int loss_fib(const void* in, void* out, int64_t N, int D, int64_t B, const void* coef, int dtype, hipStream_t s);
int loss_adj_fib(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t B, const void* coef,
                 const void* dcoef, double* partial, int nwg, int zp1, int dtype, hipStream_t s);

}  // namespace sfamd
