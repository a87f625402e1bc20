// LDS-staged beamsplitter and loss kernels for inner gate axes of stride < 16.
#pragma once

#include "common.h"

namespace sfamd {

constexpr int kSlabMaxD = 20;

// whether a slab tiling exists for gate axes of strides S_out > Sin
bool slab_supported(int64_t N, int D, int64_t S_out, int64_t Sin, int dtype, int adjoint);
// diff = 1: difference-conserving blocks (two-mode squeezing), complex64 only
int pair_bs_slab(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
                 int dtype, hipStream_t s, int diff = 0);
int pair_bs_adj_slab(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                     const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial, int nwg,
                     int zp1, int dtype, hipStream_t s, int diff = 0);
int loss_slab(const void* in, void* out, int64_t N, int D, int64_t sb, int64_t sk, const void* coef, int dtype,
              hipStream_t s);
int loss_adj_slab(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sb, int64_t sk,
                  const void* coef, const void* dcoef, double* partial, int nwg, int zp1, int dtype, hipStream_t s);

}  // namespace sfamd
