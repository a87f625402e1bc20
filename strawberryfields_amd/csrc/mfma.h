// Matrix-core kernels for dense single-mode gates (complex64).
#pragma once

#include "common.h"

namespace sfamd {

constexpr int kMfmaMaxD = 20;

bool mfma_supported(int D, int dtype);
// out = U' in along an axis of stride R (U' = U, or conj(U) if conj)
int axis_dense_mfma(const void* in, void* out, int64_t N, int D, int64_t R, const void* M, int conj, hipStream_t s);
// lam_in = U'^H lam_out and the two parameter partial sums (grid = nwg)
int axis_dense_adj_mfma(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t R,
                        const void* M, const void* dM0, const void* dM1, int conj, double* partial, int nwg, int zp1,
                        hipStream_t s);

// beamsplitter on axes of strides sp (first mode) and sq, blocks in the compact
// layout of the gate table; use when the inner gate axis run is >= 16.
// diff = 1: difference-conserving blocks (two-mode squeezing, block s holds
// n1 - n2 = s - (D - 1)) instead of photon-number blocks.
int pair_bs_mfma(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
                 hipStream_t s, int diff = 0);
int pair_bs_adj_mfma(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                     const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial, int nwg,
                     int zp1, hipStream_t s, int diff = 0);

// rho' = U rho U^H on one mode of a DM (bra stride D * sk, ket stride sk) in
// one pass; supported for sk == 1 or sk >= 16
bool dm_dense_mfma_supported(int D, int dtype, int64_t sk);
int dm_dense_mfma(const void* in, void* out, int64_t N, int D, int64_t sk, const void* M, hipStream_t s);

// rho' = G rho G^H for a beamsplitter-family gate G on modes (m0, m1) of an
// n-mode DM (axes [m: unconjugated, conjugated] per mode, mode 0 slowest), in
// one pass; blocks as for pair_bs_mfma.  Supported when the modes after
// max(m0, m1) leave an innermost run of >= 16 entries.
bool dm_pair_mfma_supported(int n, int D, int m0, int m1, int dtype);
int dm_pair_mfma(const void* in, void* out, int n, int D, int m0, int m1, const void* blocks, int diff,
                 hipStream_t s);

}  // namespace sfamd
