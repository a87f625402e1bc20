// Fused gate stages: host-side sizes and the dispatch to the per-cutoff
// launchers (the kernels are in fused_impl.h, instantiated by the
// fused_dNN_MM.hip translation units).
#include "fused_impl.h"

namespace sfamd {

bool stage_supported(int D, int K) {
  if (D < 2 || D > kStageMaxD || K < 2 || K > 4) return false;
  return ipow(D, K) <= kStageMaxTile;
}

int64_t stage_gate_elems(int kind, int D, bool derivs) {
  switch (kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      return D;
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE:
      return derivs ? 2 * (int64_t)D * D : (int64_t)D * D;
    case SFAMD_OP_BEAMSPLITTER:
      if (D <= kFullBsMaxD) return (derivs ? 2 : 1) * ipow(D, 3);
      return derivs ? 2 * bs_nb(D) : bs_nb(D);
    default:
      return 0;
  }
}

int64_t stage_lds_fwd(int D, int K, int mat_fwd, int dtype) {
  const int64_t esz = dtype == SFAMD_C64 ? 8 : 16;
  return (2 * (int64_t)stage_tile(D, K) + mat_fwd) * esz;
}

int64_t stage_lds_adj(int D, int K, int ngates, int mat_adj, int dtype) {
  const int64_t esz = dtype == SFAMD_C64 ? 8 : 16;
  return ((ngates + 3) * (int64_t)stage_tile(D, K) + mat_adj) * esz +
         (int64_t)ngates * 2 * (stage_threads(D, K) / 64) * 8;
}

int stage_fwd(int D, int K, const void* in, void* out, const void* table, const StageArgs& a, int grid, int dtype,
              hipStream_t s) {
  switch (D) {
#define CASE_D(d) \
  case d:         \
    return stage_fwd_d##d(K, in, out, table, a, grid, dtype, s);
    SFAMD_STAGE_FOR_EACH_D(CASE_D)
#undef CASE_D
    default:
      return SFAMD_E_CUTOFF;
  }
}

int stage_adj(int D, int K, const void* psi, const void* lam_out, void* lam_in, const void* table,
              const StageArgs& a, int grid, double* partial, int nwg, int dtype, hipStream_t s) {
  switch (D) {
#define CASE_D(d) \
  case d:         \
    return stage_adj_d##d(K, psi, lam_out, lam_in, table, a, grid, partial, nwg, dtype, s);
    SFAMD_STAGE_FOR_EACH_D(CASE_D)
#undef CASE_D
    default:
      return SFAMD_E_CUTOFF;
  }
}

}  // namespace sfamd
