// Fused pure-state gate stages: descriptors shared by the planner
// (program.hip) and the kernels (fused.hip).
#pragma once

#include "common.h"

namespace sfamd {

constexpr int kStageMaxGates = 40;   // gates per stage (kernel-argument space)
constexpr int kStageMaxTile = 4096;  // amplitudes per LDS tile (D^K)
constexpr int kStageMaxD = 20;       // cutoffs with compiled stage kernels: 2..20
constexpr int kStageMaxModes = 24;   // modes outside the tile
// dynamic LDS a stage may use: gfx950 has 160 KB of LDS per compute unit
constexpr int64_t kStageLdsPerCu = 160 * 1024;
constexpr int64_t kStageLdsBudget = 152 * 1024;
constexpr int64_t kComputeUnits = 256;  // MI355X

struct StageGate {
  int8_t kind;    // SFAMD_OP_ROTATION .. SFAMD_OP_BEAMSPLITTER
  int8_t a0;      // local axis of the gate's first mode (0 = slowest)
  int8_t a1;      // local axis of the second mode (beamsplitter)
  int8_t pad;
  int32_t op;     // program operation index (gradient slot)
  int32_t s0;     // tile stride of axis a0 (D^(K-1-a0))
  int32_t s1;     // tile stride of axis a1
  int32_t moff;   // LDS offset of the matrix, forward layout (matrix only)
  int32_t amoff;  // LDS offset of matrix + derivative, adjoint layout
};

// The fields a stage kernel's prologue reads come first (one or two scalar
// cache lines with the pointer arguments), so their loads are issued together.
struct StageArgs {
  int32_t ngates;
  int32_t nnl;        // modes outside the tile
  int32_t ntiles;     // D^nnl
  int32_t grid;       // workgroups of the launch (set by the launcher)
  int32_t mat_fwd;    // LDS matrix region of the forward kernel (complex elements)
  int32_t mat_adj;    // LDS matrix region of the adjoint kernel (complex elements)
  int64_t img_fwd;    // gate-table offset of the forward region's image
  int64_t img_adj;    // gate-table offset of the adjoint region's image
  int64_t loc_stride[4];               // global strides of the local modes
  int64_t nl_stride[kStageMaxModes];   // global strides of the other modes (slowest first)
  StageGate g[kStageMaxGates];
};

// tile of K local modes: D^K amplitudes
__host__ __device__ constexpr int stage_tile(int D, int K) {
  return K == 2 ? D * D : K == 3 ? D * D * D : D * D * D * D;
}
// threads per workgroup: one tile amplitude per thread up to 1024, at least
// 256 (more waves in flight per SIMD to hide LDS latency)
__host__ __device__ constexpr int stage_threads(int D, int K) {
  return stage_tile(D, K) >= 1024 ? 1024 : (stage_tile(D, K) <= 256 ? 256 : ((stage_tile(D, K) + 63) / 64) * 64);
}

bool stage_supported(int D, int K);
// Beamsplitters at cutoffs up to kFullBsMaxD are staged in the zero-padded
// per-fiber layout (D^3 entries per matrix), larger ones in the compact block
// layout of the gate table.
constexpr int kFullBsMaxD = 10;
// LDS matrix elements of one gate: forward (the matrix) or adjoint (matrix and
// the first parameter's derivative)
int64_t stage_gate_elems(int kind, int D, bool derivs);
// dynamic LDS bytes the stage kernels need
int64_t stage_lds_fwd(int D, int K, int mat_fwd, int dtype);
int64_t stage_lds_adj(int D, int K, int ngates, int mat_adj, int dtype);
int stage_fwd(int D, int K, const void* in, void* out, const void* table, const StageArgs& a, int grid, int dtype,
              hipStream_t s);
// per-cutoff launchers behind stage_fwd / stage_adj (fused_dNN_MM.hip)
#define SFAMD_STAGE_FOR_EACH_D(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)
#define SFAMD_STAGE_DECLARE_D(d)                                                                                 \
  int stage_fwd_d##d(int K, const void* in, void* out, const void* table, const StageArgs& a, int grid, int dtype, \
                     hipStream_t s);                                                                             \
  int stage_adj_d##d(int K, const void* psi, const void* lo, void* li, const void* table, const StageArgs& a,      \
                     int grid, double* partial, int nwg, int dtype, hipStream_t s);
SFAMD_STAGE_FOR_EACH_D(SFAMD_STAGE_DECLARE_D)
#undef SFAMD_STAGE_DECLARE_D
int stage_adj(int D, int K, const void* psi, const void* lam_out, void* lam_in, const void* table,
              const StageArgs& a, int grid, double* partial, int nwg, int dtype, hipStream_t s);

}  // namespace sfamd
