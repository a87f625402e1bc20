// Recorded-circuit executor and the C ABI entry points of libsfamd.
//
// A program is the list of Fock-backend operations the Engine issued
// (strawberryfields/engine.py:427-430 calls op.apply per command); the host
// loop below issues the kernels of each operation on one stream, keeps the
// state in caller-provided buffers, writes the inputs of requested operations
// as checkpoints and runs the reverse sweep.  The whole sequence can be
// captured into a HIP graph by the caller (sfamd_graph_*).
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"
#include "fiber.h"
#include "fused.h"
#include "mfma.h"
#include "slab.h"

using namespace sfamd;

namespace {

struct OpPlan {
  int kind;
  int m0, m1;
  int mixed_in;   // state at the op input is a DM
  int mix_first;  // loss on a pure state: convert to a DM first
  int64_t off;    // table offset (complex elements)
};

// strides of mode m in a ket of n modes and in a DM of n modes
inline int64_t ket_stride(int n, int D, int m) { return ipow(D, n - 1 - m); }
inline int64_t bra_stride_dm(int n, int D, int m) { return ipow(D, 2 * (n - 1 - m) + 1); }
inline int64_t ket_stride_dm(int n, int D, int m) { return ipow(D, 2 * (n - 1 - m)); }

inline int64_t elem_size(int dtype) { return dtype == SFAMD_C64 ? 8 : 16; }

inline const char* cptr(const void* p, int64_t elems, int dtype) {
  return (const char*)p + elems * elem_size(dtype);
}

bool two_mode_kind(int kind) {
  return kind == SFAMD_OP_BEAMSPLITTER || kind == SFAMD_OP_TWO_MODE_SQUEEZE || kind == SFAMD_OP_MZGATE ||
         kind == SFAMD_OP_CROSS_KERR;
}

int validate_op(int kind, int m0, int m1, int n, int D) {
  if (kind < 0 || kind > SFAMD_OP_CUBIC_PHASE) return SFAMD_E_KIND;
  if (D < 1 || D > SFAMD_MAX_CUTOFF) return SFAMD_E_CUTOFF;
  if (two_mode_kind(kind)) {
    if (D > SFAMD_MAX_CUTOFF_BS) return SFAMD_E_CUTOFF;
    if (m0 < 0 || m0 >= n || m1 < 0 || m1 >= n || m0 == m1) return SFAMD_E_ARG;
  } else if (kind != SFAMD_OP_MIX) {
    if (kind == SFAMD_OP_CUBIC_PHASE && D > SFAMD_MAX_CUTOFF_V) return SFAMD_E_CUTOFF;
    if (m0 < 0 || m0 >= n) return SFAMD_E_ARG;
  }
  return 0;
}

// second-parameter derivative weight of a photon-number-conserving two-mode
// gate: d/dp1 G[N][n] = i (wa n + wb N) G[N][n]
void bs_weights(int kind, int* wa, int* wb) {
  *wa = kind == SFAMD_OP_CROSS_KERR ? 0 : 1;
  *wb = kind == SFAMD_OP_BEAMSPLITTER ? -1 : 0;
}

// Routing between element-wise and per-fiber kernels: per-fiber when a kernel
// is compiled for D and the innermost non-gate run is long enough to coalesce.
bool use_fib(int D, int64_t run) { return fiber_supported(D) && run >= kFiberMinRun; }

// complex64 dense gates run on the matrix cores for every axis stride
int dense_pass(const void* in, void* out, int64_t N, int D, int64_t B, const void* U, int conj, int parity, int dtype,
               hipStream_t s) {
  if (mfma_supported(D, dtype)) return axis_dense_mfma(in, out, N, D, B, U, conj, s);
  if (use_fib(D, B)) return axis_dense_fib(in, out, N, D, B, U, conj, parity, dtype, s);
  return axis_dense(in, out, N, D, B, U, conj, dtype, s);
}

int dense_adj_pass(const void* psi, const void* lo, void* li, int64_t N, int D, int64_t B, const void* U,
                   const void* dU0, const void* dU1, int conj, int parity, double* pb, int nwg, int zp1, int dtype,
                   hipStream_t s) {
  if (mfma_supported(D, dtype)) return axis_dense_adj_mfma(psi, lo, li, N, D, B, U, dU0, dU1, conj, pb, nwg, zp1, s);
  if (use_fib(D, B)) return axis_dense_adj_fib(psi, lo, li, N, D, B, U, dU0, dU1, conj, parity, pb, nwg, zp1, dtype, s);
  return axis_dense_adj(psi, lo, li, N, D, B, U, dU0, dU1, conj, pb, nwg, zp1, dtype, s);
}

// Two-axis gates: per-fiber kernels when the inner gate axis has a long
// coalesced run; LDS slab tiles when it is (nearly) innermost and a tile fits;
// element-wise kernels otherwise.  Slab tiles need a large enough state to pay
// for their staging, so tiny states stay element-wise.
constexpr int64_t kSlabMinState = 1 << 16;

bool use_slab(int64_t N, int D, int64_t sa, int64_t sb, int dtype, int adjoint) {
  const int64_t so = sa > sb ? sa : sb, si = sa > sb ? sb : sa;
  return N >= kSlabMinState && slab_supported(N, D, so, si, dtype, adjoint);
}

int bs_pass(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* g, int conj, int dtype,
            hipStream_t s) {
  if (use_fib(D, sp < sq ? sp : sq)) {
    if (mfma_supported(D, dtype)) return pair_bs_mfma(in, out, N, D, sp, sq, g, conj, s);
    return pair_bs_fib(in, out, N, D, sp, sq, g, conj, dtype, s);
  }
  if (use_slab(N, D, sp, sq, dtype, 0)) return pair_bs_slab(in, out, N, D, sp, sq, g, conj, dtype, s);
  return pair_bs(in, out, N, D, sp, sq, g, conj, dtype, s);
}

int bs_adj_pass(const void* psi, const void* lo, void* li, int64_t N, int D, int64_t sp, int64_t sq, const void* g,
                const void* dB, int conj, int wa, int wb, double* pb, int nwg, int zp1, int dtype, hipStream_t s) {
  if (use_fib(D, sp < sq ? sp : sq)) {
    if (mfma_supported(D, dtype))
      return pair_bs_adj_mfma(psi, lo, li, N, D, sp, sq, g, dB, conj, wa, wb, pb, nwg, zp1, s);
    return pair_bs_adj_fib(psi, lo, li, N, D, sp, sq, g, dB, conj, wa, wb, pb, nwg, zp1, dtype, s);
  }
  if (use_slab(N, D, sp, sq, dtype, 1))
    return pair_bs_adj_slab(psi, lo, li, N, D, sp, sq, g, dB, conj, wa, wb, pb, nwg, zp1, dtype, s);
  return pair_bs_adj(psi, lo, li, N, D, sp, sq, g, dB, conj, wa, wb, pb, nwg, zp1, dtype, s);
}

// two-mode pass of any two-mode kind (S2 on its own kernels, the others on the
// beamsplitter family)
int pair_pass(int kind, const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* g, int conj,
              int dtype, hipStream_t s) {
  if (kind == SFAMD_OP_TWO_MODE_SQUEEZE) {
    // the beamsplitter's per-fiber (matrix-core for complex64) and slab kernels on
    // difference blocks; the element-wise S2 kernel otherwise
    const bool fib = use_fib(D, sp < sq ? sp : sq);
    if (mfma_supported(D, dtype)) {
      if (fib) return pair_bs_mfma(in, out, N, D, sp, sq, g, conj, s, 1);
      if (use_slab(N, D, sp, sq, dtype, 0)) return pair_bs_slab(in, out, N, D, sp, sq, g, conj, dtype, s, 1);
    } else if (fib) {
      return pair_bs_fib(in, out, N, D, sp, sq, g, conj, dtype, s, 1);
    }
    return pair_s2(in, out, N, D, sp, sq, g, conj, dtype, s);
  }
  return bs_pass(in, out, N, D, sp, sq, g, conj, dtype, s);
}

int pair_adj_pass(int kind, const void* psi, const void* lo, void* li, int64_t N, int D, int64_t sp, int64_t sq,
                  const void* g, const void* dG, int conj, double* pb, int nwg, int zp1, int dtype, hipStream_t s) {
  if (kind == SFAMD_OP_TWO_MODE_SQUEEZE) {
    // d/dphi G[N][n] = i (N - n) G[N][n]: weights wa = -1, wb = 1
    const bool fib = use_fib(D, sp < sq ? sp : sq);
    if (mfma_supported(D, dtype)) {
      if (fib) return pair_bs_adj_mfma(psi, lo, li, N, D, sp, sq, g, dG, conj, -1, 1, pb, nwg, zp1, s, 1);
      if (use_slab(N, D, sp, sq, dtype, 1))
        return pair_bs_adj_slab(psi, lo, li, N, D, sp, sq, g, dG, conj, -1, 1, pb, nwg, zp1, dtype, s, 1);
    } else if (fib) {
      return pair_bs_adj_fib(psi, lo, li, N, D, sp, sq, g, dG, conj, -1, 1, pb, nwg, zp1, dtype, s, 1);
    }
    return pair_s2_adj(psi, lo, li, N, D, sp, sq, g, dG, conj, pb, nwg, zp1, dtype, s);
  }
  int wa, wb;
  bs_weights(kind, &wa, &wb);
  return bs_adj_pass(psi, lo, li, N, D, sp, sq, g, dG, conj, wa, wb, pb, nwg, zp1, dtype, s);
}

int loss_pass(const void* in, void* out, int64_t N, int D, int64_t sb, int64_t sk, const void* c, int dtype,
              hipStream_t s) {
  if (use_fib(D, sk)) return loss_fib(in, out, N, D, sk, c, dtype, s);
  if (use_slab(N, D, sb, sk, dtype, 0)) return loss_slab(in, out, N, D, sb, sk, c, dtype, s);
  return loss_fwd(in, out, N, D, sb, sk, c, dtype, s);
}

int loss_adj_pass(const void* psi, const void* lo, void* li, int64_t N, int D, int64_t sb, int64_t sk, const void* c,
                  const void* dc, double* pb, int nwg, int zp1, int dtype, hipStream_t s) {
  if (use_fib(D, sk)) return loss_adj_fib(psi, lo, li, N, D, sk, c, dc, pb, nwg, zp1, dtype, s);
  if (use_slab(N, D, sb, sk, dtype, 1)) return loss_adj_slab(psi, lo, li, N, D, sb, sk, c, dc, pb, nwg, zp1, dtype, s);
  return loss_adj(psi, lo, li, N, D, sb, sk, c, dc, pb, nwg, zp1, dtype, s);
}

// forward of one operation: in -> out (tmp: scratch state for two-pass ops)
int apply_op(const OpPlan& op, int n, int D, const void* table, const void* in, void* out, void* tmp, int dtype,
             hipStream_t s) {
  const int64_t NK = ipow(D, n), NM = ipow(D, 2 * n);
  const void* g = cptr(table, op.off, dtype);
  switch (op.kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      if (op.mixed_in)
        return pair_diag(in, out, NM, D, bra_stride_dm(n, D, op.m0), ket_stride_dm(n, D, op.m0), g, dtype, s);
      return axis_diag(in, out, NK, D, ket_stride(n, D, op.m0), g, 0, dtype, s);
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE:
    case SFAMD_OP_CUBIC_PHASE: {
      const int par = op.kind == SFAMD_OP_SQUEEZE;
      if (op.mixed_in) {
        const int64_t sk = ket_stride_dm(n, D, op.m0);
        if (dm_dense_mfma_supported(D, dtype, sk)) return dm_dense_mfma(in, out, NM, D, sk, g, s);
        if (!tmp) return SFAMD_E_WORKSPACE;
        RET_IF(dense_pass(in, tmp, NM, D, bra_stride_dm(n, D, op.m0), g, 0, par, dtype, s));
        return dense_pass(tmp, out, NM, D, ket_stride_dm(n, D, op.m0), g, 1, par, dtype, s);
      }
      return dense_pass(in, out, NK, D, ket_stride(n, D, op.m0), g, 0, par, dtype, s);
    }
    case SFAMD_OP_BEAMSPLITTER:
    case SFAMD_OP_TWO_MODE_SQUEEZE:
    case SFAMD_OP_MZGATE:
    case SFAMD_OP_CROSS_KERR:
      if (op.mixed_in) {
        // one pass (rho' = G rho G^H) when a run of inner modes follows the pair
        if (dm_pair_mfma_supported(n, D, op.m0, op.m1, dtype))
          return dm_pair_mfma(in, out, n, D, op.m0, op.m1, g, op.kind == SFAMD_OP_TWO_MODE_SQUEEZE, s);
        if (!tmp) return SFAMD_E_WORKSPACE;
        RET_IF(pair_pass(op.kind, in, tmp, NM, D, bra_stride_dm(n, D, op.m0), bra_stride_dm(n, D, op.m1), g, 0,
                         dtype, s));
        return pair_pass(op.kind, tmp, out, NM, D, ket_stride_dm(n, D, op.m0), ket_stride_dm(n, D, op.m1), g, 1,
                         dtype, s);
      }
      return pair_pass(op.kind, in, out, NK, D, ket_stride(n, D, op.m0), ket_stride(n, D, op.m1), g, 0, dtype, s);
    case SFAMD_OP_LOSS: {
      const void* src = in;
      if (op.mix_first) {
        if (!tmp) return SFAMD_E_WORKSPACE;
        RET_IF(mix_fwd(in, tmp, n, D, dtype, s));
        src = tmp;
      }
      return loss_pass(src, out, NM, D, bra_stride_dm(n, D, op.m0), ket_stride_dm(n, D, op.m0), g, dtype, s);
    }
    case SFAMD_OP_MIX:
      return mix_fwd(in, out, n, D, dtype, s);
    default:
      return SFAMD_E_KIND;
  }
}

// adjoint of one operation.  pb = partial base of this op: pass p writes
// slot 0 at pb + (2p)*nwg and slot 1 at pb + (2p+1)*nwg.
int adjoint_op(const OpPlan& op, int n, int D, const void* table, const void* psi, const void* lo, void* li,
               void* t1, void* t2, double* pb, int nwg, int dtype, hipStream_t s) {
  const int64_t NK = ipow(D, n), NM = ipow(D, 2 * n);
  const void* g = cptr(table, op.off, dtype);
  double* pass0 = pb;
  double* pass1 = pb + 2 * (int64_t)nwg;
  const int kerr = op.kind == SFAMD_OP_KERR;
  switch (op.kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      if (op.mixed_in)
        return pair_diag_adj(psi, lo, li, NM, D, bra_stride_dm(n, D, op.m0), ket_stride_dm(n, D, op.m0), g, kerr,
                             pass0, nwg, 1, dtype, s);
      return axis_diag_adj(psi, lo, li, NK, D, ket_stride(n, D, op.m0), g, kerr, 0, pass0, nwg, 1, dtype, s);
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE:
    case SFAMD_OP_CUBIC_PHASE: {
      const int64_t DD = (int64_t)D * D;
      const void* dU0 = cptr(g, DD, dtype);
      const void* dU1 = cptr(g, 2 * DD, dtype);
      const int par = op.kind == SFAMD_OP_SQUEEZE;
      if (op.mixed_in) {
        const int64_t sb = bra_stride_dm(n, D, op.m0), sk = ket_stride_dm(n, D, op.m0);
        // rho1 = U rho (bra side) recomputed; then the ket-side pass, then the bra-side pass
        RET_IF(dense_pass(psi, t1, NM, D, sb, g, 0, par, dtype, s));
        RET_IF(dense_adj_pass(t1, lo, t2, NM, D, sk, g, dU0, dU1, 1, par, pass1, nwg, 0, dtype, s));
        return dense_adj_pass(psi, t2, li, NM, D, sb, g, dU0, dU1, 0, par, pass0, nwg, 0, dtype, s);
      }
      return dense_adj_pass(psi, lo, li, NK, D, ket_stride(n, D, op.m0), g, dU0, dU1, 0, par, pass0, nwg, 1, dtype, s);
    }
    case SFAMD_OP_BEAMSPLITTER:
    case SFAMD_OP_TWO_MODE_SQUEEZE:
    case SFAMD_OP_MZGATE:
    case SFAMD_OP_CROSS_KERR: {
      const void* dB = cptr(g, bs_nb(D), dtype);
      const int k = op.kind;
      if (op.mixed_in) {
        const int64_t bp = bra_stride_dm(n, D, op.m0), bq = bra_stride_dm(n, D, op.m1);
        const int64_t kp = ket_stride_dm(n, D, op.m0), kq = ket_stride_dm(n, D, op.m1);
        RET_IF(pair_pass(k, psi, t1, NM, D, bp, bq, g, 0, dtype, s));
        RET_IF(pair_adj_pass(k, t1, lo, t2, NM, D, kp, kq, g, dB, 1, pass1, nwg, 0, dtype, s));
        return pair_adj_pass(k, psi, t2, li, NM, D, bp, bq, g, dB, 0, pass0, nwg, 0, dtype, s);
      }
      return pair_adj_pass(k, psi, lo, li, NK, D, ket_stride(n, D, op.m0), ket_stride(n, D, op.m1), g, dB, 0, pass0,
                           nwg, 1, dtype, s);
    }
    case SFAMD_OP_LOSS: {
      const int64_t D3 = (int64_t)D * D * D;
      const char* c = (const char*)g;
      const char* dc = c + D3 * (elem_size(dtype) / 2);
      const int64_t sb = bra_stride_dm(n, D, op.m0), sk = ket_stride_dm(n, D, op.m0);
      if (op.mix_first) {
        RET_IF(mix_fwd(psi, t1, n, D, dtype, s));
        RET_IF(loss_adj_pass(t1, lo, t2, NM, D, sb, sk, c, dc, pass0, nwg, 1, dtype, s));
        return mix_adj(psi, t2, li, n, D, dtype, s);
      }
      return loss_adj_pass(psi, lo, li, NM, D, sb, sk, c, dc, pass0, nwg, 1, dtype, s);
    }
    case SFAMD_OP_MIX:
      RET_IF(fill_zero(pb, 4 * (int64_t)nwg * sizeof(double), s));
      return mix_adj(psi, lo, li, n, D, dtype, s);
    default:
      return SFAMD_E_KIND;
  }
}

}  // namespace

// One step of execution: a single operation (element-wise kernels) or a
// fused stage of gates on K local modes of a pure state (fused.hip).
struct Step {
  int fused;
  int op;  // single operation index (fused == 0)
  int K;
  int diag;  // a run of rotation / Kerr gates in one pass (fused == 0)
  std::vector<int> ops;  // operation indices in execution order
  StageArgs args;
  DiagRun run;
};

struct sfamd_prog {
  int n, D, dtype, mixed_in, nops, nwg, fuse_K;
  std::vector<OpPlan> ops;
  std::vector<int32_t> img_fwd, img_adj;  // per op: its stage-image regions, or -1
  std::vector<Step> steps;
  std::vector<int64_t> state_elems;  // elements at the input of step i (i = nsteps: output)
  int64_t table_elems, max_state;
  int final_mixed;
};

namespace {

bool fusable(const OpPlan& op) {
  return !op.mixed_in && (op.kind == SFAMD_OP_ROTATION || op.kind == SFAMD_OP_KERR ||
                          op.kind == SFAMD_OP_DISPLACEMENT || op.kind == SFAMD_OP_SQUEEZE ||
                          op.kind == SFAMD_OP_BEAMSPLITTER);
}

int op_nmodes(const OpPlan& op) { return two_mode_kind(op.kind) ? 2 : 1; }

bool op_in(const OpPlan& op, unsigned mask) {
  if (!(mask >> op.m0 & 1u)) return false;
  return !two_mode_kind(op.kind) || (mask >> op.m1 & 1u);
}

unsigned op_mask(const OpPlan& op) {
  unsigned m = 1u << op.m0;
  if (two_mode_kind(op.kind)) m |= 1u << op.m1;
  return m;
}

// ops of `rem` (program order) absorbable into a stage on `mask`: an op joins
// if its modes lie in the mask, no earlier op left outside the stage touches
// one of them (gates on disjoint modes commute) and the adjoint kernel's LDS
// (one tile per gate plus two, and every gate's matrix + derivatives) fits
void absorb(const std::vector<OpPlan>& ops, const std::vector<int>& rem, unsigned mask, int D, int Kc,
            int64_t tile_bytes, int64_t esz, int64_t budget, std::vector<int>* got) {
  got->clear();
  unsigned blocked = 0;
  int64_t used = 3 * tile_bytes;  // the stage input and two cotangent tiles
  for (int i : rem) {
    const OpPlan& op = ops[i];
    const unsigned om = op_mask(op);
    const int64_t cost = tile_bytes + (((stage_gate_elems(op.kind, D, true) + 1) & ~(int64_t)1) * esz) +
                         2 * (stage_threads(D, Kc) / 64) * 8;
    if ((int)got->size() < kStageMaxGates && op_in(op, mask) && !(om & blocked) && used + cost <= budget) {
      got->push_back(i);
      used += cost;
    } else {
      blocked |= om;
    }
  }
}

// greedy partition of a run of fusable ops into stages of Kc local modes
void plan_run(sfamd_prog* p, const std::vector<int>& run, int Kc) {
  const int n = p->n, D = p->D;
  const int64_t esz = elem_size(p->dtype);
  const int64_t tile_bytes = stage_tile(D, Kc) * esz;
  // a launch has one workgroup per tile (up to the grid cap): when that is
  // more than one per compute unit, two must fit in a CU's LDS at once
  const int64_t ntiles = ipow(D, n - Kc);
  const int64_t budget = ntiles > kComputeUnits ? kStageLdsPerCu / 2 - 512 : kStageLdsBudget;
  std::vector<int> rem = run;
  std::vector<unsigned> cands;
  // all Kc-subsets when few, else subsets grown from the next ops
  std::vector<int> idx(Kc);
  int64_t ncomb = 1;
  for (int t = 0; t < Kc; ++t) ncomb = ncomb * (n - t) / (t + 1);
  if (ncomb <= 512) {
    for (int t = 0; t < Kc; ++t) idx[t] = t;
    while (true) {
      unsigned m = 0;
      for (int t = 0; t < Kc; ++t) m |= 1u << idx[t];
      cands.push_back(m);
      int t = Kc - 1;
      while (t >= 0 && idx[t] == n - Kc + t) --t;
      if (t < 0) break;
      ++idx[t];
      for (int u = t + 1; u < Kc; ++u) idx[u] = idx[u - 1] + 1;
    }
  }
  std::vector<int> got, best, rem2, got2;
  while (!rem.empty()) {
    std::vector<unsigned> local = cands;
    if (local.empty()) {
      unsigned m = 0;
      for (int i : rem) {
        const unsigned om = op_mask(p->ops[i]);
        if (__builtin_popcount(m | om) <= Kc) m |= om;
        if (__builtin_popcount(m) == Kc) break;
      }
      for (int md = n - 1; __builtin_popcount(m) < Kc && md >= 0; --md) m |= 1u << md;
      local.push_back(m);
    }
    unsigned bm = 0;
    best.clear();
    // the occupancy budget is a preference: a gate too large for it (compact
    // beamsplitter tables at high cutoff) gets a stage of the full budget
    for (int64_t bud : {budget, kStageLdsBudget}) {
      long bkey = -1;
      for (unsigned m : local) {
        absorb(p->ops, rem, m, D, Kc, tile_bytes, esz, bud, &got);
        // one step of lookahead: the gates this stage takes plus the most the
        // next stage could take after it (on c3: 49 stages instead of 52);
        // ties go to more gates now, then to tiles that include the last
        // modes (contiguous runs)
        long nxt = 0;
        if (!got.empty() && local.size() > 1 && got.size() < rem.size()) {
          rem2.clear();
          size_t gi = 0;
          for (int i : rem) {
            if (gi < got.size() && got[gi] == i)
              ++gi;
            else
              rem2.push_back(i);
          }
          for (unsigned m2 : local) {
            absorb(p->ops, rem2, m2, D, Kc, tile_bytes, esz, bud, &got2);
            if ((long)got2.size() > nxt) nxt = (long)got2.size();
          }
        }
        const long key = (((long)got.size() + nxt) << 40) + ((long)got.size() << 26) + (long)m;
        if (key > bkey) {
          bkey = key;
          bm = m;
          best = got;
        }
      }
      if (!best.empty()) break;
    }
    if (best.empty()) {
      // no stage holds the next operation: it runs on its own, unfused
      Step one;
      one.fused = 0;
      one.op = rem.front();
      one.K = 0;
      one.diag = 0;
      one.ops.push_back(rem.front());
      memset(&one.args, 0, sizeof(one.args));
      memset(&one.run, 0, sizeof(one.run));
      p->steps.push_back(one);
      rem.erase(rem.begin());
      continue;
    }
    Step st;
    st.fused = 1;
    st.op = best.front();
    st.K = Kc;
    st.diag = 0;
    st.ops = best;
    memset(&st.run, 0, sizeof(st.run));
    StageArgs& a = st.args;
    memset(&a, 0, sizeof(a));
    a.ngates = (int)best.size();
    int lmodes[4], nl = 0, nlc = 0;
    for (int md = 0; md < n; ++md) {
      if (bm >> md & 1u)
        lmodes[nl++] = md;
      else
        a.nl_stride[nlc++] = ipow(D, n - 1 - md);
    }
    a.nnl = nlc;
    a.ntiles = (int32_t)ipow(D, nlc);
    a.grid = 0;
    for (int t = 0; t < Kc; ++t) a.loc_stride[t] = ipow(D, n - 1 - lmodes[t]);
    int64_t mf = 0, ma = 0;
    for (int gi = 0; gi < a.ngates; ++gi) {
      const OpPlan& op = p->ops[best[gi]];
      StageGate& g = a.g[gi];
      g.kind = (int8_t)op.kind;
      g.a0 = g.a1 = 0;
      for (int t = 0; t < Kc; ++t) {
        if (lmodes[t] == op.m0) g.a0 = (int8_t)t;
        if (op.kind == SFAMD_OP_BEAMSPLITTER && lmodes[t] == op.m1) g.a1 = (int8_t)t;
      }
      g.op = best[gi];
      g.s0 = (int32_t)ipow(D, Kc - 1 - g.a0);
      g.s1 = (int32_t)ipow(D, Kc - 1 - g.a1);
      // one region per gate, 16-byte aligned in LDS: the matrix in the
      // forward kernel, matrix + derivative in the adjoint
      g.moff = (int32_t)mf;
      g.amoff = (int32_t)ma;
      mf += (stage_gate_elems(op.kind, D, false) + 1) & ~(int64_t)1;
      ma += (stage_gate_elems(op.kind, D, true) + 1) & ~(int64_t)1;
    }
    a.mat_fwd = (int32_t)mf;
    a.mat_adj = (int32_t)ma;
    p->steps.push_back(st);
    std::vector<int> nrem;
    size_t bi = 0;
    for (int i : rem) {
      if (bi < best.size() && best[bi] == i)
        ++bi;
      else
        nrem.push_back(i);
    }
    rem.swap(nrem);
  }
}

void plan_steps(sfamd_prog* p) {
  const int n = p->n, D = p->D;
  const int Kc = p->fuse_K < n ? p->fuse_K : n;
  const int64_t esz = elem_size(p->dtype);
  // fusion needs compiled stage kernels and room for at least one gate of any kind
  int gmax = 0;
  if (Kc >= 2 && n <= kStageMaxModes && stage_supported(D, Kc)) {
    const int64_t one = 4 * stage_tile(D, Kc) * esz + stage_gate_elems(SFAMD_OP_BEAMSPLITTER, D, true) * esz + 256;
    gmax = one <= kStageLdsBudget ? 1 : 0;
  }
  std::vector<int> run;
  auto flush = [&]() {
    if (!run.empty()) plan_run(p, run, Kc);
    run.clear();
  };
  auto is_diag = [&](int i) {
    const int k = p->ops[i].kind;
    return k == SFAMD_OP_ROTATION || k == SFAMD_OP_KERR;
  };
  for (int i = 0; i < p->nops; ++i) {
    if (gmax >= 1 && fusable(p->ops[i])) {
      run.push_back(i);
      continue;
    }
    flush();
    // consecutive rotation / Kerr gates on the same representation share one pass
    if (is_diag(i) && !p->steps.empty()) {
      Step& last = p->steps.back();
      if (!last.fused && is_diag(last.ops.back()) && p->ops[last.ops.back()].mixed_in == p->ops[i].mixed_in &&
          (int)last.ops.size() < kDiagRunMax) {
        last.ops.push_back(i);
        last.diag = 1;
        continue;
      }
    }
    Step st;
    st.fused = 0;
    st.op = i;
    st.K = 0;
    st.diag = 0;
    st.ops.push_back(i);
    memset(&st.args, 0, sizeof(st.args));
    memset(&st.run, 0, sizeof(st.run));
    p->steps.push_back(st);
  }
  flush();
  for (Step& st : p->steps) {
    if (!st.diag) continue;
    DiagRun& r = st.run;
    const OpPlan& first = p->ops[st.ops.front()];
    r.n = n;
    r.D = D;
    r.mixed = first.mixed_in;
    r.ngates = (int)st.ops.size();
    r.N = ipow(D, r.mixed ? 2 * n : n);
    const int64_t R = r.mixed ? (int64_t)D * D : D;
    for (int g = 0; g < r.ngates; ++g) {
      const OpPlan& op = p->ops[st.ops[g]];
      r.mode[g] = op.m0;
      r.kind[g] = op.kind;
      r.op[g] = st.ops[g];
      r.toff[g] = op.off;
      int64_t dv = 1;
      for (int t = 0; t < n - 2 - op.m0; ++t) dv *= R;
      r.rowdiv[g] = dv;
    }
  }
  // state sizes at step inputs
  p->state_elems.clear();
  for (const Step& st : p->steps) {
    const OpPlan& op = p->ops[st.ops.front()];
    p->state_elems.push_back(ipow(D, op.mixed_in ? 2 * n : n));
  }
}

}  // namespace

// ====================================================================== C ABI
extern "C" int sfamd_version(void) { return SFAMD_VERSION; }

extern "C" const char* sfamd_status_string(int code) {
  switch (code) {
    case 0:
      return "success";
    case SFAMD_E_ARG:
      return "invalid argument";
    case SFAMD_E_CUTOFF:
      return "cutoff dimension outside the supported range";
    case SFAMD_E_KIND:
      return "unknown operation kind";
    case SFAMD_E_STATE:
      return "operation not valid for this state representation";
    case SFAMD_E_WORKSPACE:
      return "workspace missing or too small";
    default:
      break;
  }
  if (code > 0) return hipGetErrorString((hipError_t)code);
  return "unknown status";
}

extern "C" int64_t sfamd_gate_table_size(int kind, int D) {
  if (D < 1 || D > SFAMD_MAX_CUTOFF) return -1;
  if (kind < 0 || kind > SFAMD_OP_CUBIC_PHASE) return -1;
  if (two_mode_kind(kind) && D > SFAMD_MAX_CUTOFF_BS) return -1;
  if (kind == SFAMD_OP_CUBIC_PHASE && D > SFAMD_MAX_CUTOFF_V) return -1;
  return table_size(kind, D);
}

extern "C" int sfamd_build_gate(int kind, double p0, double p1, int D, int dtype, void* out, void* stream) {
  CHECK_ARG(out && (dtype == SFAMD_C64 || dtype == SFAMD_C128));
  RET_IF(validate_op(kind, 0, 1, 2, D));
  if (kind == SFAMD_OP_MIX) return SFAMD_E_KIND;
  BuildBatch b;
  b.first = -1;
  b.p0 = p0;
  b.p1 = p1;
  b.d[0].kind = kind;
  b.d[0].off = 0;
  b.d[0].img_fwd = b.d[0].img_adj = -1;
  return launch_build(b, 1, D, nullptr, out, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_apply_1mode(int kind, const void* in, void* out, void* tmp, int n, int D, int mode, int mixed,
                                 const void* table, int dtype, void* stream) {
  CHECK_ARG(in && out && table && in != out && n >= 1);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  if (kind != SFAMD_OP_ROTATION && kind != SFAMD_OP_KERR && kind != SFAMD_OP_DISPLACEMENT &&
      kind != SFAMD_OP_SQUEEZE && kind != SFAMD_OP_CUBIC_PHASE)
    return SFAMD_E_KIND;
  RET_IF(validate_op(kind, mode, -1, n, D));
  OpPlan op{kind, mode, -1, mixed ? 1 : 0, 0, 0};
  return apply_op(op, n, D, table, in, out, tmp, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_apply_matrix(const void* in, void* out, void* tmp, int n, int D, int mode, int mixed,
                                  const void* U, int dtype, void* stream) {
  CHECK_ARG(in && out && U && in != out && n >= 1);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  RET_IF(validate_op(SFAMD_OP_DISPLACEMENT, mode, -1, n, D));
  // the dense single-mode path reads only the leading D x D matrix of a table
  OpPlan op{SFAMD_OP_DISPLACEMENT, mode, -1, mixed ? 1 : 0, 0, 0};
  return apply_op(op, n, D, U, in, out, tmp, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_apply_bs(const void* in, void* out, void* tmp, int n, int D, int mode1, int mode2, int mixed,
                              const void* table, int dtype, void* stream) {
  CHECK_ARG(in && out && table && in != out && n >= 2);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  RET_IF(validate_op(SFAMD_OP_BEAMSPLITTER, mode1, mode2, n, D));
  OpPlan op{SFAMD_OP_BEAMSPLITTER, mode1, mode2, mixed ? 1 : 0, 0, 0};
  return apply_op(op, n, D, table, in, out, tmp, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_apply_2mode(int kind, const void* in, void* out, void* tmp, int n, int D, int mode1, int mode2,
                                 int mixed, const void* table, int dtype, void* stream) {
  CHECK_ARG(in && out && table && in != out && n >= 2);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  if (!two_mode_kind(kind)) return SFAMD_E_KIND;
  RET_IF(validate_op(kind, mode1, mode2, n, D));
  OpPlan op{kind, mode1, mode2, mixed ? 1 : 0, 0, 0};
  return apply_op(op, n, D, table, in, out, tmp, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_apply_loss(const void* in, void* out, int n, int D, int mode, const void* table, int dtype,
                                void* stream) {
  CHECK_ARG(in && out && table && in != out && n >= 1);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  RET_IF(validate_op(SFAMD_OP_LOSS, mode, -1, n, D));
  OpPlan op{SFAMD_OP_LOSS, mode, -1, 1, 0, 0};
  return apply_op(op, n, D, table, in, out, nullptr, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_mix(const void* ket, void* dm, int n, int D, int dtype, void* stream) {
  CHECK_ARG(ket && dm && n >= 1 && D >= 1);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  return mix_fwd(ket, dm, n, D, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_replace_mode(const void* in, void* out, int n, int D, int mode, const void* sigma, int dtype,
                                  void* stream) {
  CHECK_ARG(in && out && sigma && in != out && n >= 1 && D >= 1 && mode >= 0 && mode < n);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  return replace_mode(in, out, ipow(D, 2 * n), D, bra_stride_dm(n, D, mode), ket_stride_dm(n, D, mode), sigma, dtype,
                      (hipStream_t)stream);
}

// ---------------------------------------------------------------- programs
extern "C" int sfamd_prog_create(const sfamd_op_t* ops, int nops, int n, int D, int mixed_in, int dtype, int fuse_K,
                                 sfamd_prog_t** out) {
  CHECK_ARG(out && nops >= 0 && n >= 1 && n <= 32 && (ops || nops == 0));
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  if (D < 1 || D > SFAMD_MAX_CUTOFF) return SFAMD_E_CUTOFF;
  sfamd_prog* p = new (std::nothrow) sfamd_prog();
  if (!p) return SFAMD_E_ARG;
  p->n = n;
  p->D = D;
  p->dtype = dtype;
  p->mixed_in = mixed_in ? 1 : 0;
  p->nops = nops;
  p->fuse_K = fuse_K;
  int mixed = p->mixed_in;
  int64_t off = 0, max_state = ipow(D, mixed ? 2 * n : n);
  for (int i = 0; i < nops; ++i) {
    const int st = validate_op(ops[i].kind, ops[i].mode0, ops[i].mode1, n, D);
    if (st != 0) {
      delete p;
      return st;
    }
    OpPlan op{ops[i].kind, ops[i].mode0, ops[i].mode1, mixed, 0, off};
    if (op.kind == SFAMD_OP_MIX && mixed) {
      delete p;
      return SFAMD_E_STATE;
    }
    if (op.kind == SFAMD_OP_LOSS && !mixed) op.mix_first = 1;
    if (op.kind == SFAMD_OP_LOSS || op.kind == SFAMD_OP_MIX) mixed = 1;
    off += table_size(op.kind, D);
    p->ops.push_back(op);
    const int64_t se = ipow(D, mixed ? 2 * n : n);
    if (se > max_state) max_state = se;
    // a two-pass op on a DM and a pure->mixed loss need DM-sized scratch
    if (op.mix_first) max_state = ipow(D, 2 * n) > max_state ? ipow(D, 2 * n) : max_state;
  }
  p->max_state = max_state;
  p->final_mixed = mixed;
  plan_steps(p);
  // the fused stages' matrix images follow the operations' entries
  p->img_fwd.assign(nops, -1);
  p->img_adj.assign(nops, -1);
  for (Step& st : p->steps) {
    if (!st.fused) continue;
    StageArgs& a = st.args;
    a.img_fwd = off;
    off += a.mat_fwd;
    a.img_adj = off;
    off += a.mat_adj;
    for (int gi = 0; gi < a.ngates; ++gi) {
      p->img_fwd[a.g[gi].op] = (int32_t)(a.img_fwd + a.g[gi].moff);
      p->img_adj[a.g[gi].op] = (int32_t)(a.img_adj + a.g[gi].amoff);
    }
  }
  p->table_elems = off;
  p->state_elems.push_back(ipow(D, mixed ? 2 * n : n));
  int64_t nwg = (max_state + kBlock - 1) / kBlock;
  if (nwg > 4096) nwg = 4096;
  if (nwg < 1) nwg = 1;
  p->nwg = (int)nwg;
  *out = p;
  return 0;
}

extern "C" int sfamd_prog_info(const sfamd_prog_t* p, sfamd_prog_info_t* info) {
  CHECK_ARG(p && info);
  info->table_elems = p->table_elems;
  info->partial_doubles = (int64_t)p->nops * 4 * p->nwg;
  info->ws_elems = 4 * p->max_state;
  info->final_elems = p->state_elems.back();
  info->final_mixed = p->final_mixed;
  info->nops = p->nops;
  info->nsteps = (int32_t)p->steps.size();
  return 0;
}

extern "C" int64_t sfamd_prog_state_elems(const sfamd_prog_t* p, int i) {
  if (!p || i < 0 || i > (int)p->steps.size()) return -1;
  return p->state_elems[i];
}

extern "C" int sfamd_prog_step_ops(const sfamd_prog_t* p, int step, int32_t* ops_out, int cap) {
  if (!p || step < 0 || step >= (int)p->steps.size()) return SFAMD_E_ARG;
  const Step& st = p->steps[step];
  for (int i = 0; i < (int)st.ops.size() && i < cap; ++i) ops_out[i] = st.ops[i];
  return (int)st.ops.size();
}

extern "C" int sfamd_prog_destroy(sfamd_prog_t* p) {
  delete p;
  return 0;
}

extern "C" int sfamd_prog_build_table(const sfamd_prog_t* p, const double* params, void* table, void* stream) {
  CHECK_ARG(p && (p->nops == 0 || (params && table)));
  CHECK_ARG(p->table_elems < ((int64_t)1 << 31));
  BuildBatch b;
  b.p0 = b.p1 = 0.0;
  for (int first = 0; first < p->nops; first += kMaxBuildPerLaunch) {
    const int cnt = p->nops - first < kMaxBuildPerLaunch ? p->nops - first : kMaxBuildPerLaunch;
    b.first = first;
    for (int i = 0; i < cnt; ++i) {
      b.d[i].kind = p->ops[first + i].kind;  // MIX entries build nothing
      b.d[i].off = (int32_t)p->ops[first + i].off;
      b.d[i].img_fwd = p->img_fwd[first + i];
      b.d[i].img_adj = p->img_adj[first + i];
    }
    RET_IF(launch_build(b, cnt, p->D, params, table, p->dtype, (hipStream_t)stream));
  }
  return 0;
}

extern "C" int sfamd_prog_forward(const sfamd_prog_t* p, int begin, int end, const void* table, const void* in,
                                  void* out, void* const* saved, void* ws, void* stream) {
  const int nsteps = p ? (int)p->steps.size() : 0;
  CHECK_ARG(p && in && out && in != out && 0 <= begin && begin <= end && end <= nsteps);
  CHECK_ARG(ws || begin == end);
  hipStream_t s = (hipStream_t)stream;
  const int64_t esz = elem_size(p->dtype);
  char* A = (char*)ws;
  char* B = A + p->max_state * esz;
  char* Ct = B + p->max_state * esz;
  if (begin == end) {
    HIP_RET(hipMemcpyAsync(out, in, p->state_elems[begin] * esz, hipMemcpyDeviceToDevice, s));
    return 0;
  }
  const void* cur = in;
  if (saved && saved[0] && saved[0] != in)
    HIP_RET(hipMemcpyAsync(saved[0], in, p->state_elems[begin] * esz, hipMemcpyDeviceToDevice, s));
  for (int i = begin; i < end; ++i) {
    void* dst;
    if (i == end - 1)
      dst = out;
    else if (saved && saved[i + 1 - begin])
      dst = saved[i + 1 - begin];
    else
      dst = (cur == A) ? B : A;
    if (dst == cur) return SFAMD_E_ARG;
    const Step& st = p->steps[i];
    if (st.fused) {
      int grid = (int)(st.args.ntiles < p->nwg ? st.args.ntiles : p->nwg);
      RET_IF(stage_fwd(p->D, st.K, cur, dst, table, st.args, grid, p->dtype, s));
    } else if (st.diag) {
      RET_IF(diag_run_fwd(cur, dst, table, st.run, p->dtype, s));
    } else {
      RET_IF(apply_op(p->ops[st.op], p->n, p->D, table, cur, dst, Ct, p->dtype, s));
    }
    cur = dst;
  }
  return 0;
}

extern "C" int sfamd_prog_adjoint(const sfamd_prog_t* p, int begin, int end, const void* table, void* const* saved,
                                  const void* g_out, void* g_in, double* grads, double* partial, void* ws,
                                  void* stream) {
  const int nsteps = p ? (int)p->steps.size() : 0;
  CHECK_ARG(p && saved && g_out && g_in && grads && partial && ws && g_in != g_out);
  CHECK_ARG(0 <= begin && begin < end && end <= nsteps);
  hipStream_t s = (hipStream_t)stream;
  const int64_t esz = elem_size(p->dtype);
  char* LA = (char*)ws;
  char* LB = LA + p->max_state * esz;
  char* T1 = LB + p->max_state * esz;
  char* T2 = T1 + p->max_state * esz;
  const int nwg = p->nwg;
  // every adjoint launch writes its operation's whole [pass][slot][nwg]
  // partial block, so nothing needs clearing between calls
  const void* lam = g_out;
  for (int i = end - 1; i >= begin; --i) {
    if (!saved[i - begin]) return SFAMD_E_ARG;
    void* dst = (i == begin) ? g_in : ((lam == LA) ? LB : LA);
    const Step& st = p->steps[i];
    if (st.fused) {
      // one workgroup per tile at most, one partial column per workgroup
      const int grid = (int)(st.args.ntiles < nwg ? st.args.ntiles : nwg);
      RET_IF(stage_adj(p->D, st.K, saved[i - begin], lam, dst, table, st.args, grid, partial, nwg, p->dtype, s));
    } else if (st.diag) {
      RET_IF(diag_run_adj(saved[i - begin], lam, dst, table, st.run, partial, nwg, p->dtype, s));
    } else {
      RET_IF(adjoint_op(p->ops[st.op], p->n, p->D, table, saved[i - begin], lam, dst, T1, T2,
                        partial + (int64_t)st.op * 4 * nwg, nwg, p->dtype, s));
    }
    lam = dst;
  }
  // gradients of the operations of these steps, one launch per OpList chunk
  OpList ol;
  ol.n = 0;
  for (int i = begin; i < end; ++i) {
    const Step& st = p->steps[i];
    const int ncol = st.fused ? (int)(st.args.ntiles < nwg ? st.args.ntiles : nwg) : nwg;
    for (int op : st.ops) {
      ol.cols[ol.n] = (uint16_t)(st.fused ? ((unsigned)ncol | kOpColsOnePass) : (unsigned)ncol);
      ol.ops[ol.n++] = op;
      if (ol.n == kOpListMax) {
        RET_IF(reduce_grads(partial, grads, ol, nwg, s));
        ol.n = 0;
      }
    }
  }
  return reduce_grads(partial, grads, ol, nwg, s);
}

// ---------------------------------------------------------------- graphs
extern "C" int sfamd_graph_begin(void* stream) {
  CHECK_ARG(stream);
  HIP_RET(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
  return 0;
}

extern "C" int sfamd_graph_end(void* stream, void** graph_exec) {
  CHECK_ARG(stream && graph_exec);
  hipGraph_t g = nullptr;
  HIP_RET(hipStreamEndCapture((hipStream_t)stream, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_RET(e);
  *graph_exec = (void*)ex;
  return 0;
}

extern "C" int sfamd_graph_launch(void* graph_exec, void* stream) {
  CHECK_ARG(graph_exec);
  HIP_RET(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
  return 0;
}

extern "C" int sfamd_graph_destroy(void* graph_exec) {
  if (!graph_exec) return 0;
  HIP_RET(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return 0;
}
