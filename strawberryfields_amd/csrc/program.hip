// Recorded-circuit executor and the C ABI entry points of libsfamd.
//
// A program is the list of Fock-backend operations the Engine issued
// (strawberryfields/engine.py:427-430 calls op.apply per command); the host
// loop below issues the kernels of each operation on one stream, keeps the
// state in caller-provided buffers, writes the inputs of requested operations
// as checkpoints and runs the reverse sweep.  The whole sequence can be
// captured into a HIP graph by the caller (sfamd_graph_*).
#include <new>
#include <vector>

#include "common.h"

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

int validate_op(int kind, int m0, int m1, int n, int D) {
  if (kind < 0 || kind > SFAMD_OP_MIX) return SFAMD_E_KIND;
  if (D < 1 || D > SFAMD_MAX_CUTOFF) return SFAMD_E_CUTOFF;
  if (kind == SFAMD_OP_BEAMSPLITTER) {
    if (D > SFAMD_MAX_CUTOFF_BS) return SFAMD_E_CUTOFF;
    if (m0 < 0 || m0 >= n || m1 < 0 || m1 >= n || m0 == m1) return SFAMD_E_ARG;
  } else if (kind != SFAMD_OP_MIX) {
    if (m0 < 0 || m0 >= n) return SFAMD_E_ARG;
  }
  return 0;
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
      if (op.mixed_in) {
        if (!tmp) return SFAMD_E_WORKSPACE;
        RET_IF(axis_dense(in, tmp, NM, D, bra_stride_dm(n, D, op.m0), g, 0, dtype, s));
        return axis_dense(tmp, out, NM, D, ket_stride_dm(n, D, op.m0), g, 1, dtype, s);
      }
      return axis_dense(in, out, NK, D, ket_stride(n, D, op.m0), g, 0, dtype, s);
    case SFAMD_OP_BEAMSPLITTER:
      if (op.mixed_in) {
        if (!tmp) return SFAMD_E_WORKSPACE;
        RET_IF(pair_bs(in, tmp, NM, D, bra_stride_dm(n, D, op.m0), bra_stride_dm(n, D, op.m1), g, 0, dtype, s));
        return pair_bs(tmp, out, NM, D, ket_stride_dm(n, D, op.m0), ket_stride_dm(n, D, op.m1), g, 1, dtype, s);
      }
      return pair_bs(in, out, NK, D, ket_stride(n, D, op.m0), ket_stride(n, D, op.m1), g, 0, dtype, s);
    case SFAMD_OP_LOSS: {
      const void* src = in;
      if (op.mix_first) {
        if (!tmp) return SFAMD_E_WORKSPACE;
        RET_IF(mix_fwd(in, tmp, n, D, dtype, s));
        src = tmp;
      }
      return loss_fwd(src, out, NM, D, bra_stride_dm(n, D, op.m0), ket_stride_dm(n, D, op.m0), g, dtype, s);
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
                             pass0, nwg, dtype, s);
      return axis_diag_adj(psi, lo, li, NK, D, ket_stride(n, D, op.m0), g, kerr, 0, pass0, nwg, dtype, s);
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE: {
      const int64_t DD = (int64_t)D * D;
      const void* dU0 = cptr(g, DD, dtype);
      const void* dU1 = cptr(g, 2 * DD, dtype);
      if (op.mixed_in) {
        const int64_t sb = bra_stride_dm(n, D, op.m0), sk = ket_stride_dm(n, D, op.m0);
        // rho1 = U rho (bra side) recomputed; then the ket-side pass, then the bra-side pass
        RET_IF(axis_dense(psi, t1, NM, D, sb, g, 0, dtype, s));
        RET_IF(axis_dense_adj(t1, lo, t2, NM, D, sk, g, dU0, dU1, 1, pass1, nwg, dtype, s));
        return axis_dense_adj(psi, t2, li, NM, D, sb, g, dU0, dU1, 0, pass0, nwg, dtype, s);
      }
      return axis_dense_adj(psi, lo, li, NK, D, ket_stride(n, D, op.m0), g, dU0, dU1, 0, pass0, nwg, dtype, s);
    }
    case SFAMD_OP_BEAMSPLITTER: {
      const void* dB = cptr(g, bs_nb(D), dtype);
      if (op.mixed_in) {
        const int64_t bp = bra_stride_dm(n, D, op.m0), bq = bra_stride_dm(n, D, op.m1);
        const int64_t kp = ket_stride_dm(n, D, op.m0), kq = ket_stride_dm(n, D, op.m1);
        RET_IF(pair_bs(psi, t1, NM, D, bp, bq, g, 0, dtype, s));
        RET_IF(pair_bs_adj(t1, lo, t2, NM, D, kp, kq, g, dB, 1, pass1, nwg, dtype, s));
        return pair_bs_adj(psi, t2, li, NM, D, bp, bq, g, dB, 0, pass0, nwg, dtype, s);
      }
      return pair_bs_adj(psi, lo, li, NK, D, ket_stride(n, D, op.m0), ket_stride(n, D, op.m1), g, dB, 0, pass0, nwg,
                         dtype, s);
    }
    case SFAMD_OP_LOSS: {
      const int64_t D3 = (int64_t)D * D * D;
      const char* c = (const char*)g;
      const char* dc = c + D3 * (elem_size(dtype) / 2);
      const int64_t sb = bra_stride_dm(n, D, op.m0), sk = ket_stride_dm(n, D, op.m0);
      if (op.mix_first) {
        RET_IF(mix_fwd(psi, t1, n, D, dtype, s));
        RET_IF(loss_adj(t1, lo, t2, NM, D, sb, sk, c, dc, pass0, nwg, dtype, s));
        return mix_adj(psi, t2, li, n, D, dtype, s);
      }
      return loss_adj(psi, lo, li, NM, D, sb, sk, c, dc, pass0, nwg, dtype, s);
    }
    case SFAMD_OP_MIX:
      return mix_adj(psi, lo, li, n, D, dtype, s);
    default:
      return SFAMD_E_KIND;
  }
}

}  // namespace

struct sfamd_prog {
  int n, D, dtype, mixed_in, nops, nwg;
  std::vector<OpPlan> ops;
  std::vector<int64_t> state_elems;  // elements at the input of op i (i = nops: output)
  int64_t table_elems, max_state;
  int final_mixed;
};

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
  if (kind == SFAMD_OP_BEAMSPLITTER && D > SFAMD_MAX_CUTOFF_BS) return -1;
  return table_size(kind, D);
}

extern "C" int sfamd_build_gate(int kind, double p0, double p1, int D, int dtype, void* out, void* stream) {
  CHECK_ARG(out && (dtype == SFAMD_C64 || dtype == SFAMD_C128));
  RET_IF(validate_op(kind, 0, 1, 2, D));
  if (kind == SFAMD_OP_MIX) return SFAMD_E_KIND;
  BuildBatch b;
  b.d[0].kind = kind;
  b.d[0].pidx = -1;
  b.d[0].off = 0;
  b.d[0].p0 = p0;
  b.d[0].p1 = p1;
  return launch_build(b, 1, D, nullptr, out, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_apply_1mode(int kind, const void* in, void* out, void* tmp, int n, int D, int mode, int mixed,
                                 const void* table, int dtype, void* stream) {
  CHECK_ARG(in && out && table && in != out && n >= 1);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  if (kind != SFAMD_OP_ROTATION && kind != SFAMD_OP_KERR && kind != SFAMD_OP_DISPLACEMENT &&
      kind != SFAMD_OP_SQUEEZE)
    return SFAMD_E_KIND;
  RET_IF(validate_op(kind, mode, -1, n, D));
  OpPlan op{kind, mode, -1, mixed ? 1 : 0, 0, 0};
  return apply_op(op, n, D, table, in, out, tmp, dtype, (hipStream_t)stream);
}

extern "C" int sfamd_apply_bs(const void* in, void* out, void* tmp, int n, int D, int mode1, int mode2, int mixed,
                              const void* table, int dtype, void* stream) {
  CHECK_ARG(in && out && table && in != out && n >= 2);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  RET_IF(validate_op(SFAMD_OP_BEAMSPLITTER, mode1, mode2, n, D));
  OpPlan op{SFAMD_OP_BEAMSPLITTER, mode1, mode2, mixed ? 1 : 0, 0, 0};
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
extern "C" int sfamd_prog_create(const sfamd_op_t* ops, int nops, int n, int D, int mixed_in, int dtype,
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
  int mixed = p->mixed_in;
  int64_t off = 0, max_state = ipow(D, mixed ? 2 * n : n);
  p->state_elems.push_back(max_state);
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
    p->state_elems.push_back(se);
    if (se > max_state) max_state = se;
    // a two-pass op on a DM and a pure->mixed loss need DM-sized scratch
    if (op.mix_first) max_state = ipow(D, 2 * n) > max_state ? ipow(D, 2 * n) : max_state;
  }
  p->table_elems = off;
  p->max_state = max_state;
  p->final_mixed = mixed;
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
  return 0;
}

extern "C" int64_t sfamd_prog_state_elems(const sfamd_prog_t* p, int i) {
  if (!p || i < 0 || i > p->nops) return -1;
  return p->state_elems[i];
}

extern "C" int sfamd_prog_destroy(sfamd_prog_t* p) {
  delete p;
  return 0;
}

extern "C" int sfamd_prog_build_table(const sfamd_prog_t* p, const double* params, void* table, void* stream) {
  CHECK_ARG(p && (p->nops == 0 || (params && table)));
  BuildBatch b;
  int cnt = 0;
  for (int i = 0; i < p->nops; ++i) {
    const OpPlan& op = p->ops[i];
    if (op.kind == SFAMD_OP_MIX) continue;
    b.d[cnt].kind = op.kind;
    b.d[cnt].pidx = i;
    b.d[cnt].off = op.off;
    b.d[cnt].p0 = 0.0;
    b.d[cnt].p1 = 0.0;
    if (++cnt == kMaxBuildPerLaunch) {
      RET_IF(launch_build(b, cnt, p->D, params, table, p->dtype, (hipStream_t)stream));
      cnt = 0;
    }
  }
  return launch_build(b, cnt, p->D, params, table, p->dtype, (hipStream_t)stream);
}

extern "C" int sfamd_prog_forward(const sfamd_prog_t* p, int begin, int end, const void* table, const void* in,
                                  void* out, void* const* saved, void* ws, void* stream) {
  CHECK_ARG(p && in && out && in != out && 0 <= begin && begin <= end && end <= p->nops);
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
    RET_IF(apply_op(p->ops[i], p->n, p->D, table, cur, dst, Ct, p->dtype, s));
    cur = dst;
  }
  return 0;
}

extern "C" int sfamd_prog_adjoint(const sfamd_prog_t* p, int begin, int end, const void* table, void* const* saved,
                                  const void* g_out, void* g_in, double* grads, double* partial, void* ws,
                                  void* stream) {
  CHECK_ARG(p && saved && g_out && g_in && grads && partial && ws && g_in != g_out);
  CHECK_ARG(0 <= begin && begin < end && end <= p->nops);
  hipStream_t s = (hipStream_t)stream;
  const int64_t esz = elem_size(p->dtype);
  char* LA = (char*)ws;
  char* LB = LA + p->max_state * esz;
  char* T1 = LB + p->max_state * esz;
  char* T2 = T1 + p->max_state * esz;
  const int nwg = p->nwg;
  RET_IF(fill_zero(partial + (int64_t)begin * 4 * nwg, (int64_t)(end - begin) * 4 * nwg * sizeof(double), s));
  const void* lam = g_out;
  for (int i = end - 1; i >= begin; --i) {
    if (!saved[i - begin]) return SFAMD_E_ARG;
    void* dst = (i == begin) ? g_in : ((lam == LA) ? LB : LA);
    RET_IF(adjoint_op(p->ops[i], p->n, p->D, table, saved[i - begin], lam, dst, T1, T2,
                      partial + (int64_t)i * 4 * nwg, nwg, p->dtype, s));
    lam = dst;
  }
  return reduce_grads(partial, grads, begin, end, nwg, s);
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
