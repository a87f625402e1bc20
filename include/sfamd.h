/*
 * sfamd.h -- C ABI of the MI355X Fock-basis gate-application library.
 *
 * The library applies Strawberry Fields' Fock-backend operations (rotation,
 * Kerr, displacement, squeezing, beamsplitter, loss) to a truncated Fock-basis
 * state held in device memory, reads probabilities and reduced density
 * matrices out of it, and runs the adjoint (reverse-mode) sweep of a recorded
 * circuit.  It replaces the arithmetic of
 *   strawberryfields/backends/fockbackend/circuit.py (Circuit)  and
 *   strawberryfields/backends/fockbackend/ops.py     (gate matrices)
 * and, for gradients, the tape of strawberryfields/backends/tfbackend.
 *
 * Conventions (all functions):
 *   - Every function returns int: 0 = success, a positive hipError_t value on a
 *     HIP runtime error, or a negative SFAMD_E_* code on invalid arguments.
 *     No function throws across the ABI.  sfamd_status_string() names a code.
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     All work is enqueued asynchronously on that stream.
 *   - Buffers are device pointers owned by the caller.  The library never
 *     allocates device memory; scratch space is caller-provided (except for
 *     sfamd_fidelity_loss_noalloc, whose scratch is a static per-GPU array).
 *   - `dtype` selects the amplitude type: SFAMD_C64 (interleaved float re,im)
 *     or SFAMD_C128 (interleaved double re,im).
 *   - A state of `n` modes and cutoff `D` is a ket of n axes [D]*n in C order
 *     (mode 0 most significant, fockbackend/ops.py:74-78) or a density matrix
 *     of 2n axes interleaved [i0,j0,i1,j1,...] (fockbackend/ops.py:110-120).
 *     The kernels see a tensor of R = n (ket) or 2n (DM) axes of extent D.
 */
#ifndef SFAMD_H
#define SFAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants ---------------------------------------------------------- */
enum {
  SFAMD_C64 = 0,
  SFAMD_C128 = 1
};

/* operation kinds (gate families of fockbackend/backend.py) */
enum {
  SFAMD_OP_ROTATION = 0,     /* R(theta)      fockbackend/ops.py:309-314       */
  SFAMD_OP_KERR = 1,         /* K(kappa)      fockbackend/ops.py:274-281       */
  SFAMD_OP_DISPLACEMENT = 2, /* D(r, phi)     fockbackend/ops.py:219-235       */
  SFAMD_OP_SQUEEZE = 3,      /* S(r, phi)     fockbackend/ops.py:238-254       */
  SFAMD_OP_BEAMSPLITTER = 4, /* BS(theta,phi) fockbackend/ops.py:318-329       */
  SFAMD_OP_LOSS = 5,         /* N(T)          fockbackend/ops.py:471-490       */
  SFAMD_OP_MIX = 6,          /* pure -> mixed fockbackend/ops.py:110-120       */
  SFAMD_OP_TWO_MODE_SQUEEZE = 7, /* S2(r, phi)  fockbackend/ops.py:257-271,
                                    circuit.py:572-577 (_apply_S2 337-365)      */
  SFAMD_OP_MZGATE = 8,       /* MZ(phi_in, phi_ex) fockbackend/ops.py:332-343  */
  SFAMD_OP_CROSS_KERR = 9,   /* CK(kappa)     fockbackend/ops.py:284-294       */
  SFAMD_OP_CUBIC_PHASE = 10  /* V(gamma), hbar = 2  fockbackend/ops.py:297-306 */
};

/* error codes (negative) */
enum {
  SFAMD_E_ARG = -1,       /* invalid argument (sizes, modes, dtype)          */
  SFAMD_E_CUTOFF = -2,    /* cutoff outside the supported range              */
  SFAMD_E_KIND = -3,      /* unknown operation kind                          */
  SFAMD_E_STATE = -4,     /* operation not valid for this state kind         */
  SFAMD_E_WORKSPACE = -5  /* caller-provided workspace too small             */
};

/* maximum cutoff accepted by each gate family */
#define SFAMD_MAX_CUTOFF 64      /* R, K, D, S, loss                          */
#define SFAMD_MAX_CUTOFF_BS 40   /* beamsplitter, S2, MZ, CK (two FP64 blocks in LDS) */
#define SFAMD_MAX_CUTOFF_V 32    /* cubic phase (FP64 matrix exponential in LDS) */

/* ---- library ------------------------------------------------------------ */
int sfamd_version(void);
const char* sfamd_status_string(int code);

/* Size, in complex elements of `dtype`, of the gate-table entry of one
 * operation of `kind` at cutoff D (matrix plus its parameter derivatives). */
int64_t sfamd_gate_table_size(int kind, int D);

/* ---- gate matrices (FP64 recurrences, rounded once to dtype) ------------ */
/* Builds the Fock-basis matrix of one gate and its derivatives with respect to
 * its two real parameters into `out` (sfamd_gate_table_size(kind, D) complex
 * elements).  Replaces thewalrus.fock_gradients.{displacement,squeezing,
 * beamsplitter} called at fockbackend/ops.py:233,252,326 and the analytic
 * R/K/loss matrices of fockbackend/ops.py:274-281,309-314,471-490.
 * Layout of `out`:
 *   R, K : diag[D]
 *   D, S : U[D][D] (row = out index), dU/dp0[D][D], dU/dp1[D][D]
 *   BS   : photon-number blocks s = 0..2D-2, block s holds B[N][n]
 *          (out1 = N, in1 = n, out2 = s-N, in2 = s-n, lo(s) <= N,n <= hi(s)),
 *          followed by the same blocks of dB/dtheta.
 *   loss : real c[p][q][l] then real dc/dT[p][q][l], packed two reals per
 *          complex slot; rho'[p,q] = sum_l c[p][q][l] rho[p+l][q+l].
 *   MZ, CK: photon-number blocks as BS (both conserve N1 + N2), then the same
 *          blocks of dG/dp0 (p0 = phi_in, kappa).  The second parameter's
 *          derivative is i w G[N][n] with w = n (MZ: phi_ex rotates input
 *          mode 1) and none for CK; for BS it is w = n - N.
 *   S2   : blocks of constant out1 - out2 = in1 - in2 = d, stored as BS block
 *          s = D - 1 + d holding G[N][n] (out1 = N, in1 = n, out2 = N - d,
 *          in2 = n - d), then the blocks of dG/dr; dG/dphi = i (N - n) G.
 *   V    : U[D][D], dU/dgamma[D][D], zeros[D][D] (as D, S).
 */
int sfamd_build_gate(int kind, double p0, double p1, int D, int dtype,
                     void* out, void* stream);

/* ---- single operations on a state --------------------------------------- */
/* Dense or diagonal single-mode gate.  Replaces Circuit.apply_gate_BLAS
 * (fockbackend/circuit.py:118-217) for one mode.  `table` is the gate-table
 * entry of a R/K/D/S/V gate.  For a DM (mixed=1) the gate is applied on both
 * sides, U rho U^dagger.  `tmp` (same size as the state) is needed for a dense
 * gate on a DM and may be NULL otherwise.  in == out is not allowed. */
int sfamd_apply_1mode(int kind, const void* in, void* out, void* tmp,
                      int n, int D, int mode, int mixed, const void* table,
                      int dtype, void* stream);

/* Beamsplitter on (mode1, mode2); the matrix's first pair acts on mode1.
 * Replaces Circuit.apply_twomode_gate / _apply_two_mode_passive
 * (fockbackend/circuit.py:219-335).  `tmp` as above for a DM. */
int sfamd_apply_bs(const void* in, void* out, void* tmp, int n, int D,
                   int mode1, int mode2, int mixed, const void* table,
                   int dtype, void* stream);

/* Any two-mode gate (kind BS, S2, MZ or CK) on (mode1, mode2), the table's
 * first pair on mode1.  Replaces Circuit.apply_twomode_gate with the
 * BSgate / MZgate / S2gate branches (fockbackend/circuit.py:219-304) and the
 * cross-Kerr apply_gate_BLAS call (circuit.py:586-591). */
int sfamd_apply_2mode(int kind, const void* in, void* out, void* tmp, int n,
                      int D, int mode1, int mode2, int mixed,
                      const void* table, int dtype, void* stream);

/* Any single-mode matrix U[D][D] (row = output index, state dtype) on `mode`;
 * U rho U^dagger for a DM.  The generic case of Circuit.apply_gate_BLAS
 * (fockbackend/circuit.py:118-217), used by the homodyne conditional
 * projection (circuit.py:787-795).  `tmp` as for sfamd_apply_1mode. */
int sfamd_apply_matrix(const void* in, void* out, void* tmp, int n, int D,
                       int mode, int mixed, const void* U, int dtype,
                       void* stream);

/* Loss channel on a DM.  Replaces Circuit._apply_channel with the Kraus set of
 * fockbackend/ops.py:471-490 (fockbackend/circuit.py:65-87). */
int sfamd_apply_loss(const void* in, void* out, int n, int D, int mode,
                     const void* table, int dtype, void* stream);

/* Pure -> mixed: rho = psi (x) psi^*, interleaved axes (fockbackend/ops.py:110-120). */
int sfamd_mix(const void* ket, void* dm, int n, int D, int dtype, void* stream);

/* Replace one mode by the single-mode density matrix `sigma` [D][D]:
 * rho' = sigma (x) Tr_mode(rho).  Implements prepare_multimode for one mode
 * (fockbackend/circuit.py:393-473) on a DM. */
int sfamd_replace_mode(const void* in, void* out, int n, int D, int mode,
                       const void* sigma, int dtype, void* stream);

/* ---- readout (FP64 accumulation) ---------------------------------------- */
/* out_dev[0] = <psi|psi> (ket) or Tr rho (DM), as double.  `partial` must hold
 * sfamd_reduce_partials(n, D, mixed) doubles. */
int64_t sfamd_reduce_partials(int n, int D, int mixed);
int sfamd_norm(const void* state, int n, int D, int mixed, double* partial,
               double* out_dev, int dtype, void* stream);

/* probabilities of all Fock basis states, [D]*n, real of the state's precision
 * (BaseFockState.all_fock_probs, states.py:580-608) */
int sfamd_all_fock_probs(const void* state, void* probs, int n, int D, int mixed,
                         int dtype, void* stream);

/* reduced density matrix of the sorted `modes` (k of them), straight from the
 * ket or DM without materialising the full DM (states.py:613-642,
 * fockbackend/backend.py:209-264).  out: [D]*2k interleaved, complex dtype.
 * `partial` must hold sfamd_reduced_dm_partials(n, D, k, mixed) doubles. */
int64_t sfamd_reduced_dm_partials(int n, int D, int k, int mixed);
int sfamd_reduced_dm(const void* state, void* out, int n, int D, int mixed,
                     const int* modes, int k, double* partial, int dtype,
                     void* stream);

/* sum over basis states of p(I) * prod_{m in modes} values[I_m], the
 * expectation of a number-diagonal operator (states.py:947-980).  values has
 * D doubles.  Result written as double to out_dev[0]. */
int sfamd_diagonal_expectation(const void* state, int n, int D, int mixed,
                               const int* modes, int k, const double* values_dev,
                               double* partial, double* out_dev, int dtype,
                               void* stream);

/* cotangent of all_fock_probs: g_state from g_probs (real [D]*n). */
int sfamd_all_fock_probs_vjp(const void* state, const void* g_probs, void* g_state,
                             int n, int D, int mixed, int dtype, void* stream);

/* <a|b> over n_elems amplitudes with FP64 accumulation: out_dev[0] = Re,
 * out_dev[1] = Im, out_dev[2] = 1 - |<a|b>|^2.  `partial` holds
 * sfamd_overlap_partials(n_elems) doubles.  (BaseFockState.fidelity-style
 * readout, states.py:665-730, for a full-register target ket.) */
int64_t sfamd_overlap_partials(int64_t n_elems);
int sfamd_overlap(const void* a, const void* b, int64_t n_elems, double* partial,
                  double* out_dev, int dtype, void* stream);
/* As sfamd_overlap with a = target, b = psi, and, if g_psi != NULL, the
 * cotangent of L = 1 - |<target|psi>|^2: g_psi = -2 <target|psi> target. */
int sfamd_fidelity_loss(const void* target, const void* psi, int64_t n_elems,
                        double* partial, double* out_dev, void* g_psi,
                        int dtype, void* stream);

/* Allocation-free variant of sfamd_fidelity_loss; no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * partial sums live in a file-scope device array of
 * SFAMD_NOALLOC_SCRATCH_DOUBLES doubles per GPU, held until the process ends
 * (n_elems whose sfamd_overlap_partials exceed it return SFAMD_E_WORKSPACE).
 * The device whose array is used is the current device, which must be the
 * stream's.  Calls that overlap on one GPU share that array and race. */
#define SFAMD_NOALLOC_SCRATCH_DOUBLES 2048
int sfamd_fidelity_loss_noalloc(const void* target, const void* psi, int64_t n_elems,
                                double* out_dev, void* g_psi, int dtype, void* stream);

/* Total mean photon number of a ket, L = sum_I |psi_I|^2 sum_m I_m (FP64,
 * out_dev[0]), and if g_psi != NULL its cotangent g = 2 (sum_m I_m) psi.
 * (BaseFockState.mean_photon / number_expectation summed over modes,
 * states.py:657-663,968-974.)  `partial` holds sfamd_reduce_partials(n, D, 0)
 * doubles. */
int sfamd_number_loss(const void* psi, void* g_psi, int n, int D, double* partial,
                      double* out_dev, int dtype, void* stream);

/* ---- measurements and multi-mode preparation ---------------------------- */
/* Marginal Fock distribution of the sorted `modes` (k of them): out_dev[D^k]
 * (double, first listed mode most significant) = sum over the other modes of
 * |psi|^2 or of the DM diagonal.  Replaces the partial trace + diagonal of
 * Circuit.measure_fock (fockbackend/circuit.py:671-677).  `partial` holds
 * sfamd_marginal_partials(n, D, k) doubles. */
int64_t sfamd_marginal_partials(int n, int D, int k);
int sfamd_marginal_probs(const void* state, int n, int D, int mixed,
                         const int* modes, int k, double* partial,
                         double* out_dev, int dtype, void* stream);

/* out = scale * (|0><x| on each of `modes`) in, outcome x[t] for modes[t]
 * (any order), both sides of a DM.  Replaces ops.project_reset
 * (fockbackend/ops.py:179-198) followed by the division by the norm
 * (circuit.py:689-699); the caller passes scale = 1/norm.  in != out. */
int sfamd_project_reset(const void* in, void* out, int n, int D, int mixed,
                        const int* modes, const int* outcome, int k,
                        double scale, int dtype, void* stream);

/* Homodyne density of a single-mode DM rho[D][D] (bra, ket) rotated by
 * R(-phi), on the grid q_b = -q_mag + b * 2 q_mag / (num_bins - 1):
 * out_dev[b] = sqrt(1/(pi hbar)) dq sum_{n,m} rho'_nm h_n(x) h_m(x),
 * x = q / sqrt(hbar), h_n the normalised Hermite functions (double).
 * Replaces the pdf of Circuit.measure_homodyne (circuit.py:729-765) and
 * ops.hermiteVals (ops.py:500-515).  D <= SFAMD_MAX_CUTOFF. */
int sfamd_homodyne_pdf(const void* rho, int D, double phi, double q_mag,
                       int num_bins, double hbar, double* out_dev, int dtype,
                       void* stream);

/* Product state of prepare_multimode (fockbackend/circuit.py:393-473): out =
 * rest (x) sigma with sigma's axis t on register mode modes[t] (any order) and
 * `rest` the (n-k)-mode DM of the other modes in increasing order.  DMs
 * (mixed = 1); for k == n, rest may be NULL and sigma may be a ket
 * (mixed = 0), which gives the permuted ket. */
int sfamd_embed_modes(const void* rest, const void* sigma, void* out, int n,
                      int D, const int* modes, int k, int mixed, int dtype,
                      void* stream);

/* ---- recorded circuits (forward, adjoint, graphs) ----------------------- */
typedef struct {
  int32_t kind;   /* SFAMD_OP_*                                              */
  int32_t mode0;  /* first mode (mode1 of a beamsplitter)                   */
  int32_t mode1;  /* second mode of a beamsplitter, else -1                 */
  int32_t pad;
} sfamd_op_t;

typedef struct sfamd_prog sfamd_prog_t;

typedef struct {
  int64_t table_elems;    /* complex elements of the gate table             */
  int64_t partial_doubles;/* doubles of the adjoint partial-sum buffer      */
  int64_t ws_elems;       /* complex elements of the scratch workspace      */
  int64_t final_elems;    /* elements of the output state                   */
  int32_t final_mixed;    /* 1 if the output state is a DM                  */
  int32_t nops;
  int32_t nsteps;         /* execution steps (fused stages or single ops)   */
  int32_t pad;
} sfamd_prog_info_t;

/* Plans a circuit of `nops` operations on an n-mode state at cutoff D,
 * starting pure (mixed_in = 0) or mixed.  A loss on a pure state inserts the
 * pure -> mixed conversion as Circuit._apply_channel does
 * (fockbackend/circuit.py:79-81).  With fuse_K >= 2, runs of gates on a pure
 * state are grouped into steps that each apply many gates to LDS tiles of
 * fuse_K local modes (gates on disjoint modes commute, so a gate may move to
 * an earlier step when no earlier gate shares a mode); fuse_K = 0 gives one
 * step per operation.  Host memory only. */
int sfamd_prog_create(const sfamd_op_t* ops, int nops, int n, int D,
                      int mixed_in, int dtype, int fuse_K, sfamd_prog_t** out);
int sfamd_prog_info(const sfamd_prog_t* prog, sfamd_prog_info_t* info);
/* state elements at the input of step i (i = nsteps: the output) */
int64_t sfamd_prog_state_elems(const sfamd_prog_t* prog, int i);
/* operation indices of step `step` in execution order; returns their count */
int sfamd_prog_step_ops(const sfamd_prog_t* prog, int step, int32_t* ops_out,
                        int cap);
int sfamd_prog_destroy(sfamd_prog_t* prog);

/* Gate table of all operations from the device parameter vector
 * params[2*i + slot] (slot 0 = first gate parameter, 1 = second). */
int sfamd_prog_build_table(const sfamd_prog_t* prog, const double* params,
                           void* table, void* stream);

/* Runs steps [begin, end) on `in`, writing the result to `out`.
 * saved[i - begin], if saved != NULL and the entry is non-NULL, receives the
 * input state of step i (steps write their output straight into the next
 * saved slot, so checkpoints cost no copies; saved[0] may equal `in`).
 * `ws` is scratch of sfamd_prog_info.ws_elems elements.  `in` is not
 * modified; out must not alias in. */
int sfamd_prog_forward(const sfamd_prog_t* prog, int begin, int end,
                       const void* table, const void* in, void* out,
                       void* const* saved, void* ws, void* stream);

/* Adjoint sweep over steps [begin, end) in reverse.  saved[i - begin] holds
 * the input state of step i; g_out is the cotangent of the state after step
 * end-1 (PyTorch convention dL/dRe + i dL/dIm); g_in receives the cotangent
 * of the state before step begin.  The gradients of the operations of these
 * steps are written to grads[2*op + slot] (double, device); other entries are
 * left untouched.  `partial` holds sfamd_prog_info.partial_doubles doubles. */
int sfamd_prog_adjoint(const sfamd_prog_t* prog, int begin, int end,
                       const void* table, void* const* saved, const void* g_out,
                       void* g_in, double* grads, double* partial, void* ws,
                       void* stream);

/* HIP graph capture of a sequence of calls issued on `stream`. */
int sfamd_graph_begin(void* stream);
int sfamd_graph_end(void* stream, void** graph_exec);
int sfamd_graph_launch(void* graph_exec, void* stream);
int sfamd_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif

#endif /* SFAMD_H */
