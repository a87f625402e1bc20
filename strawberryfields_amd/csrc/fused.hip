// Fused gate stages on a pure state (ket): many gates per pass, in LDS.
//
// A stage is a set of K "local" modes and an ordered list of gates that act
// only on those modes (R, K, D, S on one local mode; BS on two).  The ket is
// cut into tiles: a tile holds the D^K amplitudes of the local modes for one
// assignment of the other modes' digits.  One workgroup loads a tile into LDS,
// applies every gate of the stage in LDS (ping-pong buffers, one barrier per
// gate) and writes the tile back, so a stage costs one read and one write of
// the state however many gates it holds.  Gates on disjoint modes commute, so
// the planner (program.hip) may move a gate into an earlier stage as long as
// no earlier gate shares a mode with it.
//
// The adjoint of a stage loads the stage's input tile (its checkpoint) and the
// incoming cotangent tile, recomputes the stage forward keeping every
// intermediate tile in LDS, and runs the gates backward: lam <- G^H lam and,
// per gate parameter, Re sum conj(lam_out) (dG/dp psi_in).  Parameter partial
// sums are reduced per wavefront with shuffles and per workgroup in a fixed
// order (deterministic).
//
// D and K are compile-time, so every digit extraction is a multiply-shift and
// every contraction loop is unrolled.  Matrices are staged in LDS per gate.
#include "common.h"
#include "fused.h"

namespace {

template <typename T>
using C = cpx<T>;


template <int D, int E>
struct Pow {
  static constexpr int v = D * Pow<D, E - 1>::v;
};
template <int D>
struct Pow<D, 0> {
  static constexpr int v = 1;
};

template <int D, int K>
__device__ __forceinline__ int pw(int e) {
  // D^e for 0 <= e < K (e runtime, small)
  int r = 1;
#pragma unroll
  for (int t = 0; t < K; ++t)
    if (t < e) r *= D;
  return r;
}

// digit of local axis `a` (0 = slowest) of tile element e; the loop is fully
// unrolled over compile-time D and K, so no array (and no scratch) is needed
template <int D, int K>
__device__ __forceinline__ int digit(int e, int a) {
  int r = e, v = 0;
#pragma unroll
  for (int t = K - 1; t >= 0; --t) {
    const int q = r / D;
    const int dd = r - q * D;
    if (t == a) v = dd;
    r = q;
  }
  return v;
}

// digits of two local axes at once
template <int D, int K>
__device__ __forceinline__ void digit2(int e, int a, int b, int& va, int& vb) {
  int r = e;
  va = 0;
  vb = 0;
#pragma unroll
  for (int t = K - 1; t >= 0; --t) {
    const int q = r / D;
    const int dd = r - q * D;
    if (t == a) va = dd;
    if (t == b) vb = dd;
    r = q;
  }
}

// ---- staging of a stage's gate tables into LDS (once per workgroup) -------
template <int D>
__device__ __forceinline__ int gate_table_elems(int kind, bool derivs) {
  switch (kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      return D;
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE:
      return derivs ? 3 * D * D : D * D;
    case SFAMD_OP_BEAMSPLITTER:
      return derivs ? 2 * (int)bs_nb(D) : (int)bs_nb(D);
    default:
      return 0;
  }
}

// all loads are issued before the single barrier that follows, so the stage
// pays one global-memory latency for its matrices, not one per gate
template <typename T, int D, int NT>
__device__ __forceinline__ void stage_all_matrices(const C<T>* __restrict__ table, const sfamd::StageArgs& a,
                                                   bool derivs, C<T>* mats) {
  for (int gi = 0; gi < a.ngates; ++gi) {
    const sfamd::StageGate& g = a.g[gi];
    const int cnt = gate_table_elems<D>(g.kind, derivs);
    const C<T>* src = table + g.off;
    C<T>* dst = mats + (derivs ? g.amoff : g.moff);
    for (int i = threadIdx.x; i < cnt; i += NT) dst[i] = src[i];
  }
}

// ---- one gate on a tile: dst = G src --------------------------------------
// `put(e, v)` stores output element e: into the next LDS tile, or straight to
// global memory for the last gate of a forward stage.
template <typename T, int D, int K, int NT, typename Put>
__device__ __forceinline__ void apply_gate_to(const sfamd::StageGate& g, const C<T>* __restrict__ mat,
                                              const C<T>* __restrict__ src, Put put) {
  constexpr int DK = Pow<D, K>::v;
  const int a0 = g.a0, a1 = g.a1;
  const int s0 = pw<D, K>(K - 1 - a0);
  if (g.kind == SFAMD_OP_ROTATION || g.kind == SFAMD_OP_KERR) {
    for (int e = threadIdx.x; e < DK; e += NT) put(e, cmul(mat[digit<D, K>(e, a0)], src[e]));
  } else if (g.kind == SFAMD_OP_BEAMSPLITTER) {
    const int s1 = pw<D, K>(K - 1 - a1);
    for (int e = threadIdx.x; e < DK; e += NT) {
      int i, j;
      digit2<D, K>(e, a0, a1, i, j);
      const int s = i + j;
      const int lo = s - D + 1 > 0 ? s - D + 1 : 0;
      const int hi = s < D - 1 ? s : D - 1;
      const int c = hi - lo + 1;
      const int base = e - i * s0 - j * s1;
      const C<T>* row = mat + bs_block_off(s, D) + (i - lo) * c;
      C<T> acc = mk<T>(0, 0);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (k >= lo && k <= hi) cmac(acc, row[k - lo], src[base + k * s0 + (s - k) * s1]);
      }
      put(e, acc);
    }
  } else if (g.kind == SFAMD_OP_SQUEEZE) {  // zero unless i + k even: half the terms
    for (int e = threadIdx.x; e < DK; e += NT) {
      const int i = digit<D, K>(e, a0);
      const int base = e - i * s0;
      const C<T>* row = mat + i * D;
      C<T> acc = mk<T>(0, 0);
      const int p = i & 1;
#pragma unroll
      for (int h = 0; h < (D + 1) / 2; ++h) {
        const int k = 2 * h + p;
        if (k < D) cmac(acc, row[k], src[base + k * s0]);
      }
      put(e, acc);
    }
  } else {  // dense single-mode
    for (int e = threadIdx.x; e < DK; e += NT) {
      const int i = digit<D, K>(e, a0);
      const int base = e - i * s0;
      const C<T>* row = mat + i * D;
      C<T> acc = mk<T>(0, 0);
#pragma unroll
      for (int k = 0; k < D; ++k) cmac(acc, row[k], src[base + k * s0]);
      put(e, acc);
    }
  }
}

template <typename T, int D, int K, int NT>
__device__ __forceinline__ void apply_gate(const sfamd::StageGate& g, const C<T>* __restrict__ mat,
                                           const C<T>* __restrict__ src, C<T>* __restrict__ dst) {
  apply_gate_to<T, D, K, NT>(g, mat, src, [dst](int e, C<T> v) { dst[e] = v; });
}

// ---- adjoint of one gate on a tile -----------------------------------------
// Each thread takes its element e as an INPUT index k of the gate, reads the
// cotangent fiber once and forms
//   lam_in[k] = sum_i conj(G[i][k]) lam_out[i]
//   g_p      += Re( (sum_i conj(lam_out[i]) dG/dp[i][k]) psi_in[k] ),
// which summed over k is Re sum conj(lam_out) (dG/dp psi_in).  The phase
// derivatives come from G itself: d/dphi of D, S, BS is i (m - n) f G with
// f = 1, 1/2, 1 (row m = out index, column n = in index).
// `put(e, v)` stores lam_in element e: into the other LDS cotangent tile, or
// straight to global memory for the first gate of the stage (the last one the
// backward sweep visits).
template <typename T, int D, int K, int NT, typename Put>
__device__ __forceinline__ void adjoint_gate(const sfamd::StageGate& g, const C<T>* __restrict__ mat,
                                             const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_,
                                             Put put, double& g0, double& g1) {
  constexpr int DK = Pow<D, K>::v;
  const int a0 = g.a0, a1 = g.a1;
  const int s0 = pw<D, K>(K - 1 - a0);
  if (g.kind == SFAMD_OP_ROTATION || g.kind == SFAMD_OP_KERR) {
    const bool kerr = g.kind == SFAMD_OP_KERR;
    for (int e = threadIdx.x; e < DK; e += NT) {
      const int i = digit<D, K>(e, a0);
      const C<T> f = mat[i];
      const C<T> l = lo_[e];
      put(e, cmul(cconj(f), l));
      const T w = (T)(kerr ? i * i : i);
      const C<T> fx = cmul(f, psi[e]);
      g0 += (double)re_conj_mul(l, mk<T>(-w * fx.im, w * fx.re));
    }
  } else if (g.kind == SFAMD_OP_BEAMSPLITTER) {
    const int s1 = pw<D, K>(K - 1 - a1);
    const int nb = (int)bs_nb(D);
    for (int e = threadIdx.x; e < DK; e += NT) {
      int n0, m0;
      digit2<D, K>(e, a0, a1, n0, m0);
      const int s = n0 + m0;
      const int lo = s - D + 1 > 0 ? s - D + 1 : 0;
      const int hi = s < D - 1 ? s : D - 1;
      const int c = hi - lo + 1;
      const int base = e - n0 * s0 - m0 * s1;
      const C<T>* col = mat + bs_block_off(s, D) + (n0 - lo);
      C<T> acc = mk<T>(0, 0), w0 = mk<T>(0, 0), w1 = mk<T>(0, 0);
#pragma unroll
      for (int N = 0; N < D; ++N) {
        if (N >= lo && N <= hi) {
          const C<T> l = lo_[base + N * s0 + (s - N) * s1];
          const C<T> b = col[(N - lo) * c];
          cmac_conj(acc, b, l);                    // conj(B[N][n]) lam[N]
          cmac_conj(w0, l, col[nb + (N - lo) * c]);  // conj(lam[N]) dB[N][n]
          const C<T> lb = mk<T>(l.re * b.re + l.im * b.im, l.re * b.im - l.im * b.re);  // conj(lam) B
          const T w = (T)(n0 - N);                 // i (n - N) conj(lam) B
          w1.re += -w * lb.im;
          w1.im += w * lb.re;
        }
      }
      put(e, acc);
      const C<T> x = psi[e];
      g0 += (double)(w0.re * x.re - w0.im * x.im);
      g1 += (double)(w1.re * x.re - w1.im * x.im);
    }
  } else {
    constexpr int DD = D * D;
    const bool sq = g.kind == SFAMD_OP_SQUEEZE;
    const T f = sq ? (T)0.5 : (T)1;
    for (int e = threadIdx.x; e < DK; e += NT) {
      const int k = digit<D, K>(e, a0);
      const int base = e - k * s0;
      C<T> acc = mk<T>(0, 0), w0 = mk<T>(0, 0), w1 = mk<T>(0, 0);
#pragma unroll
      for (int i = 0; i < D; ++i) {
        if (!sq || ((i + k) & 1) == 0) {
          const C<T> l = lo_[base + i * s0];
          const C<T> u = mat[i * D + k];
          cmac_conj(acc, u, l);                  // conj(U[i][k]) lam[i]
          cmac_conj(w0, l, mat[DD + i * D + k]);  // conj(lam[i]) dU/dr[i][k]
          const C<T> lu = mk<T>(l.re * u.re + l.im * u.im, l.re * u.im - l.im * u.re);  // conj(lam) U
          const T w = f * (T)(i - k);            // i (i - k) f conj(lam) U
          w1.re += -w * lu.im;
          w1.im += w * lu.re;
        }
      }
      put(e, acc);
      const C<T> x = psi[e];
      g0 += (double)(w0.re * x.re - w0.im * x.im);
      g1 += (double)(w1.re * x.re - w1.im * x.im);
    }
  }
}

// One DPP lane move of a double (two 32-bit moves); lanes of rows outside
// RowMask receive 0.
template <int Ctrl, int RowMask>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, Ctrl, RowMask, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), Ctrl, RowMask, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes of a wave by DPP moves, which stay in the VALU instead
// of a round trip through the LDS crossbar per step as shuffles do: swaps
// within pairs and quads, then the half-row and row mirrors leave every lane of
// a 16-lane row with the row's sum; row broadcast 15 adds row 0 into row 1 and
// row 2 into row 3, row broadcast 31 adds rows 0-1 into row 3.  The total is
// valid in lane 63 (the summation order is fixed, so results are
// deterministic).
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xF>(v);  // row_mirror
  v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
  return v;
}
constexpr int kWaveSumLane = 63;

// global offset of tile element e (relative to the tile base)
template <int D, int K>
__device__ __forceinline__ int64_t local_offset(int e, const sfamd::StageArgs& a) {
  int r = e;
  int64_t off = 0;
#pragma unroll
  for (int t = K - 1; t >= 0; --t) {
    const int q = r / D;
    off += (int64_t)(r - q * D) * a.loc_stride[t];
    r = q;
  }
  return off;
}

__device__ __forceinline__ int64_t tile_base(int64_t t, const sfamd::StageArgs& a, int D) {
  int64_t base = 0;
  for (int u = a.nnl - 1; u >= 0; --u) {
    const int64_t dg = t % D;
    t /= D;
    base += dg * a.nl_stride[u];
  }
  return base;
}

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8),
// and each XCD has its own L2.  Neighbouring tiles share cache lines whenever
// the innermost non-local mode has a small stride, so the tiles are cut into 8
// contiguous ranges and XCD x's blocks (b = x + 8 k) walk range x; every line
// is then fetched into one L2 instead of eight.
struct TileRange {
  int64_t first, end, step;
};

__device__ __forceinline__ TileRange xcd_tiles(int64_t ntiles) {
  const int64_t G = gridDim.x, b = blockIdx.x;
  if (G % 8) return TileRange{b, ntiles, G};
  const int64_t per = (ntiles + 7) / 8;
  const int64_t x = b % 8;
  const int64_t end = (x + 1) * per < ntiles ? (x + 1) * per : ntiles;
  return TileRange{x * per + b / 8, end, G / 8};
}

// ---- kernels ------------------------------------------------------------
template <typename T, int D, int K, int NT = sfamd::stage_threads(Pow<D, K>::v)>
__global__ void __launch_bounds__(NT) k_stage_fwd(const C<T>* __restrict__ in, C<T>* __restrict__ out,
                                                        const C<T>* __restrict__ table, sfamd::StageArgs a) {
  constexpr int DK = Pow<D, K>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  C<T>* bufA = reinterpret_cast<C<T>*>(smem);
  C<T>* bufB = bufA + DK;
  C<T>* mats = bufB + DK;
  stage_all_matrices<T, D, NT>(table, a, false, mats);
  const TileRange tr = xcd_tiles(a.ntiles);
  for (int64_t t = tr.first; t < tr.end; t += tr.step) {
    const int64_t base = tile_base(t, a, D);
    for (int e = threadIdx.x; e < DK; e += NT) bufA[e] = in[base + local_offset<D, K>(e, a)];
    __syncthreads();
    C<T>* src = bufA;
    C<T>* dst = bufB;
    for (int gi = 0; gi < a.ngates - 1; ++gi) {
      apply_gate<T, D, K, NT>(a.g[gi], mats + a.g[gi].moff, src, dst);
      __syncthreads();
      C<T>* tmp = src;
      src = dst;
      dst = tmp;
    }
    // the last gate writes its outputs straight to the state
    const int gl = a.ngates - 1;
    apply_gate_to<T, D, K, NT>(a.g[gl], mats + a.g[gl].moff, src,
                               [&](int e, C<T> v) { out[base + local_offset<D, K>(e, a)] = v; });
    __syncthreads();
  }
}

template <typename T, int D, int K, int NT = sfamd::stage_threads(Pow<D, K>::v)>
__global__ void __launch_bounds__(NT) k_stage_adj(const C<T>* __restrict__ psi_in, const C<T>* __restrict__ lam_out,
                                                        C<T>* __restrict__ lam_in, const C<T>* __restrict__ table,
                                                        sfamd::StageArgs a, double* __restrict__ partial, int nwg) {
  constexpr int DK = Pow<D, K>::v;
  extern __shared__ __align__(16) unsigned char smem[];
  // layout: [ngates psi tiles (each gate's input)][2 lam tiles][matrices + derivatives][grad slots [ngates][2][kWaves]]
  C<T>* tiles = reinterpret_cast<C<T>*>(smem);
  C<T>* lamA = tiles + a.ngates * DK;
  C<T>* lamB = lamA + DK;
  C<T>* mats = lamB + DK;
  double* gsl = reinterpret_cast<double*>(mats + a.mat_adj);
  constexpr int kWaves = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int idx = threadIdx.x; idx < a.ngates * 2 * kWaves; idx += NT) gsl[idx] = 0.0;
  stage_all_matrices<T, D, NT>(table, a, true, mats);
  const TileRange tr = xcd_tiles(a.ntiles);
  for (int64_t t = tr.first; t < tr.end; t += tr.step) {
    const int64_t base = tile_base(t, a, D);
    for (int e = threadIdx.x; e < DK; e += NT) {
      const int64_t off = base + local_offset<D, K>(e, a);
      tiles[e] = psi_in[off];
      lamA[e] = lam_out[off];
    }
    __syncthreads();
    // recompute the stage forward, keeping every intermediate tile; the adjoint
    // of gate gi reads only its input tile, so the last gate's output is not formed
    for (int gi = 0; gi < a.ngates - 1; ++gi) {
      apply_gate<T, D, K, NT>(a.g[gi], mats + a.g[gi].amoff, tiles + gi * DK, tiles + (gi + 1) * DK);
      __syncthreads();
    }
    C<T>* lo_ = lamA;
    C<T>* li = lamB;
    for (int gi = a.ngates - 1; gi >= 0; --gi) {
      double g0 = 0.0, g1 = 0.0;
      if (gi > 0)
        adjoint_gate<T, D, K, NT>(a.g[gi], mats + a.g[gi].amoff, tiles + gi * DK, lo_,
                                  [li](int e, C<T> v) { li[e] = v; }, g0, g1);
      else  // the stage's first gate: lam_in goes straight to the output state
        adjoint_gate<T, D, K, NT>(a.g[gi], mats + a.g[gi].amoff, tiles + gi * DK, lo_,
                                  [&](int e, C<T> v) { lam_in[base + local_offset<D, K>(e, a)] = v; }, g0, g1);
      // R and K have one parameter: their second slot stays at its cleared zero
      const bool two = a.g[gi].kind != SFAMD_OP_ROTATION && a.g[gi].kind != SFAMD_OP_KERR;
      g0 = wave_sum(g0);
      if (two) g1 = wave_sum(g1);
      if (lane == kWaveSumLane) {  // each wave owns its slot: deterministic, no atomics
        gsl[(gi * 2 + 0) * kWaves + wave] += g0;
        if (two) gsl[(gi * 2 + 1) * kWaves + wave] += g1;
      }
      __syncthreads();
      C<T>* tmp = lo_;
      lo_ = li;
      li = tmp;
    }  // the sweep's last barrier also frees the tiles for the next tile's loads
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 2 * a.ngates; idx += NT) {
    const int gi = idx >> 1, slot = idx & 1;
    double s = 0.0;
    for (int w = 0; w < kWaves; ++w) s += gsl[idx * kWaves + w];
    // program partial layout [((op * 2 + pass) * 2 + slot) * nwg + workgroup]:
    // pass 0 gets the sum, pass 1 is cleared.  The launch may have fewer groups
    // than nwg (one per tile at most), so group b also clears the columns
    // b + gridDim.x, b + 2 gridDim.x, ... below nwg.
    double* col = partial + ((int64_t)a.g[gi].op * 4 + slot) * nwg;
    col[blockIdx.x] = s;
    col[2 * (int64_t)nwg + blockIdx.x] = 0.0;
    for (int w = blockIdx.x + gridDim.x; w < nwg; w += gridDim.x) {
      col[w] = 0.0;
      col[2 * (int64_t)nwg + w] = 0.0;
    }
  }
}

// ---- dispatch ---------------------------------------------------------------
// dynamic LDS above the default per-launch limit needs an opt-in per kernel
template <typename F>
int allow_lds(F kernel, size_t lds) {
  if (lds <= 65536) return 0;
  return (int)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
}

template <typename T, int D, int K>
int launch_fwd(const void* in, void* out, const void* table, const sfamd::StageArgs& a, int grid, hipStream_t s) {
  const size_t lds = (size_t)sfamd::stage_lds_fwd(D, K, a.mat_fwd, sizeof(C<T>) == 8 ? SFAMD_C64 : SFAMD_C128);
  RET_IF(allow_lds(k_stage_fwd<T, D, K>, lds));
  hipLaunchKernelGGL((k_stage_fwd<T, D, K>), dim3(grid), dim3(sfamd::stage_threads(Pow<D, K>::v)), lds, s, (const C<T>*)in, (C<T>*)out,
                     (const C<T>*)table, a);
  return (int)hipGetLastError();
}

template <typename T, int D, int K>
int launch_adj(const void* psi, const void* lo, void* li, const void* table, const sfamd::StageArgs& a, int grid,
               double* partial, int nwg, hipStream_t s) {
  const size_t lds =
      (size_t)sfamd::stage_lds_adj(D, K, a.ngates, a.mat_adj, sizeof(C<T>) == 8 ? SFAMD_C64 : SFAMD_C128);
  RET_IF(allow_lds(k_stage_adj<T, D, K>, lds));
  hipLaunchKernelGGL((k_stage_adj<T, D, K>), dim3(grid), dim3(sfamd::stage_threads(Pow<D, K>::v)), lds, s, (const C<T>*)psi, (const C<T>*)lo,
                     (C<T>*)li, (const C<T>*)table, a, partial, nwg);
  return (int)hipGetLastError();
}

template <typename T, int D>
int fwd_k(int K, const void* in, void* out, const void* table, const sfamd::StageArgs& a, int grid, hipStream_t s) {
  switch (K) {
    case 2:
      return launch_fwd<T, D, 2>(in, out, table, a, grid, s);
    case 3:
      if constexpr (D * D * D <= sfamd::kStageMaxTile) return launch_fwd<T, D, 3>(in, out, table, a, grid, s);
      break;
    case 4:
      if constexpr (D * D * D * D <= sfamd::kStageMaxTile) return launch_fwd<T, D, 4>(in, out, table, a, grid, s);
      break;
    default:
      break;
  }
  return SFAMD_E_ARG;
}

template <typename T, int D>
int adj_k(int K, const void* psi, const void* lo, void* li, const void* table, const sfamd::StageArgs& a, int grid,
          double* partial, int nwg, hipStream_t s) {
  switch (K) {
    case 2:
      return launch_adj<T, D, 2>(psi, lo, li, table, a, grid, partial, nwg, s);
    case 3:
      if constexpr (D * D * D <= sfamd::kStageMaxTile)
        return launch_adj<T, D, 3>(psi, lo, li, table, a, grid, partial, nwg, s);
      break;
    case 4:
      if constexpr (D * D * D * D <= sfamd::kStageMaxTile)
        return launch_adj<T, D, 4>(psi, lo, li, table, a, grid, partial, nwg, s);
      break;
    default:
      break;
  }
  return SFAMD_E_ARG;
}

#define SFAMD_FOR_EACH_D(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)

template <typename T>
int fwd_dispatch(int D, int K, const void* in, void* out, const void* table, const sfamd::StageArgs& a, int grid,
                 hipStream_t s) {
  switch (D) {
#define CASE_D(d) \
  case d:         \
    return fwd_k<T, d>(K, in, out, table, a, grid, s);
    SFAMD_FOR_EACH_D(CASE_D)
#undef CASE_D
    default:
      return SFAMD_E_CUTOFF;
  }
}

template <typename T>
int adj_dispatch(int D, int K, const void* psi, const void* lo, void* li, const void* table, const sfamd::StageArgs& a,
                 int grid, double* partial, int nwg, hipStream_t s) {
  switch (D) {
#define CASE_D(d) \
  case d:         \
    return adj_k<T, d>(K, psi, lo, li, table, a, grid, partial, nwg, s);
    SFAMD_FOR_EACH_D(CASE_D)
#undef CASE_D
    default:
      return SFAMD_E_CUTOFF;
  }
}

}  // namespace

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
      return derivs ? 3 * (int64_t)D * D : (int64_t)D * D;
    case SFAMD_OP_BEAMSPLITTER:
      return derivs ? 2 * bs_nb(D) : bs_nb(D);
    default:
      return 0;
  }
}

int64_t stage_lds_fwd(int D, int K, int mat_fwd, int dtype) {
  const int64_t esz = dtype == SFAMD_C64 ? 8 : 16;
  return (2 * ipow(D, K) + mat_fwd) * esz;
}

int64_t stage_lds_adj(int D, int K, int ngates, int mat_adj, int dtype) {
  const int64_t esz = dtype == SFAMD_C64 ? 8 : 16;
  return ((ngates + 2) * ipow(D, K) + mat_adj) * esz + (int64_t)ngates * 2 * (sfamd::stage_threads((int)ipow(D, K)) / 64) * 8;
}

int stage_fwd(int D, int K, const void* in, void* out, const void* table, const StageArgs& a, int grid, int dtype,
              hipStream_t s) {
  if (dtype == SFAMD_C64) return fwd_dispatch<float>(D, K, in, out, table, a, grid, s);
  return fwd_dispatch<double>(D, K, in, out, table, a, grid, s);
}

int stage_adj(int D, int K, const void* psi, const void* lam_out, void* lam_in, const void* table,
              const StageArgs& a, int grid, double* partial, int nwg, int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64) return adj_dispatch<float>(D, K, psi, lam_out, lam_in, table, a, grid, partial, nwg, s);
  return adj_dispatch<double>(D, K, psi, lam_out, lam_in, table, a, grid, partial, nwg, s);
}

}  // namespace sfamd
