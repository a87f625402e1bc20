// Fused gate stages on a pure state (ket): many gates per pass, in LDS.
//
// A stage is a set of K "local" modes and an ordered list of gates that act
// only on those modes (R, K, D, S on one local mode; BS on two).  The ket is
// cut into tiles: a tile holds the D^K amplitudes of the local modes for one
// assignment of the other modes' digits.  One workgroup loads a tile into LDS,
// applies every gate of the stage there (ping-pong tiles, one barrier per
// gate) and writes the tile back, so a stage costs one read and one write of
// the state however many gates it holds.  Gates on disjoint modes commute, so
// the planner (program.hip) may move a gate into an earlier stage as long as
// no earlier gate shares a mode with it.
//
// Work decomposition of a gate: the tile's D^K amplitudes form F = D^(K-1)
// fibers of D amplitudes, and a work item computes R output rows of one fiber
// from the whole fiber, held in registers, so each LDS read of the fiber
// feeds R outputs.
//  * single-mode gate on local axis a: the fibers run along a; the items of
//    one wave share their rows, so the matrix reads are wave-wide broadcasts;
//  * beamsplitter on local axes (a0, a1): photon number s = n0 + n1 is
//    conserved, and fiber (z, p) collects the modular anti-diagonal
//    {(n0 = k, n1 = (p - k) mod D)}: the full block s = p (k <= p) and the
//    partial block s = p + D (k > p).  The gate acts on each fiber as
//    blockdiag(B_p, B_(p+D)); a row of B_p reads the fiber with its k > p
//    entries zeroed, a row of B_(p+D) the opposite, so every item runs the same
//    unrolled D-term loop over the compact block table (out-of-block table
//    reads are finite and multiply zeros).
// The LDS tile has one padding slot after every D amplitudes, which spreads
// the fibers of the innermost axis (D amplitudes apart) over the banks.
//
// The adjoint of a stage loads the stage's input tile (its checkpoint) and the
// incoming cotangent tile, recomputes the stage forward keeping every gate's
// input and output tile in LDS, and runs the gates backward:
//   lam_in = G^H lam_out,  mu = (dG/dp0)^H lam_out,  g0 = Re sum conj(mu) psi_in,
// and the phase parameter from the identity dG/dphi = i f (out - in) G
// (D: f = 1 on the mode's digit, S: f = 1/2; BS: i (n - N) on the first mode),
//   g1 = f Re i [sum_m m conj(lam_out,m) psi_out,m - sum_n n conj(lam_in,n) psi_in,n],
// which needs no second matrix product.  R and K: g = Re sum conj(lam) i w f psi.
// Parameter sums are reduced per wavefront by DPP lane moves and per
// workgroup in a fixed order (deterministic), one partial per workgroup.
//
// D and K are compile-time, so every digit extraction is a multiply-shift and
// every contraction loop is unrolled.
#pragma once

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
struct Geo {
  static constexpr int DK = Pow<D, K>::v;
  static constexpr int F = DK / D;  // fibers of one axis
  static constexpr int PDK = DK + F;  // padded LDS tile
  static constexpr int R = sfamd::stage_rows(D);
  static constexpr int G = (D + R - 1) / R;  // row groups
  static constexpr int ITEMS = F * G;
  static constexpr int NT = sfamd::stage_threads(D, K);
};

// padded LDS position of tile element e
template <int D>
__device__ __forceinline__ int pad(int e) {
  return e + e / D;
}

// global offset of tile element e relative to the tile base
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

// ---- staging of a stage's gate tables into LDS (once per workgroup) -------
// Each gate region holds the matrix (forward) or the matrix and its first
// parameter's derivative (adjoint).  Beamsplitters at D <= kFullBsMaxD are
// expanded to the zero-padded per-fiber layout M[r][k][p] (row r, column k of
// the fiber matrix blockdiag(B_p, B_(p+D)) of anti-diagonal p), so the lanes of
// a wave (consecutive p) read consecutive entries and no masking is needed;
// at larger D the compact blocks are kept and followed by D zeros (their rows
// read up to D - 1 entries past a block).
constexpr int kFullBsMaxD = 10;

template <int D>
__device__ __forceinline__ int gate_table_elems(int kind, bool derivs) {
  switch (kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      return D;
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE:
      return derivs ? 2 * D * D : D * D;
    case SFAMD_OP_BEAMSPLITTER:
      return derivs ? 2 * (int)bs_nb(D) : (int)bs_nb(D);
    default:
      return 0;
  }
}

// compact-table row bases of fiber p: row r of the fiber reads
// table[row_base(r) + k] for its block's columns k
struct BsRows {
  int o1, c1, o2, c2;
};
template <int D>
__device__ __forceinline__ BsRows bs_rows(int p) {
  BsRows b;
  b.c1 = p + 1;
  b.c2 = D - 1 - p;
  b.o1 = (int)bs_block_off(p, D);
  b.o2 = (int)bs_block_off(p + D, D) - (p + 1) * (b.c2 + 1);
  return b;
}
__device__ __forceinline__ int bs_row_base(const BsRows& b, int p, int r) {
  return r <= p ? b.o1 + r * b.c1 : b.o2 + r * b.c2;
}

// all loads are issued before the barrier that follows, so the stage pays one
// global-memory latency for its matrices, not one per gate
template <typename T, int D, int NT>
__device__ __forceinline__ void stage_all_matrices(const C<T>* __restrict__ table, const sfamd::StageArgs& a,
                                                   bool derivs, C<T>* mats) {
  for (int gi = 0; gi < a.ngates; ++gi) {
    const sfamd::StageGate& g = a.g[gi];
    const C<T>* src = table + g.off;
    C<T>* dst = mats + g.moff;
    if (D <= kFullBsMaxD && g.kind == SFAMD_OP_BEAMSPLITTER) {
      constexpr int D3 = D * D * D;
      const int nb = (int)bs_nb(D);
      for (int i = threadIdx.x; i < (derivs ? 2 : 1) * D3; i += NT) {
        const int part = i / D3, e = i - part * D3;
        const int r = e / (D * D), k = (e / D) % D, p = e % D;
        C<T> v = mk<T>(0, 0);
        if ((r <= p) == (k <= p)) v = src[part * nb + bs_row_base(bs_rows<D>(p), p, r) + k];
        dst[i] = v;
      }
    } else {
      const int cnt = gate_table_elems<D>(g.kind, derivs);
      for (int i = threadIdx.x; i < cnt + D; i += NT) dst[i] = i < cnt ? src[i] : mk<T>(0, 0);
    }
  }
}

// ---- complex multiply-adds as explicit FMAs ---------------------------------
__device__ __forceinline__ float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }
// acc += a b
template <typename T>
__device__ __forceinline__ void cfma(C<T>& acc, C<T> a, C<T> b) {
  acc.re = fmat(a.re, b.re, acc.re);
  acc.re = fmat(-a.im, b.im, acc.re);
  acc.im = fmat(a.re, b.im, acc.im);
  acc.im = fmat(a.im, b.re, acc.im);
}
// acc += conj(a) b
template <typename T>
__device__ __forceinline__ void cfma_conj(C<T>& acc, C<T> a, C<T> b) {
  acc.re = fmat(a.re, b.re, acc.re);
  acc.re = fmat(a.im, b.im, acc.re);
  acc.im = fmat(a.re, b.im, acc.im);
  acc.im = fmat(-a.im, b.re, acc.im);
}
// Im(conj(a) b)
template <typename T>
__device__ __forceinline__ T im_conj_mul(C<T> a, C<T> b) {
  return a.re * b.im - a.im * b.re;
}

// ---- fiber geometry with compile-time axes ----------------------------------
// axis A of a tile of K local modes: tile stride D^(K-1-A) and the stride in
// the padded LDS tile
template <int D, int K, int A>
struct Ax {
  static constexpr int S = Pow<D, K - 1 - A>::v;
  static constexpr int PS = S == 1 ? 1 : S + S / D;
};

// tile element with digit 0 on axes A0 and A1 (A1 = A0 for one axis), the
// other digits taken from f (slowest first)
template <int D, int K, int A0, int A1>
__device__ __forceinline__ int fiber_base(int f) {
  int e = 0, rem = f, mul = 1;
#pragma unroll
  for (int t = K - 1; t >= 0; --t) {
    if (t != A0 && t != A1) {
      const int q = rem / D;
      e += (rem - q * D) * mul;
      rem = q;
    }
    mul *= D;
  }
  return e;
}

// beamsplitter fiber (z, p) on axes (A0, A1): element r has first-axis digit r
// and second-axis digit (p - r) mod D
template <int D>
__device__ __forceinline__ int bs_q(int p, int r) {
  return r <= p ? p - r : p + D - r;
}

// ---- forward of one gate on a tile ------------------------------------------
// `put(P, goff, v)` stores one output element: into the next LDS tile at
// padded position P, or straight to global memory at tile offset goff for the
// last gate of a forward stage (goff is computed only when the store uses it).
template <typename T, int D, int K, int A, typename Put>
__device__ __forceinline__ void fwd_single(int kind, const C<T>* __restrict__ mat, const C<T>* __restrict__ src,
                                           const sfamd::StageArgs& a, Put put) {
  using Gm = Geo<D, K>;
  using X = Ax<D, K, A>;
  constexpr int R = Gm::R;
  constexpr bool full_rows = Gm::G * R == D;
  for (int it = threadIdx.x; it < Gm::ITEMS; it += Gm::NT) {
    const int f = it % Gm::F;
    const int r0 = (it / Gm::F) * R;
    const int e0 = fiber_base<D, K, A, A>(f);
    const int P0 = pad<D>(e0);
    auto go = [&](int r) { return local_offset<D, K>(e0, a) + (int64_t)r * a.loc_stride[A]; };
    if (kind == SFAMD_OP_ROTATION || kind == SFAMD_OP_KERR) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = r0 + j;
        if (full_rows || r < D) put(P0 + r * X::PS, go(r), cmul(mat[r], src[P0 + r * X::PS]));
      }
      continue;
    }
    C<T> x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = src[P0 + k * X::PS];
    const bool sq = kind == SFAMD_OP_SQUEEZE;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = r0 + j;
      if (!full_rows && r >= D) continue;
      const C<T>* row = mat + r * D;
      C<T> acc = mk<T>(0, 0);
      if (sq) {  // zero unless r + k even; r0 is even
#pragma unroll
        for (int k = j & 1; k < D; k += 2) cfma(acc, row[k], x[k]);
      } else {
#pragma unroll
        for (int k = 0; k < D; ++k) cfma(acc, row[k], x[k]);
      }
      put(P0 + r * X::PS, go(r), acc);
    }
  }
}

template <typename T, int D, int K, int A0, int A1, typename Put>
__device__ __forceinline__ void fwd_bs(const C<T>* __restrict__ mat, const C<T>* __restrict__ src,
                                       const sfamd::StageArgs& a, Put put) {
  using Gm = Geo<D, K>;
  constexpr int R = Gm::R;
  constexpr bool full_rows = Gm::G * R == D;
  constexpr int PS0 = Ax<D, K, A0>::PS, PS1 = Ax<D, K, A1>::PS;
  for (int it = threadIdx.x; it < Gm::ITEMS; it += Gm::NT) {
    const int f = it % Gm::F;
    const int r0 = (it / Gm::F) * R;
    const int p = f % D;
    const int e0 = fiber_base<D, K, A0, A1>(f / D);
    const int Pp = pad<D>(e0) + p * PS1;  // element r at Pp + r (PS0 - PS1) (+ D PS1 when r > p)
    auto pos = [&](int r) { return Pp + r * (PS0 - PS1) + (r > p ? D * PS1 : 0); };
    auto go = [&](int r) {
      return local_offset<D, K>(e0, a) + (int64_t)r * a.loc_stride[A0] + (int64_t)bs_q<D>(p, r) * a.loc_stride[A1];
    };
    C<T> x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = src[pos(k)];
    if constexpr (D <= kFullBsMaxD) {
      const C<T>* mp = mat + p;  // M[r][k][p] at (r D + k) D + p
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = r0 + j;
        if (!full_rows && r >= D) continue;
        const C<T>* row = mp + r * D * D;
        C<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int k = 0; k < D; ++k) cfma(acc, row[k * D], x[k]);
        put(pos(r), go(r), acc);
      }
    } else {
      const BsRows br = bs_rows<D>(p);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = r0 + j;
        if (!full_rows && r >= D) continue;
        const bool up = r <= p;
        const C<T>* row = mat + bs_row_base(br, p, r);
        C<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int k = 0; k < D; ++k) cfma(acc, row[k], ((k <= p) == up) ? x[k] : mk<T>(0, 0));
        put(pos(r), go(r), acc);
      }
    }
  }
}

// ---- adjoint of one gate on a tile -------------------------------------------
// Item rows are INPUT indices k of the gate: lam_in[k] = sum_i conj(G[i][k])
// lam_out[i] over the fiber's lam_out (in registers), and the parameter sums
// of the header over the item's elements (the items' rows partition the tile).
// `put(P, goff, v)` stores lam_in: into the other LDS cotangent tile, or
// straight to global memory for the stage's first gate.
template <typename T, int D, int K, int A, typename Put>
__device__ __forceinline__ void adj_single(int kind, const C<T>* __restrict__ mat, const C<T>* __restrict__ pin,
                                           const C<T>* __restrict__ pout, const C<T>* __restrict__ lo,
                                           const sfamd::StageArgs& a, Put put, double& g0, double& g1) {
  using Gm = Geo<D, K>;
  using X = Ax<D, K, A>;
  constexpr int R = Gm::R;
  constexpr bool full_rows = Gm::G * R == D;
  constexpr int DD = D * D;
  for (int it = threadIdx.x; it < Gm::ITEMS; it += Gm::NT) {
    const int f = it % Gm::F;
    const int r0 = (it / Gm::F) * R;
    const int e0 = fiber_base<D, K, A, A>(f);
    const int P0 = pad<D>(e0);
    auto go = [&](int r) { return local_offset<D, K>(e0, a) + (int64_t)r * a.loc_stride[A]; };
    if (kind == SFAMD_OP_ROTATION || kind == SFAMD_OP_KERR) {
      const bool kerr = kind == SFAMD_OP_KERR;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int k = r0 + j;
        if (!full_rows && k >= D) continue;
        const int P = P0 + k * X::PS;
        const C<T> l = lo[P], fo = pout[P];
        put(P, go(k), cmul(cconj(mat[k]), l));
        const T w = (T)(kerr ? k * k : k);
        g0 += (double)(-w * im_conj_mul(l, fo));  // Re conj(l) i w psi_out
      }
      continue;
    }
    C<T> l[D];
#pragma unroll
    for (int i = 0; i < D; ++i) l[i] = lo[P0 + i * X::PS];
    const bool sq = kind == SFAMD_OP_SQUEEZE;
    const T fph = sq ? (T)0.5 : (T)1;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int k = r0 + j;
      if (!full_rows && k >= D) continue;
      C<T> acc = mk<T>(0, 0), mu = mk<T>(0, 0);
      const C<T>* col = mat + k;
      if (sq) {  // i + k even; r0 is even
#pragma unroll
        for (int i = j & 1; i < D; i += 2) {
          cfma_conj(acc, col[i * D], l[i]);
          cfma_conj(mu, col[DD + i * D], l[i]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < D; ++i) {
          cfma_conj(acc, col[i * D], l[i]);
          cfma_conj(mu, col[DD + i * D], l[i]);
        }
      }
      const int P = P0 + k * X::PS;
      put(P, go(k), acc);
      const C<T> xi = pin[P], xo = pout[P], lk = lo[P];
      g0 += (double)re_conj_mul(mu, xi);
      // dG/dphi = i f (m - n) G with m, n the mode's out / in digits
      g1 += (double)(fph * (T)k * (im_conj_mul(acc, xi) - im_conj_mul(lk, xo)));
    }
  }
}

template <typename T, int D, int K, int A0, int A1, typename Put>
__device__ __forceinline__ void adj_bs(const C<T>* __restrict__ mat, const C<T>* __restrict__ pin,
                                       const C<T>* __restrict__ pout, const C<T>* __restrict__ lo,
                                       const sfamd::StageArgs& a, Put put, double& g0, double& g1) {
  using Gm = Geo<D, K>;
  constexpr int R = Gm::R;
  constexpr bool full_rows = Gm::G * R == D;
  constexpr int PS0 = Ax<D, K, A0>::PS, PS1 = Ax<D, K, A1>::PS;
  for (int it = threadIdx.x; it < Gm::ITEMS; it += Gm::NT) {
    const int f = it % Gm::F;
    const int r0 = (it / Gm::F) * R;
    const int p = f % D;
    const int e0 = fiber_base<D, K, A0, A1>(f / D);
    const int Pp = pad<D>(e0) + p * PS1;
    auto pos = [&](int r) { return Pp + r * (PS0 - PS1) + (r > p ? D * PS1 : 0); };
    auto go = [&](int r) {
      return local_offset<D, K>(e0, a) + (int64_t)r * a.loc_stride[A0] + (int64_t)bs_q<D>(p, r) * a.loc_stride[A1];
    };
    C<T> l[D];
#pragma unroll
    for (int i = 0; i < D; ++i) l[i] = lo[pos(i)];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int k = r0 + j;
      if (!full_rows && k >= D) continue;
      C<T> acc = mk<T>(0, 0), mu = mk<T>(0, 0);
      if constexpr (D <= kFullBsMaxD) {
        constexpr int D3 = D * D * D;
        const C<T>* col = mat + p + k * D;  // M[i][k][p] at (i D + k) D + p
#pragma unroll
        for (int i = 0; i < D; ++i) {
          cfma_conj(acc, col[i * D * D], l[i]);
          cfma_conj(mu, col[D3 + i * D * D], l[i]);
        }
      } else {
        const BsRows br = bs_rows<D>(p);
        const int nb = (int)bs_nb(D);
        const bool up = k <= p;
#pragma unroll
        for (int i = 0; i < D; ++i) {
          const C<T> li = ((i <= p) == up) ? l[i] : mk<T>(0, 0);
          const int idx = bs_row_base(br, p, i) + k;
          cfma_conj(acc, mat[idx], li);
          cfma_conj(mu, mat[nb + idx], li);
        }
      }
      const int P = pos(k);
      put(P, go(k), acc);
      const C<T> xi = pin[P], xo = pout[P], lk = lo[P];
      g0 += (double)re_conj_mul(mu, xi);
      // dB/dphi = i (n - N) B with n, N the first-mode digits (k here)
      g1 += (double)((T)k * (im_conj_mul(lk, xo) - im_conj_mul(acc, xi)));
    }
  }
}

// ---- dispatch on the gate's local axes (uniform per workgroup) ----------------
template <typename T, int D, int K, int A0, typename Put>
__device__ __forceinline__ void fwd_bs_a1(int a1, const C<T>* mat, const C<T>* src, const sfamd::StageArgs& a,
                                          Put put) {
#define SFAMD_BS_A1(X)                                                         \
  if constexpr (X < K && X != A0) {                                            \
    if (a1 == X) return fwd_bs<T, D, K, A0, X>(mat, src, a, put);              \
  }
  SFAMD_BS_A1(0) SFAMD_BS_A1(1) SFAMD_BS_A1(2) SFAMD_BS_A1(3)
#undef SFAMD_BS_A1
}

template <typename T, int D, int K, typename Put>
__device__ __forceinline__ void gate_fwd(const sfamd::StageGate& g, const C<T>* __restrict__ mat,
                                         const C<T>* __restrict__ src, const sfamd::StageArgs& a, Put put) {
  const int kind = g.kind, a0 = g.a0;
  if (kind == SFAMD_OP_BEAMSPLITTER) {
#define SFAMD_BS_A0(X) \
  if constexpr (X < K) \
    if (a0 == X) return fwd_bs_a1<T, D, K, X>(g.a1, mat, src, a, put);
    SFAMD_BS_A0(0) SFAMD_BS_A0(1) SFAMD_BS_A0(2) SFAMD_BS_A0(3)
#undef SFAMD_BS_A0
    return;
  }
#define SFAMD_1_A(X)   \
  if constexpr (X < K) \
    if (a0 == X) return fwd_single<T, D, K, X>(kind, mat, src, a, put);
  SFAMD_1_A(0) SFAMD_1_A(1) SFAMD_1_A(2) SFAMD_1_A(3)
#undef SFAMD_1_A
}

template <typename T, int D, int K, int A0, typename Put>
__device__ __forceinline__ void adj_bs_a1(int a1, const C<T>* mat, const C<T>* pin, const C<T>* pout,
                                          const C<T>* lo, const sfamd::StageArgs& a, Put put, double& g0,
                                          double& g1) {
#define SFAMD_BS_A1(X)                                                              \
  if constexpr (X < K && X != A0) {                                                 \
    if (a1 == X) return adj_bs<T, D, K, A0, X>(mat, pin, pout, lo, a, put, g0, g1); \
  }
  SFAMD_BS_A1(0) SFAMD_BS_A1(1) SFAMD_BS_A1(2) SFAMD_BS_A1(3)
#undef SFAMD_BS_A1
}

template <typename T, int D, int K, typename Put>
__device__ __forceinline__ void gate_adj(const sfamd::StageGate& g, const C<T>* __restrict__ mat,
                                         const C<T>* __restrict__ pin, const C<T>* __restrict__ pout,
                                         const C<T>* __restrict__ lo, const sfamd::StageArgs& a, Put put, double& g0,
                                         double& g1) {
  const int kind = g.kind, a0 = g.a0;
  if (kind == SFAMD_OP_BEAMSPLITTER) {
#define SFAMD_BS_A0(X) \
  if constexpr (X < K) \
    if (a0 == X) return adj_bs_a1<T, D, K, X>(g.a1, mat, pin, pout, lo, a, put, g0, g1);
    SFAMD_BS_A0(0) SFAMD_BS_A0(1) SFAMD_BS_A0(2) SFAMD_BS_A0(3)
#undef SFAMD_BS_A0
    return;
  }
#define SFAMD_1_A(X)   \
  if constexpr (X < K) \
    if (a0 == X) return adj_single<T, D, K, X>(kind, mat, pin, pout, lo, a, put, g0, g1);
  SFAMD_1_A(0) SFAMD_1_A(1) SFAMD_1_A(2) SFAMD_1_A(3)
#undef SFAMD_1_A
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
template <typename T, int D, int K>
__global__ void __launch_bounds__(sfamd::stage_threads(D, K), 2) k_stage_fwd(const C<T>* __restrict__ in, C<T>* __restrict__ out,
                                                             const C<T>* __restrict__ table, sfamd::StageArgs a) {
  using Gm = Geo<D, K>;
  extern __shared__ __align__(16) unsigned char smem[];
  C<T>* bufA = reinterpret_cast<C<T>*>(smem);
  C<T>* bufB = bufA + Gm::PDK;
  C<T>* mats = bufB + Gm::PDK;
  constexpr int PER = (Gm::DK + Gm::NT - 1) / Gm::NT;  // tile elements per thread
  const TileRange tr = xcd_tiles(a.ntiles);
  // the first tile's loads are in flight while the matrices are staged, and
  // each next tile's while the gates of the current one run
  C<T> pre[PER];
  auto fetch = [&](int64_t t) {
    const int64_t b = tile_base(t, a, D);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * Gm::NT;
      if (e < Gm::DK) pre[i] = in[b + local_offset<D, K>(e, a)];
    }
  };
  if (tr.first < tr.end) fetch(tr.first);
  stage_all_matrices<T, D, Gm::NT>(table, a, false, mats);
  for (int64_t t = tr.first; t < tr.end; t += tr.step) {
    const int64_t base = tile_base(t, a, D);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * Gm::NT;
      if (e < Gm::DK) bufA[pad<D>(e)] = pre[i];
    }
    __syncthreads();
    if (t + tr.step < tr.end) fetch(t + tr.step);
    C<T>* src = bufA;
    C<T>* dst = bufB;
    for (int gi = 0; gi < a.ngates - 1; ++gi) {
      gate_fwd<T, D, K>(a.g[gi], mats + a.g[gi].moff, src, a, [dst](int P, int64_t, C<T> v) { dst[P] = v; });
      __syncthreads();
      C<T>* tmp = src;
      src = dst;
      dst = tmp;
    }
    // the last gate writes its outputs straight to the state
    const int gl = a.ngates - 1;
    gate_fwd<T, D, K>(a.g[gl], mats + a.g[gl].moff, src, a,
                      [&](int, int64_t go, C<T> v) { out[base + go] = v; });
    __syncthreads();
  }
}

template <typename T, int D, int K>
__global__ void __launch_bounds__(sfamd::stage_threads(D, K), 2) k_stage_adj(const C<T>* __restrict__ psi_in,
                                                             const C<T>* __restrict__ lam_out,
                                                             C<T>* __restrict__ lam_in,
                                                             const C<T>* __restrict__ table, sfamd::StageArgs a,
                                                             double* __restrict__ partial, int nwg) {
  using Gm = Geo<D, K>;
  extern __shared__ __align__(16) unsigned char smem[];
  // layout: [ngates + 1 psi tiles (every gate's input, then the last output)]
  //         [2 lam tiles][matrices + derivatives][grad slots [ngates][2][waves]]
  C<T>* tiles = reinterpret_cast<C<T>*>(smem);
  C<T>* lamA = tiles + (a.ngates + 1) * Gm::PDK;
  C<T>* lamB = lamA + Gm::PDK;
  C<T>* mats = lamB + Gm::PDK;
  double* gsl = reinterpret_cast<double*>(mats + a.mat_adj);
  constexpr int kWaves = Gm::NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int PER = (Gm::DK + Gm::NT - 1) / Gm::NT;  // tile elements per thread
  const TileRange tr = xcd_tiles(a.ntiles);
  // tile loads in flight while the matrices are staged (first tile) and while
  // the current tile's gates run (the next one)
  C<T> pre_psi[PER], pre_lam[PER];
  auto fetch = [&](int64_t t) {
    const int64_t b = tile_base(t, a, D);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * Gm::NT;
      if (e < Gm::DK) {
        const int64_t off = b + local_offset<D, K>(e, a);
        pre_psi[i] = psi_in[off];
        pre_lam[i] = lam_out[off];
      }
    }
  };
  if (tr.first < tr.end) fetch(tr.first);
  for (int idx = threadIdx.x; idx < a.ngates * 2 * kWaves; idx += Gm::NT) gsl[idx] = 0.0;
  stage_all_matrices<T, D, Gm::NT>(table, a, true, mats);
  for (int64_t t = tr.first; t < tr.end; t += tr.step) {
    const int64_t base = tile_base(t, a, D);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = threadIdx.x + i * Gm::NT;
      if (e < Gm::DK) {
        tiles[pad<D>(e)] = pre_psi[i];
        lamA[pad<D>(e)] = pre_lam[i];
      }
    }
    __syncthreads();
    if (t + tr.step < tr.end) fetch(t + tr.step);
    // recompute the stage forward, keeping every gate's input and output
    for (int gi = 0; gi < a.ngates; ++gi) {
      C<T>* dst = tiles + (gi + 1) * Gm::PDK;
      gate_fwd<T, D, K>(a.g[gi], mats + a.g[gi].moff, tiles + gi * Gm::PDK, a,
                        [dst](int P, int64_t, C<T> v) { dst[P] = v; });
      __syncthreads();
    }
    C<T>* lo_ = lamA;
    C<T>* li = lamB;
    for (int gi = a.ngates - 1; gi >= 0; --gi) {
      double g0 = 0.0, g1 = 0.0;
      const C<T>* pin = tiles + gi * Gm::PDK;
      const C<T>* pout = pin + Gm::PDK;
      if (gi > 0)
        gate_adj<T, D, K>(a.g[gi], mats + a.g[gi].moff, pin, pout, lo_, a,
                          [li](int P, int64_t, C<T> v) { li[P] = v; }, g0, g1);
      else  // the stage's first gate: lam_in goes straight to the output state
        gate_adj<T, D, K>(a.g[gi], mats + a.g[gi].moff, pin, pout, lo_, a,
                          [&](int, int64_t go, C<T> v) { lam_in[base + go] = v; }, g0, g1);
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
  // one partial per workgroup and gate parameter, at
  // partial[((op * 2 + 0) * 2 + slot) * nwg + blockIdx.x]; the gradient
  // reduction reads the launch's gridDim.x columns of pass 0 only
  for (int idx = threadIdx.x; idx < 2 * a.ngates; idx += Gm::NT) {
    const int gi = idx >> 1, slot = idx & 1;
    double s = 0.0;
    for (int w = 0; w < kWaves; ++w) s += gsl[idx * kWaves + w];
    partial[((int64_t)a.g[gi].op * 4 + slot) * nwg + blockIdx.x] = s;
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
  hipLaunchKernelGGL((k_stage_fwd<T, D, K>), dim3(grid), dim3(Geo<D, K>::NT), lds, s, (const C<T>*)in, (C<T>*)out,
                     (const C<T>*)table, a);
  return (int)hipGetLastError();
}

template <typename T, int D, int K>
int launch_adj(const void* psi, const void* lo, void* li, const void* table, const sfamd::StageArgs& a, int grid,
               double* partial, int nwg, hipStream_t s) {
  const size_t lds =
      (size_t)sfamd::stage_lds_adj(D, K, a.ngates, a.mat_adj, sizeof(C<T>) == 8 ? SFAMD_C64 : SFAMD_C128);
  RET_IF(allow_lds(k_stage_adj<T, D, K>, lds));
  hipLaunchKernelGGL((k_stage_adj<T, D, K>), dim3(grid), dim3(Geo<D, K>::NT), lds, s, (const C<T>*)psi,
                     (const C<T>*)lo, (C<T>*)li, (const C<T>*)table, a, partial, nwg);
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

}  // namespace

// Defines sfamd::stage_fwd_d<d> / stage_adj_d<d>, the launchers of cutoff d
// (both dtypes, every K the tile limit allows); each fused_dNN_MM.hip
// translation unit instantiates a range of cutoffs so the kernels compile in
// parallel.
#define SFAMD_STAGE_DEFINE_D(d)                                                                                   \
  namespace sfamd {                                                                                               \
  int stage_fwd_d##d(int K, const void* in, void* out, const void* table, const StageArgs& a, int grid, int dtype, \
                     hipStream_t s) {                                                                             \
    if (dtype == SFAMD_C64) return fwd_k<float, d>(K, in, out, table, a, grid, s);                                \
    return fwd_k<double, d>(K, in, out, table, a, grid, s);                                                       \
  }                                                                                                               \
  int stage_adj_d##d(int K, const void* psi, const void* lo, void* li, const void* table, const StageArgs& a,      \
                     int grid, double* partial, int nwg, int dtype, hipStream_t s) {                              \
    if (dtype == SFAMD_C64) return adj_k<float, d>(K, psi, lo, li, table, a, grid, partial, nwg, s);              \
    return adj_k<double, d>(K, psi, lo, li, table, a, grid, partial, nwg, s);                                     \
  }                                                                                                               \
  }
