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
// Each thread owns one tile amplitude (or a few, for tiles above 1024) for the
// whole kernel; its digits on the K local axes are computed once, and a gate
// on runtime (workgroup-uniform) axes picks the digits it needs.  A gate's
// output amplitude reads its fiber from the LDS tile and the matrix row from
// the LDS matrix region: R and K one term, D a row of D terms, S the terms of
// the same parity, BS a row of the zero-padded per-fiber matrix (cutoffs up
// to kFullBsMaxD: every lane runs the same D terms, no branches) or of the
// photon-number block s = n0 + n1 of the compact block table.  One small code path per gate kind keeps the instruction
// footprint small (the stage kernels run once per stage on small states).
//
// The stage prologue issues the tile's global loads first, then the loads of
// the stage's matrix image (one contiguous range of the gate table, written
// in LDS layout by the table build), and only then waits, so a stage pays
// about one memory latency before its first gate.
//
// The adjoint of a stage loads the stage's input tile (its checkpoint) and the
// incoming cotangent tile, recomputes the stage forward keeping every gate's
// input and output tile in LDS, and runs the gates backward with each thread's
// amplitude taken as an INPUT index k of the gate:
//   lam_in[k] = sum_i conj(G[i][k]) lam_out[i],  mu[k] = sum_i conj(dG/dp0[i][k]) lam_out[i],
//   g0 += Re conj(mu[k]) psi_in[k],
// and the phase parameter from the identity dG/dphi = i f (out - in) G
// (D: f = 1 on the mode's digit, S: f = 1/2; BS: i (n - N) on the first mode),
//   g1 += f k [Im conj(lam_in[k]) psi_in[k] - Im conj(lam_out[k]) psi_out[k]],
// which needs no second matrix product.  R and K: g = Re sum conj(lam) i w psi.
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
  static constexpr int NT = sfamd::stage_threads(D, K);
  static constexpr int PER = (DK + NT - 1) / NT;  // tile amplitudes per thread
};

// digit of the thread's amplitude on the (uniform, runtime) axis a: the K
// digits are packed 8 bits each (axis t at bit 8 t), so the pick is one
// bitfield extract with a scalar offset
__device__ __forceinline__ int pick(unsigned packed, int a) { return (int)__builtin_amdgcn_ubfe(packed, 8 * a, 8); }

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

// global offset of tile t: the digits of t (radix D, the last non-local
// mode fastest) times the non-local strides.  Tile indices fit in 32 bits.
// The kernels load the strides of up to kNlPre non-local modes with the other
// prologue arguments (one scalar-load round trip); more modes read the rest.
constexpr int kNlPre = 8;
template <int D>
__device__ __forceinline__ int64_t tile_base(int t, const sfamd::StageArgs& a, const int64_t (&st)[kNlPre]) {
  int64_t base = 0;
  const int nnl = a.nnl;
  for (int u = nnl - 1; u >= kNlPre; --u) {  // (rarely) modes beyond the preloaded strides: the fastest
    const int q = t / D;
    base += (int64_t)(t - q * D) * a.nl_stride[u];
    t = q;
  }
#pragma unroll
  for (int u = kNlPre - 1; u >= 0; --u) {
    if (u < nnl) {
      const int q = t / D;
      base += (int64_t)(t - q * D) * st[u];
      t = q;
    }
  }
  return base;
}

// The thread's amplitudes e = threadIdx.x + j NT: their local digits and their
// global offsets within a tile, fixed for the kernel.
template <int D, int K>
struct Mine {
  unsigned dg[Geo<D, K>::PER];  // digit of axis t at bits 8 t
  int64_t goff[Geo<D, K>::PER];
  __device__ __forceinline__ Mine(const sfamd::StageArgs& a) {
#pragma unroll
    for (int j = 0; j < Geo<D, K>::PER; ++j) {
      int r = threadIdx.x + j * Geo<D, K>::NT;
      unsigned p = 0;
#pragma unroll
      for (int t = K - 1; t >= 0; --t) {
        const int q = r / D;
        p |= (unsigned)(r - q * D) << (8 * t);
        r = q;
      }
      dg[j] = p;
      goff[j] = local_offset<D, K>(threadIdx.x + j * Geo<D, K>::NT, a);
    }
  }
};

template <int D, int K>
__device__ __forceinline__ bool owns(int j) {
  return Geo<D, K>::DK % Geo<D, K>::NT == 0 || (int)threadIdx.x + j * Geo<D, K>::NT < Geo<D, K>::DK;
}

// the thread's amplitudes of the tile at global offset b, into registers
template <typename T, int D, int K>
__device__ __forceinline__ void fetch_tile(const C<T>* __restrict__ src, int64_t b, const Mine<D, K>& me,
                                           C<T> (&v)[Geo<D, K>::PER]) {
#pragma unroll
  for (int j = 0; j < Geo<D, K>::PER; ++j)
    if (owns<D, K>(j)) v[j] = src[b + me.goff[j]];
}

// ---- staging of a stage's matrices into LDS (once per workgroup) ----------
// The table build wrote the stage's regions, in LDS layout, as one contiguous
// image (the forward kernel's matrices, or the adjoint's matrices and
// derivatives): staging is a coalesced copy.  The kernels issue the first
// batch of image loads and the tile loads together and wait once.
constexpr int kStageBatch = 8;
template <typename T, int NT>
__device__ __forceinline__ void image_load(const C<T>* __restrict__ img, int total, int b0, C<T> (&v)[kStageBatch]) {
  // unconditional loads (index clamped into the image): no branch separates
  // them, so one wait covers the batch
#pragma unroll
  for (int u = 0; u < kStageBatch; ++u) {
    const int l = b0 + (int)threadIdx.x + u * NT;
    v[u] = img[l < total ? l : total - 1];
  }
}
template <typename T, int NT>
__device__ __forceinline__ void image_store(C<T>* mats, int total, int b0, const C<T> (&v)[kStageBatch]) {
#pragma unroll
  for (int u = 0; u < kStageBatch; ++u) {
    const int l = b0 + (int)threadIdx.x + u * NT;
    if (l < total) mats[l] = v[u];
  }
}
// the image beyond the first batch (large stages only)
template <typename T, int NT>
__device__ __forceinline__ void image_rest(const C<T>* __restrict__ img, int total, C<T>* mats) {
  for (int b0 = kStageBatch * NT; b0 < total; b0 += kStageBatch * NT) {
    C<T> v[kStageBatch];
    image_load<T, NT>(img, total, b0, v);
    image_store<T, NT>(mats, total, b0, v);
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

// ---- forward of one gate: the thread's amplitudes -----------------------------
// `put(j, e, v)` stores the output amplitude of the thread's j-th element e:
// into the next LDS tile, or straight to global memory for the last gate of a
// forward stage.
template <typename T, int D, int K, typename Put>
__device__ __forceinline__ void gate_fwd(const sfamd::StageGate& g, const Mine<D, K>& me,
                                         const C<T>* __restrict__ mat, const C<T>* __restrict__ src, Put put) {
  using Gm = Geo<D, K>;
  const int kind = g.kind, a0 = g.a0;
  const int s0 = g.s0;
#pragma unroll
  for (int j = 0; j < Gm::PER; ++j) {
    if (!owns<D, K>(j)) continue;
    const int e = threadIdx.x + j * Gm::NT;
    const int i = pick(me.dg[j], a0);
    C<T> acc = mk<T>(0, 0);
    if (kind == SFAMD_OP_ROTATION || kind == SFAMD_OP_KERR) {
      acc = cmul(mat[i], src[e]);
    } else if (kind == SFAMD_OP_BEAMSPLITTER) {
      const int a1 = g.a1, s1 = g.s1;
      const int m = pick(me.dg[j], a1);
      const int s = i + m;
      const int ds = s0 - s1;
      if constexpr (D <= sfamd::kFullBsMaxD) {
        // fiber p = s mod D: inputs (k, (p - k) mod D) for every k, weights
        // M[i][k][p] (zero across the two photon-number blocks); input k sits
        // at b0 + k (s0 - s1), plus D s1 where the second digit wraps (k > p)
        const bool upper = s >= D;
        const int p = upper ? s - D : s, wrap = D * s1;
        const int b0 = e - i * ds - (upper ? wrap : 0);
        const C<T>* w = mat + i * D * D + p;  // M[i][k][p] at w[k D]
#pragma unroll
        for (int k = 0; k < D; ++k) cfma(acc, w[k * D], src[b0 + k * ds + (k > p ? wrap : 0)]);
      } else {
        // block s = i + m: out (i, m) <- in (k, s - k), k in [lo, hi], at e + (k - i) (s0 - s1)
        const int lo = bs_lo(s, D), hi = bs_hi(s, D), c = hi - lo + 1;
        const C<T>* row = mat + bs_block_off(s, D) + (i - lo) * c - lo;  // B_s[i][k] at row[k]
        const int bk = e - i * ds;
#pragma unroll
        for (int k = 0; k < D; ++k)
          if (k >= lo && k <= hi) cfma(acc, row[k], src[bk + k * ds]);
      }
    } else if (kind == SFAMD_OP_SQUEEZE) {  // zero unless i + k even
      const int base = e - i * s0;
      const C<T>* row = mat + i * D;
      const int par = i & 1;
#pragma unroll
      for (int h = 0; h < (D + 1) / 2; ++h) {
        const int k = 2 * h + par;
        if (k < D) cfma(acc, row[k], src[base + k * s0]);
      }
    } else {  // displacement: dense row
      const int base = e - i * s0;
      const C<T>* row = mat + i * D;
#pragma unroll
      for (int k = 0; k < D; ++k) cfma(acc, row[k], src[base + k * s0]);
    }
    put(j, e, acc);
  }
}

// ---- adjoint of one gate: the thread's amplitudes as input indices -----------
// `put(j, e, v)` stores lam_in: into the other LDS cotangent tile, or straight
// to global memory for the stage's first gate.
template <typename T, int D, int K, typename Put>
__device__ __forceinline__ void gate_adj(const sfamd::StageGate& g, const Mine<D, K>& me,
                                         const C<T>* __restrict__ mat, const C<T>* __restrict__ pin,
                                         const C<T>* __restrict__ pout, const C<T>* __restrict__ lo_, Put put,
                                         T& g0, T& g1) {
  using Gm = Geo<D, K>;
  constexpr int DD = D * D;
  const int kind = g.kind, a0 = g.a0;
  const int s0 = g.s0;
#pragma unroll
  for (int j = 0; j < Gm::PER; ++j) {
    if (!owns<D, K>(j)) continue;
    const int e = threadIdx.x + j * Gm::NT;
    const int k = pick(me.dg[j], a0);
    const C<T> lk = lo_[e], xo = pout[e], xi = pin[e];
    C<T> acc = mk<T>(0, 0), mu = mk<T>(0, 0);
    T f = (T)0;
    if (kind == SFAMD_OP_ROTATION || kind == SFAMD_OP_KERR) {
      acc = cmul(cconj(mat[k]), lk);
    } else if (kind == SFAMD_OP_BEAMSPLITTER) {
      const int a1 = g.a1, s1 = g.s1;
      const int m = pick(me.dg[j], a1);
      const int s = k + m;
      const int ds = s0 - s1;
      if constexpr (D <= sfamd::kFullBsMaxD) {
        // input (k, m) on fiber p = s mod D feeds outputs (N, (p - N) mod D),
        // weights M[N][k][p], derivatives at D^3 further
        constexpr int D3 = D * D * D;
        const bool upper = s >= D;
        const int p = upper ? s - D : s, wrap = D * s1;
        const int b0 = e - k * ds - (upper ? wrap : 0);
        const C<T>* w = mat + k * D + p;  // M[N][k][p] at w[N D^2]
#pragma unroll
        for (int N = 0; N < D; ++N) {
          const C<T> l = lo_[b0 + N * ds + (N > p ? wrap : 0)];
          cfma_conj(acc, w[N * DD], l);
          cfma_conj(mu, w[D3 + N * DD], l);
        }
      } else {
        // input (k, m) of block s = k + m feeds outputs (N, s - N), N in [lo, hi],
        // at e + (N - k) (s0 - s1)
        const int lo = bs_lo(s, D), hi = bs_hi(s, D), c = hi - lo + 1;
        const int nb = (int)bs_nb(D);
        const C<T>* col = mat + bs_block_off(s, D) + (k - lo) - lo * c;  // B[N][k] at col[N c]
        const int bk = e - k * ds;
#pragma unroll
        for (int N = 0; N < D; ++N) {
          if (N >= lo && N <= hi) {
            const C<T> l = lo_[bk + N * ds];
            cfma_conj(acc, col[N * c], l);
            cfma_conj(mu, col[nb + N * c], l);
          }
        }
      }
      f = (T)-1;  // dB/dphi = i (n - N) B: the opposite sign of D and S
    } else {
      const int base = e - k * s0;
      const C<T>* col = mat + k;
      if (kind == SFAMD_OP_SQUEEZE) {  // i + k even
        const int par = k & 1;
#pragma unroll
        for (int h = 0; h < (D + 1) / 2; ++h) {
          const int i = 2 * h + par;
          if (i < D) {
            const C<T> l = lo_[base + i * s0];
            cfma_conj(acc, col[i * D], l);
            cfma_conj(mu, col[DD + i * D], l);
          }
        }
        f = (T)0.5;
      } else {
#pragma unroll
        for (int i = 0; i < D; ++i) {
          const C<T> l = lo_[base + i * s0];
          cfma_conj(acc, col[i * D], l);
          cfma_conj(mu, col[DD + i * D], l);
        }
        f = (T)1;
      }
    }
    put(j, e, acc);
    // R, K: Re conj(lam) i w psi_out with w = k, k^2 (one parameter)
    const bool diag = kind == SFAMD_OP_ROTATION || kind == SFAMD_OP_KERR;
    const T w = (T)(kind == SFAMD_OP_KERR ? k * k : k);
    const T d0 = diag ? -w * im_conj_mul(lk, xo) : re_conj_mul(mu, xi);
    const T d1 = f * (T)k * (im_conj_mul(acc, xi) - im_conj_mul(lk, xo));
    g0 += d0;
    g1 += d1;
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

// the same reduction in single precision (complex64 stages: 64 lanes' partial
// products, well inside single precision; the per-wave sums are added in
// double), one DPP add per step
template <int Ctrl, int RowMask>
__device__ __forceinline__ float dpp_move_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, RowMask, 0xf, false));
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_move_f<0xB1, 0xF>(v);
  v += dpp_move_f<0x4E, 0xF>(v);
  v += dpp_move_f<0x141, 0xF>(v);
  v += dpp_move_f<0x140, 0xF>(v);
  v += dpp_move_f<0x142, 0xA>(v);
  v += dpp_move_f<0x143, 0xC>(v);
  return v;
}

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8),
// and each XCD has its own L2.  Neighbouring tiles share cache lines whenever
// the innermost non-local mode has a small stride, so the tiles are cut into 8
// contiguous ranges and XCD x's blocks (b = x + 8 k) walk range x; every line
// is then fetched into one L2 instead of eight.
struct TileRange {
  int first, end, step;
};

__device__ __forceinline__ TileRange xcd_tiles(int ntiles, int G) {
  const int b = blockIdx.x;
  if (G % 8) return TileRange{b, ntiles, G};
  const int per = (ntiles + 7) / 8;
  const int x = b % 8;
  const int end = (x + 1) * per < ntiles ? (x + 1) * per : ntiles;
  return TileRange{x * per + b / 8, end, G / 8};
}

// ---- kernels ------------------------------------------------------------
template <typename T, int D, int K>
__global__ void __launch_bounds__(sfamd::stage_threads(D, K)) k_stage_fwd(const C<T>* __restrict__ in, C<T>* __restrict__ out,
                                                             const C<T>* __restrict__ table, sfamd::StageArgs a) {
  using Gm = Geo<D, K>;
  int64_t st[kNlPre];
#pragma unroll
  for (int u = 0; u < kNlPre; ++u) st[u] = a.nl_stride[u];
  extern __shared__ __align__(16) unsigned char smem[];
  C<T>* bufA = reinterpret_cast<C<T>*>(smem);
  C<T>* bufB = bufA + Gm::DK;
  C<T>* mats = bufB + Gm::DK;
  const TileRange tr = xcd_tiles(a.ntiles, a.grid);
  const Mine<D, K> me(a);
  // the first tile's loads are in flight while the matrices are staged, and
  // each next tile's while the gates of the current one run
  C<T> pre[Gm::PER];
#pragma unroll
  for (int j = 0; j < Gm::PER; ++j) pre[j] = mk<T>(0, 0);
  const C<T>* img = table + a.img_fwd;
  C<T> iv[kStageBatch];
  image_load<T, Gm::NT>(img, a.mat_fwd, 0, iv);
  fetch_tile<T, D, K>(in, tile_base<D>(tr.first < tr.end ? tr.first : 0, a, st), me, pre);
  image_store<T, Gm::NT>(mats, a.mat_fwd, 0, iv);
  image_rest<T, Gm::NT>(img, a.mat_fwd, mats);
  for (int t = tr.first; t < tr.end; t += tr.step) {
    const int64_t base = tile_base<D>(t, a, st);
#pragma unroll
    for (int j = 0; j < Gm::PER; ++j)
      if (owns<D, K>(j)) bufA[threadIdx.x + j * Gm::NT] = pre[j];
    __syncthreads();
    if (t + tr.step < tr.end) fetch_tile<T, D, K>(in, tile_base<D>(t + tr.step, a, st), me, pre);
    C<T>* src = bufA;
    C<T>* dst = bufB;
    for (int gi = 0; gi < a.ngates - 1; ++gi) {
      const sfamd::StageGate& g = a.g[gi];
      gate_fwd<T, D, K>(g, me, mats + g.moff, src, [dst](int, int e, C<T> v) { dst[e] = v; });
      __syncthreads();
      C<T>* tmp = src;
      src = dst;
      dst = tmp;
    }
    // the last gate writes its outputs straight to the state
    const sfamd::StageGate& gl = a.g[a.ngates - 1];
    gate_fwd<T, D, K>(gl, me, mats + gl.moff, src, [&](int j, int, C<T> v) { out[base + me.goff[j]] = v; });
    __syncthreads();
  }
}

template <typename T, int D, int K>
__global__ void __launch_bounds__(sfamd::stage_threads(D, K)) k_stage_adj(const C<T>* __restrict__ psi_in,
                                                             const C<T>* __restrict__ lam_out,
                                                             C<T>* __restrict__ lam_in,
                                                             const C<T>* __restrict__ table, sfamd::StageArgs a,
                                                             double* __restrict__ partial, int nwg) {
  using Gm = Geo<D, K>;
  int64_t st[kNlPre];
#pragma unroll
  for (int u = 0; u < kNlPre; ++u) st[u] = a.nl_stride[u];
  extern __shared__ __align__(16) unsigned char smem[];
  // layout: [ngates + 1 psi tiles (every gate's input, then the last output)]
  //         [2 lam tiles][matrices + derivatives][grad slots [ngates][2][waves]]
  C<T>* tiles = reinterpret_cast<C<T>*>(smem);
  C<T>* lamA = tiles + (a.ngates + 1) * Gm::DK;
  C<T>* lamB = lamA + Gm::DK;
  C<T>* mats = lamB + Gm::DK;
  double* gsl = reinterpret_cast<double*>(mats + a.mat_adj);
  constexpr int kWaves = Gm::NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const TileRange tr = xcd_tiles(a.ntiles, a.grid);
  const Mine<D, K> me(a);
  // tile loads in flight while the matrices are staged (first tile) and while
  // the current tile's gates run (the next one)
  C<T> pre_psi[Gm::PER], pre_lam[Gm::PER];
#pragma unroll
  for (int j = 0; j < Gm::PER; ++j) pre_psi[j] = pre_lam[j] = mk<T>(0, 0);
  const C<T>* img = table + a.img_adj;
  C<T> iv[kStageBatch];
  image_load<T, Gm::NT>(img, a.mat_adj, 0, iv);
  {
    const int64_t b = tile_base<D>(tr.first < tr.end ? tr.first : 0, a, st);
    fetch_tile<T, D, K>(psi_in, b, me, pre_psi);
    fetch_tile<T, D, K>(lam_out, b, me, pre_lam);
  }
  for (int idx = threadIdx.x; idx < a.ngates * 2 * kWaves; idx += Gm::NT) gsl[idx] = 0.0;
  image_store<T, Gm::NT>(mats, a.mat_adj, 0, iv);
  image_rest<T, Gm::NT>(img, a.mat_adj, mats);
  for (int t = tr.first; t < tr.end; t += tr.step) {
    const int64_t base = tile_base<D>(t, a, st);
#pragma unroll
    for (int j = 0; j < Gm::PER; ++j) {
      if (owns<D, K>(j)) {
        tiles[threadIdx.x + j * Gm::NT] = pre_psi[j];
        lamA[threadIdx.x + j * Gm::NT] = pre_lam[j];
      }
    }
    __syncthreads();
    if (t + tr.step < tr.end) {
      const int64_t b = tile_base<D>(t + tr.step, a, st);
      fetch_tile<T, D, K>(psi_in, b, me, pre_psi);
      fetch_tile<T, D, K>(lam_out, b, me, pre_lam);
    }
    // recompute the stage forward, keeping every gate's input and output
    for (int gi = 0; gi < a.ngates; ++gi) {
      const sfamd::StageGate& g = a.g[gi];
      C<T>* dst = tiles + (gi + 1) * Gm::DK;
      gate_fwd<T, D, K>(g, me, mats + g.amoff, tiles + gi * Gm::DK, [dst](int, int e, C<T> v) { dst[e] = v; });
      // the last gate's output is read only by the adjoint of that gate, at
      // each thread's own elements (gate_fwd writes, and gate_adj reads pin and
      // pout, at the thread's own e), so it needs no barrier
      if (gi + 1 < a.ngates) __syncthreads();
    }
    C<T>* lo_ = lamA;
    C<T>* li = lamB;
    for (int gi = a.ngates - 1; gi >= 0; --gi) {
      const sfamd::StageGate& g = a.g[gi];
      // per-thread parameter sums in the state's precision, reduced per wave
      // and then added in double
      T g0 = (T)0, g1 = (T)0;
      const C<T>* pin = tiles + gi * Gm::DK;
      const C<T>* pout = pin + Gm::DK;
      if (gi > 0)
        gate_adj<T, D, K>(g, me, mats + g.amoff, pin, pout, lo_, [li](int, int e, C<T> v) { li[e] = v; }, g0, g1);
      else  // the stage's first gate: lam_in goes straight to the output state
        gate_adj<T, D, K>(g, me, mats + g.amoff, pin, pout, lo_,
                          [&](int j, int, C<T> v) { lam_in[base + me.goff[j]] = v; }, g0, g1);
      // R and K have one parameter: their second slot stays at its cleared zero
      const bool two = g.kind != SFAMD_OP_ROTATION && g.kind != SFAMD_OP_KERR;
      g0 = wave_sum(g0);
      if (two) g1 = wave_sum(g1);
      if (lane == kWaveSumLane) {  // each wave owns its slot: deterministic, no atomics
        gsl[(gi * 2 + 0) * kWaves + wave] += (double)g0;
        if (two) gsl[(gi * 2 + 1) * kWaves + wave] += (double)g1;
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
  sfamd::StageArgs b = a;
  b.grid = grid;
  hipLaunchKernelGGL((k_stage_fwd<T, D, K>), dim3(grid), dim3(Geo<D, K>::NT), lds, s, (const C<T>*)in, (C<T>*)out,
                     (const C<T>*)table, b);
  return (int)hipGetLastError();
}

template <typename T, int D, int K>
int launch_adj(const void* psi, const void* lo, void* li, const void* table, const sfamd::StageArgs& a, int grid,
               double* partial, int nwg, hipStream_t s) {
  const size_t lds =
      (size_t)sfamd::stage_lds_adj(D, K, a.ngates, a.mat_adj, sizeof(C<T>) == 8 ? SFAMD_C64 : SFAMD_C128);
  RET_IF(allow_lds(k_stage_adj<T, D, K>, lds));
  sfamd::StageArgs b = a;
  b.grid = grid;
  hipLaunchKernelGGL((k_stage_adj<T, D, K>), dim3(grid), dim3(Geo<D, K>::NT), lds, s, (const C<T>*)psi,
                     (const C<T>*)lo, (C<T>*)li, (const C<T>*)table, b, partial, nwg);
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
