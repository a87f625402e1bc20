// LDS-staged kernels for two-axis gates whose inner axis is (nearly) innermost.
//
// When the inner gate axis has stride < 16, consecutive fibers of a
// beamsplitter (the photon-number blocks of fockbackend/circuit.py:306-335)
// or of the loss channel (diagonals of one mode's (bra, ket) pair,
// fockbackend/circuit.py:65-87) share cache lines, so reading fibers straight
// from HBM gathers.  Here a workgroup copies a tile into LDS with coalesced
// loads, transforms every fiber of the tile in place, and writes the tile back
// with coalesced stores.
//
// Tile geometry.  A slab is the contiguous range spanned by the outer gate
// axis: D * S_out amplitudes, made of the outer gate axis (stride S_out), the
// axes between the two gate axes ("mid", Mid values, stride RW = D * Sin) and a
// run of RW amplitudes holding the inner gate axis (stride Sin) and the axes
// inside it.  A tile holds G whole slabs (contiguous), or one slab restricted
// to Mt consecutive mid values (D runs of Mt * RW contiguous amplitudes) when a
// slab is larger than the LDS budget.  In LDS the tile is dense in the order
// (slab, outer axis, mid, run).
//
// The element-wise (VALU) kernels stage the gate coefficients (beamsplitter
// blocks, loss coefficients) in LDS as well when they fit, since lanes of one
// wave work on different blocks; the matrix-core kernels build each block's
// operator once per tile from the table in global memory.
#include "common.h"
#include "mfma_ops.h"
#include "slab.h"

namespace {

template <typename T>
using C = cpx<T>;

constexpr int kSlabBlock = 512;
constexpr int64_t kSlabLds = 150 * 1024;  // per workgroup (gfx950: 160 KB)

struct Tile {
  int64_t slab, S_out, nslab, ntile;
  int RW, Sin, Mid, Mt, G, nmc, cap;
};

struct View {
  int64_t base;
  int gs, mt, run, elems;
};

__device__ __forceinline__ View tile_view(const Tile& t, int64_t tile, int D) {
  View v;
  const int64_t sg = tile / t.nmc;
  const int mc = (int)(tile - sg * t.nmc);
  const int64_t left = t.nslab - sg * t.G;
  v.gs = left < t.G ? (int)left : t.G;
  const int m0 = mc * t.Mt;
  v.mt = t.Mid - m0 < t.Mt ? t.Mid - m0 : t.Mt;
  v.run = v.mt * t.RW;
  v.base = sg * t.G * t.slab + (int64_t)m0 * t.RW;
  v.elems = v.gs * D * v.run;
  return v;
}

// element e of the tile -> offset from v.base; runs of v.run are contiguous
__device__ __forceinline__ int64_t tile_off(const Tile& t, const View& v, int e, int D) {
  if (v.run * D == t.slab) return e;  // whole slabs: the tile is contiguous
  const int j = e / v.run;             // run index = slab * D + outer-axis position
  const int r = e - j * v.run;
  const int gsl = j / D;
  return gsl * t.slab + (int64_t)(j - gsl * D) * t.S_out + r;
}

// Tile copies move 16-byte chunks, eight nontemporal loads in flight per
// thread before the LDS writes.  A tile is gs * D runs of v.run contiguous
// amplitudes (run j = slab * D + outer-axis position starts at
// (j / D) * slab + (j % D) * S_out); tiles whose runs are not whole 16-byte
// chunks fall back to amplitude-by-amplitude copies.
typedef float f4s __attribute__((ext_vector_type(4)));
constexpr int kCopyBatch = 8;

template <typename T>
__device__ __forceinline__ bool tile_chunked(const Tile& t, const View& v, const C<T>* g) {
  const uintptr_t esz = sizeof(C<T>);
  return ((((uintptr_t)(g + v.base)) | (uintptr_t)(v.run * esz) | (uintptr_t)(t.S_out * esz)) & 15) == 0;
}

// byte offset from g + v.base of 16-byte chunk c (cpr chunks per run);
// cpr == 0 marks a contiguous tile
template <typename T>
__device__ __forceinline__ int64_t chunk_off(const Tile& t, int c, int cpr, int D) {
  if (cpr == 0) return (int64_t)c * 16;
  const int j = c / cpr;
  const int r = c - j * cpr;
  const int gsl = j / D;
  return ((int64_t)gsl * t.slab + (int64_t)(j - gsl * D) * t.S_out) * (int64_t)sizeof(C<T>) + (int64_t)r * 16;
}

template <typename T>
__device__ __forceinline__ void tile_load(C<T>* sh, const C<T>* __restrict__ g, const Tile& t, const View& v, int D) {
  if (tile_chunked(t, v, g)) {
    const char* src = (const char*)(g + v.base);
    f4s* dst = (f4s*)sh;
    const int cpr = v.run * D == t.slab ? 0 : (int)(v.run * sizeof(C<T>) / 16);
    const int n = (int)(v.elems * sizeof(C<T>) / 16);
    // uniform trip count; every load is issued (chunk index clamped into the
    // tile) so that no branch separates them: conditional loads into the
    // batch array made the compiler wait for each load before the next
    for (int b0 = 0; b0 < n; b0 += kCopyBatch * (int)blockDim.x) {
      f4s r[kCopyBatch];
#pragma unroll
      for (int k = 0; k < kCopyBatch; ++k) {
        const int c = b0 + (int)threadIdx.x + k * (int)blockDim.x;
        r[k] = __builtin_nontemporal_load((const f4s*)(src + chunk_off<T>(t, c < n ? c : n - 1, cpr, D)));
      }
#pragma unroll
      for (int k = 0; k < kCopyBatch; ++k) {
        const int c = b0 + (int)threadIdx.x + k * (int)blockDim.x;
        if (c < n) dst[c] = r[k];
      }
    }
    return;
  }
  for (int e = threadIdx.x; e < v.elems; e += blockDim.x) sh[e] = g[v.base + tile_off(t, v, e, D)];
}

template <typename T>
__device__ __forceinline__ void tile_store(const C<T>* sh, C<T>* __restrict__ g, const Tile& t, const View& v, int D) {
  if (tile_chunked(t, v, (const C<T>*)g)) {
    char* dst = (char*)(g + v.base);
    const f4s* src = (const f4s*)sh;
    const int cpr = v.run * D == t.slab ? 0 : (int)(v.run * sizeof(C<T>) / 16);
    const int n = (int)(v.elems * sizeof(C<T>) / 16);
    for (int c0 = threadIdx.x; c0 < n; c0 += kCopyBatch * blockDim.x) {
#pragma unroll
      for (int k = 0; k < kCopyBatch; ++k) {
        const int c = c0 + k * (int)blockDim.x;
        if (c < n) __builtin_nontemporal_store(src[c], (f4s*)(dst + chunk_off<T>(t, c, cpr, D)));
      }
    }
    return;
  }
  for (int e = threadIdx.x; e < v.elems; e += blockDim.x) g[v.base + tile_off(t, v, e, D)] = sh[e];
}

template <typename T>
__device__ __forceinline__ void stage(T* dst, const T* __restrict__ src, int64_t n) {
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ void write_partials(double g0, double g1, double* partial, int nwg, int zp1) {
  __shared__ double sh[8];
  const double s0 = block_sum(g0, sh);
  const double s1 = block_sum(g1, sh);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s0;
    partial[nwg + blockIdx.x] = s1;
    if (zp1) {
      partial[2 * nwg + blockIdx.x] = 0.0;
      partial[3 * nwg + blockIdx.x] = 0.0;
    }
  }
}

// fiber phi of a beamsplitter tile -> block s and the LDS offset of amplitude
// N = lo on the p axis; step moves N -> N + 1
struct BsFiber {
  int s, lo, c;
  int pos0, step;
};

__device__ __forceinline__ BsFiber bs_fiber(int phi, int per_s, const Tile& t, const View& v, int p_outer, int D,
                                            int diff = 0) {
  BsFiber b;
  b.s = phi / per_s;
  const int r = phi - b.s * per_s;
  const int cidx = r % t.Sin;
  const int r2 = r / t.Sin;
  const int m = r2 % v.mt;
  const int gsl = r2 / v.mt;
  const int sA = v.run;  // local stride of the outer gate axis
  const int lp = p_outer ? sA : t.Sin, lq = p_outer ? t.Sin : sA;
  b.lo = bs_lo(b.s, D);
  b.c = bs_hi(b.s, D) - b.lo + 1;
  // photon-number block: amplitudes (k, s - k); difference block (diff): (k, k - d)
  const int q0 = diff ? b.lo - (b.s - (D - 1)) : b.s - b.lo;
  b.pos0 = gsl * D * sA + m * t.RW + cidx + b.lo * lp + q0 * lq;
  b.step = diff ? lp + lq : lp - lq;
  return b;
}

// ------------------------------------------------------------------ beamsplitter
template <typename T, int D, bool ML>
__global__ void __launch_bounds__(kSlabBlock)
k_bs_slab(const C<T>* __restrict__ in, C<T>* __restrict__ out, Tile t, int p_outer, const C<T>* __restrict__ blk,
          int conj) {
  extern __shared__ __align__(16) unsigned char smem[];
  C<T>* sh = (C<T>*)smem;
  const C<T>* mat = blk;
  if (ML) {
    C<T>* m = sh + t.cap;
    stage(m, blk, bs_nb(D));
    mat = m;
  }
  for (int64_t tile = blockIdx.x; tile < t.ntile; tile += gridDim.x) {
    const View v = tile_view(t, tile, D);
    tile_load(sh, in, t, v, D);
    __syncthreads();
    const int per_s = v.gs * v.mt * t.Sin;
    const int nf = per_s * (2 * D - 1);
    for (int phi = threadIdx.x; phi < nf; phi += blockDim.x) {
      const BsFiber f = bs_fiber(phi, per_s, t, v, p_outer, D);
      const C<T>* bb = mat + bs_block_off(f.s, D);
      C<T> x[D];
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (k < f.c) x[k] = sh[f.pos0 + k * f.step];
#pragma unroll
      for (int u = 0; u < D; ++u) {
        if (u >= f.c) continue;
        C<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
          if (k >= f.c) continue;
          if (conj)
            cmac_conj(acc, bb[u * f.c + k], x[k]);
          else
            cmac(acc, bb[u * f.c + k], x[k]);
        }
        sh[f.pos0 + u * f.step] = acc;
      }
    }
    __syncthreads();
    tile_store(sh, out, t, v, D);
    __syncthreads();
  }
}

template <typename T, int D, bool ML>
__global__ void __launch_bounds__(kSlabBlock)
k_bs_slab_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, Tile t, int p_outer,
              const C<T>* __restrict__ blk, const C<T>* __restrict__ dblk, int conj, int wa, int wb, double* partial,
              int nwg, int zp1) {
  extern __shared__ __align__(16) unsigned char smem[];
  C<T>* shx = (C<T>*)smem;
  C<T>* shl = shx + t.cap;
  const C<T>* mat = blk;
  const C<T>* dmat = dblk;
  if (ML) {
    C<T>* m = shl + t.cap;
    stage(m, blk, bs_nb(D));
    stage(m + bs_nb(D), dblk, bs_nb(D));
    mat = m;
    dmat = m + bs_nb(D);
  }
  const T sg = conj ? (T)-1 : (T)1;
  double g0 = 0.0, g1 = 0.0;
  for (int64_t tile = blockIdx.x; tile < t.ntile; tile += gridDim.x) {
    const View v = tile_view(t, tile, D);
    tile_load(shx, psi, t, v, D);
    tile_load(shl, lo_, t, v, D);
    __syncthreads();
    const int per_s = v.gs * v.mt * t.Sin;
    const int nf = per_s * (2 * D - 1);
    for (int phi = threadIdx.x; phi < nf; phi += blockDim.x) {
      const BsFiber f = bs_fiber(phi, per_s, t, v, p_outer, D);
      const int64_t off = bs_block_off(f.s, D);
      const C<T>* bb = mat + off;
      const C<T>* db = dmat + off;
      C<T> x[D], l[D];
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (k < f.c) {
          x[k] = shx[f.pos0 + k * f.step];
          l[k] = shl[f.pos0 + k * f.step];
        }
      T s0 = 0, s1 = 0;
#pragma unroll
      for (int n = 0; n < D; ++n) {
        if (n >= f.c) continue;
        // lam_in[n] = sum_N conj(B'[N][n]) lam[N]; w0 = sum_N conj(lam[N]) dB'[N][n];
        // w1 = sum_N conj(lam[N]) i sg (wa n + wb N) B'[N][n]
        C<T> acc = mk<T>(0, 0), w0 = mk<T>(0, 0), w1 = mk<T>(0, 0);
#pragma unroll
        for (int N = 0; N < D; ++N) {
          if (N >= f.c) continue;
          C<T> b = bb[N * f.c + n];
          C<T> d = db[N * f.c + n];
          if (conj) {
            b.im = -b.im;
            d.im = -d.im;
          }
          cmac_conj(acc, b, l[N]);
          cmac_conj(w0, l[N], d);
          const C<T> lb = mk<T>(l[N].re * b.re + l[N].im * b.im, l[N].re * b.im - l[N].im * b.re);
          const T wt = sg * (T)(wa * (f.lo + n) + wb * (f.lo + N));
          w1.re += -wt * lb.im;
          w1.im += wt * lb.re;
        }
        s0 += w0.re * x[n].re - w0.im * x[n].im;
        s1 += w1.re * x[n].re - w1.im * x[n].im;
        x[n] = acc;  // lam_in, written back after the loop (l is still needed)
      }
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (k < f.c) shl[f.pos0 + k * f.step] = x[k];
      g0 += (double)s0;
      g1 += (double)s1;
    }
    __syncthreads();
    tile_store(shl, li, t, v, D);
    __syncthreads();
  }
  write_partials(g0, g1, partial, nwg, zp1);
}

// ------------------------------------------------------------------ loss
// fiber phi of a loss tile -> diagonal d = q - p and the LDS offset of its
// first amplitude; step moves (p, q) -> (p + 1, q + 1).  Mid = 1 here.
struct LossFiber {
  int d, p0, L, pos0, step;
};

__device__ __forceinline__ LossFiber loss_fiber(int phi, int per_d, const Tile& t, const View& v, int D) {
  LossFiber f;
  const int di = phi / per_d;
  const int r = phi - di * per_d;
  const int cidx = r % t.Sin;
  const int gsl = r / t.Sin;
  f.d = di - (D - 1);
  f.p0 = f.d < 0 ? -f.d : 0;
  f.L = D - (f.d < 0 ? -f.d : f.d);
  const int sA = v.run;  // bra stride in the tile (= D * Sin)
  f.pos0 = gsl * D * sA + cidx + f.p0 * sA + (f.p0 + f.d) * t.Sin;
  f.step = sA + t.Sin;
  return f;
}

template <typename T, int D, bool ML>
__global__ void __launch_bounds__(kSlabBlock)
k_loss_slab(const C<T>* __restrict__ in, C<T>* __restrict__ out, Tile t, const T* __restrict__ coef) {
  extern __shared__ __align__(16) unsigned char smem[];
  C<T>* sh = (C<T>*)smem;
  const T* cf = coef;
  if (ML) {
    T* m = (T*)(sh + t.cap);
    stage(m, coef, (int64_t)D * D * D);
    cf = m;
  }
  for (int64_t tile = blockIdx.x; tile < t.ntile; tile += gridDim.x) {
    const View v = tile_view(t, tile, D);
    tile_load(sh, in, t, v, D);
    __syncthreads();
    const int per_d = v.gs * t.Sin;
    const int nf = per_d * (2 * D - 1);
    for (int phi = threadIdx.x; phi < nf; phi += blockDim.x) {
      const LossFiber f = loss_fiber(phi, per_d, t, v, D);
      const T* c0 = cf + ((f.p0 * D) + f.p0 + f.d) * D;
      C<T> x[D];
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (k < f.L) x[k] = sh[f.pos0 + k * f.step];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (k >= f.L) continue;
        const T* c = c0 + k * (D + 1) * D;
        C<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int l = 0; l < D; ++l) {
          if (k + l >= f.L) continue;
          acc.re += c[l] * x[k + l].re;
          acc.im += c[l] * x[k + l].im;
        }
        sh[f.pos0 + k * f.step] = acc;
      }
    }
    __syncthreads();
    tile_store(sh, out, t, v, D);
    __syncthreads();
  }
}

template <typename T, int D, bool ML>
__global__ void __launch_bounds__(kSlabBlock)
k_loss_slab_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, Tile t,
                const T* __restrict__ coef, const T* __restrict__ dcoef, double* partial, int nwg, int zp1) {
  extern __shared__ __align__(16) unsigned char smem[];
  C<T>* shx = (C<T>*)smem;
  C<T>* shl = shx + t.cap;
  const T* cf = coef;
  const T* dcf = dcoef;
  if (ML) {
    T* m = (T*)(shl + t.cap);
    stage(m, coef, (int64_t)D * D * D);
    stage(m + D * D * D, dcoef, (int64_t)D * D * D);
    cf = m;
    dcf = m + D * D * D;
  }
  double g0 = 0.0;
  for (int64_t tile = blockIdx.x; tile < t.ntile; tile += gridDim.x) {
    const View v = tile_view(t, tile, D);
    tile_load(shx, psi, t, v, D);
    tile_load(shl, lo_, t, v, D);
    __syncthreads();
    const int per_d = v.gs * t.Sin;
    const int nf = per_d * (2 * D - 1);
    for (int phi = threadIdx.x; phi < nf; phi += blockDim.x) {
      const LossFiber f = loss_fiber(phi, per_d, t, v, D);
      const int c0 = ((f.p0 * D) + f.p0 + f.d) * D;  // coefficient row of position 0
      C<T> x[D], l[D];
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (k < f.L) {
          x[k] = shx[f.pos0 + k * f.step];
          l[k] = shl[f.pos0 + k * f.step];
        }
      T s0 = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (k >= f.L) continue;
        // gradient: Re conj(lam[k]) sum_l dc[k][l] x[k + l]
        const T* dc = dcf + c0 + k * (D + 1) * D;
        C<T> u = mk<T>(0, 0);
#pragma unroll
        for (int l = 0; l < D; ++l) {
          if (k + l >= f.L) continue;
          u.re += dc[l] * x[k + l].re;
          u.im += dc[l] * x[k + l].im;
        }
        s0 += re_conj_mul(l[k], u);
      }
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (k >= f.L) continue;
        // lam_in[k] = sum over outputs u = k - lp (lp >= 0) of c[u][lp] lam[u]
        C<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int lp = 0; lp < D; ++lp) {
          if (lp > k) continue;
          const T cc = cf[c0 + (k - lp) * (D + 1) * D + lp];
          acc.re += cc * l[k - lp].re;
          acc.im += cc * l[k - lp].im;
        }
        shl[f.pos0 + k * f.step] = acc;
      }
      g0 += (double)s0;
    }
    __syncthreads();
    tile_store(shl, li, t, v, D);
    __syncthreads();
  }
  write_partials(g0, 0.0, partial, nwg, zp1);
}

// ------------------------------------------------------------------ matrix-core variants (complex64)
// The same tiles, with each block (beamsplitter photon number s, loss
// diagonal d) applied as a real-form MFMA operator: a wave takes whole blocks,
// builds the block's operator once and applies it to 16 fibers of the tile per
// chain, reading and writing the tile in place.  A policy describes the gate:
// block count, block size, fiber positions and operator entries.
using sfamd::mf::Blk;
using sfamd::mf::f4;
using Cf = cpx<float>;

template <int D>
struct BsPol {
  static constexpr int NB = 2 * D - 1;
  static constexpr bool kTwoParams = true;
  const Cf* blk;
  const Cf* dblk;
  int conj, p_outer;
  int wa, wb;  // second-parameter derivative weight w = wa n + wb N
  int diff;    // difference-conserving blocks (two-mode squeezing)
  __device__ int size(int s) const { return bs_hi(s, D) - bs_lo(s, D) + 1; }
  __device__ int per_block(const Tile& t, const View& v) const { return v.gs * v.mt * t.Sin; }
  __device__ void fiber(int s, int phi, const Tile& t, const View& v, int& pos0, int& step) const {
    const BsFiber f = bs_fiber(s * per_block(t, v) + phi, per_block(t, v), t, v, p_outer, D, diff);
    pos0 = f.pos0;
    step = f.step;
  }
  __device__ Cf entry(int s, int n, int i, int k) const {
    return sfamd::mf::op_entry(blk + bs_block_off(s, D), n, conj ? 1 : 0, i, k);
  }
  __device__ Cf entry_h(int s, int n, int i, int k) const {
    return sfamd::mf::op_entry(blk + bs_block_off(s, D), n, conj ? 3 : 2, i, k);
  }
  __device__ Cf dentry0(int s, int n, int i, int k) const {
    return sfamd::mf::op_entry(dblk + bs_block_off(s, D), n, conj ? 1 : 0, i, k);
  }
  // second-parameter derivative i sg (wa n + wb N) B'[N][n], sg = -1 on the conjugated side
  __device__ Cf dentry1(int s, int n, int i, int k) const {
    const Cf m = entry(s, n, i, k);
    const int lo = bs_lo(s, D);
    const float w = (conj ? -1.f : 1.f) * (float)(wa * (lo + k) + wb * (lo + i));
    return mk<float>(-w * m.im, w * m.re);
  }
};

template <int D>
struct LossPol {
  static constexpr int NB = 2 * D - 1;
  static constexpr bool kTwoParams = false;
  const float* coef;
  const float* dcoef;
  __device__ int size(int di) const {
    const int d = di - (D - 1);
    return D - (d < 0 ? -d : d);
  }
  __device__ int per_block(const Tile& t, const View& v) const { return v.gs * t.Sin; }
  __device__ void fiber(int di, int phi, const Tile& t, const View& v, int& pos0, int& step) const {
    const LossFiber f = loss_fiber(di * per_block(t, v) + phi, per_block(t, v), t, v, D);
    pos0 = f.pos0;
    step = f.step;
  }
  // out[i] = sum_l c[p0 + i][p0 + i + d][l] x[i + l]: entry (i, k = i + l);
  // zero outside 0 <= i <= k < n, the load always issued (indices clamped
  // into that range, as in op_entry)
  __device__ Cf coef_entry(const float* cf, int di, int n, int i, int k) const {
    const int d = di - (D - 1);
    const int p0 = d < 0 ? -d : 0;
    const int ic = i < n ? i : n - 1;
    const int l = k - i;
    const int lc = l < 0 ? 0 : (l > n - 1 - ic ? n - 1 - ic : l);
    const float c = cf[((p0 * D) + p0 + d) * D + ic * (D + 1) * D + lc];
    return mk<float>((i < n && k < n && k >= i) ? c : 0.f, 0.f);
  }
  __device__ Cf entry(int di, int n, int i, int k) const { return coef_entry(coef, di, n, i, k); }
  __device__ Cf entry_h(int di, int n, int i, int k) const { return coef_entry(coef, di, n, k, i); }
  __device__ Cf dentry0(int di, int n, int i, int k) const { return coef_entry(dcoef, di, n, i, k); }
  __device__ Cf dentry1(int, int, int, int) const { return mk<float>(0.f, 0.f); }
};

// PA = group_amps(n), the amplitudes per lane group, is Q - 1 or Q for every
// block size n; the blocks are instantiated for both so that the operator
// products stop after 2 PA k-steps without a branch per k-step.
template <int MB, int PA, int D, class P>
__device__ __forceinline__ void slab_block(Cf* sh, const P& pol, int b, int n, const Tile& t, const View& v) {
  constexpr int Q = Blk<MB>::Q;
  float A[MB][Blk<MB>::NS];
  constexpr int pa = PA;
  sfamd::mf::load_op<MB>([&](int i, int k) { return pol.entry(b, n, i, k); }, A, pa);
  const int per = pol.per_block(t, v);
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15, g = lane >> 4;
  for (int fb = 0; fb * 16 < per; ++fb) {
    const int phi = fb * 16 + cl;
    const bool valid = phi < per;
    int pos0, step;
    pol.fiber(b, valid ? phi : 0, t, v, pos0, step);
    Cf x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int a = g * pa + q;
      const Cf xv = sh[pos0 + sfamd::mf::clamp_amp(a, n) * step];
      x[q] = (valid && q < pa && a < n) ? xv : mk<float>(0.f, 0.f);
    }
    pos0 += g * pa * step;
    f4 acc[MB];
    sfamd::mf::apply_op<MB>(A, x, acc, pa);
    if (valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q < pa && g * pa + q < n) sh[pos0 + q * step] = sfamd::mf::acc_at<MB>(acc, q);
    }
  }
}

template <int D, class P>
__global__ void __launch_bounds__(256) k_slab_mfma(const Cf* __restrict__ in, Cf* __restrict__ out, Tile t, P gpol) {
  extern __shared__ __align__(16) unsigned char smem[];
  Cf* sh = (Cf*)smem;
  const P pol = gpol;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int64_t tile = blockIdx.x; tile < t.ntile; tile += gridDim.x) {
    const View v = tile_view(t, tile, D);
    tile_load(sh, in, t, v, D);
    __syncthreads();
    for (int b = wave; b < P::NB; b += nw) {
      const int n = pol.size(b);
      const int mb = (2 * n + 15) / 16;
#define SFAMD_SLAB_BLOCK(MB)                            \
  if (sfamd::mf::group_amps(n) == 2 * MB)               \
    slab_block<MB, 2 * MB, D>(sh, pol, b, n, t, v);     \
  else                                                  \
    slab_block<MB, 2 * MB - 1, D>(sh, pol, b, n, t, v);
      if (mb == 1) {
        SFAMD_SLAB_BLOCK(1)
      } else if constexpr (D > 8) {
        if (mb == 2) {
          SFAMD_SLAB_BLOCK(2)
        } else if constexpr (D > 16) {
          SFAMD_SLAB_BLOCK(3)
        }
      }
#undef SFAMD_SLAB_BLOCK
    }
    __syncthreads();
    tile_store(sh, out, t, v, D);
    __syncthreads();
  }
}

// adjoint of one block in one sweep over its fibers: the gradient operators
// (dB/dp0, dB/dp1) act on the forward fiber x and accumulate Re conj(lam) (A x);
// then lam_in = B^H lam overwrites lam in place (each fiber belongs to this wave
// alone).  The three operators live in registers: the kernel runs one 512-thread
// workgroup per CU (its LDS), so registers are not what bounds occupancy.
template <int MB, int PA, int D, class P>
__device__ __forceinline__ void slab_block_adj(const Cf* shx, Cf* shl, const P& pol, int b, int n, const Tile& t,
                                               const View& v, double& g0, double& g1) {
  constexpr int Q = Blk<MB>::Q;
  float A0[MB][Blk<MB>::NS], A1[MB][Blk<MB>::NS], AH[MB][Blk<MB>::NS];
  constexpr int pa = PA;
  sfamd::mf::load_op<MB>([&](int i, int k) { return pol.dentry0(b, n, i, k); }, A0, pa);
  if (P::kTwoParams) sfamd::mf::load_op<MB>([&](int i, int k) { return pol.dentry1(b, n, i, k); }, A1, pa);
  sfamd::mf::load_op<MB>([&](int i, int k) { return pol.entry_h(b, n, i, k); }, AH, pa);
  const int per = pol.per_block(t, v);
  const int lane = threadIdx.x & 63;
  const int cl = lane & 15, gq = (lane >> 4) * pa;
  float s0 = 0.f, s1 = 0.f;
  for (int fb = 0; fb * 16 < per; ++fb) {
    const int phi = fb * 16 + cl;
    const bool valid = phi < per;
    int pos0, step;
    pol.fiber(b, valid ? phi : 0, t, v, pos0, step);
    Cf x[Q], l[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const bool in_range = valid && q < pa && gq + q < n;
      const int o = pos0 + sfamd::mf::clamp_amp(gq + q, n) * step;
      const Cf lv = shl[o], xv = shx[o];
      l[q] = in_range ? lv : mk<float>(0.f, 0.f);
      x[q] = in_range ? xv : mk<float>(0.f, 0.f);
    }
    pos0 += gq * step;
    f4 acc[MB];
    sfamd::mf::apply_op<MB>(A0, x, acc, pa);
#pragma unroll
    for (int q = 0; q < Q; ++q) s0 += re_conj_mul(l[q], sfamd::mf::acc_at<MB>(acc, q));
    if (P::kTwoParams) {
      sfamd::mf::apply_op<MB>(A1, x, acc, pa);
#pragma unroll
      for (int q = 0; q < Q; ++q) s1 += re_conj_mul(l[q], sfamd::mf::acc_at<MB>(acc, q));
    }
    sfamd::mf::apply_op<MB>(AH, l, acc, pa);
    if (valid) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (q < pa && gq + q < n) shl[pos0 + q * step] = sfamd::mf::acc_at<MB>(acc, q);
    }
  }
  g0 += (double)s0;
  g1 += (double)s1;
}

template <int D, class P>
__global__ void __launch_bounds__(512)
k_slab_mfma_adj(const Cf* __restrict__ psi, const Cf* __restrict__ lo_, Cf* __restrict__ li, Tile t, P gpol,
                double* partial, int nwg, int zp1) {
  extern __shared__ __align__(16) unsigned char smem[];
  Cf* shx = (Cf*)smem;
  Cf* shl = shx + t.cap;
  const P pol = gpol;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double g0 = 0.0, g1 = 0.0;
  for (int64_t tile = blockIdx.x; tile < t.ntile; tile += gridDim.x) {
    const View v = tile_view(t, tile, D);
    tile_load(shx, psi, t, v, D);
    tile_load(shl, lo_, t, v, D);
    __syncthreads();
    for (int b = wave; b < P::NB; b += nw) {
      const int n = pol.size(b);
      const int mb = (2 * n + 15) / 16;
#define SFAMD_SLAB_BLOCK_ADJ(MB)                                        \
  if (sfamd::mf::group_amps(n) == 2 * MB)                               \
    slab_block_adj<MB, 2 * MB, D>(shx, shl, pol, b, n, t, v, g0, g1);   \
  else                                                                  \
    slab_block_adj<MB, 2 * MB - 1, D>(shx, shl, pol, b, n, t, v, g0, g1);
      if (mb == 1) {
        SFAMD_SLAB_BLOCK_ADJ(1)
      } else if constexpr (D > 8) {
        if (mb == 2) {
          SFAMD_SLAB_BLOCK_ADJ(2)
        } else if constexpr (D > 16) {
          SFAMD_SLAB_BLOCK_ADJ(3)
        }
      }
#undef SFAMD_SLAB_BLOCK_ADJ
    }
    __syncthreads();
    tile_store(shl, li, t, v, D);
    __syncthreads();
  }
  write_partials(g0, g1, partial, nwg, zp1);
}

// ------------------------------------------------------------------ host side
#define SFAMD_SLAB_D(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)

// tile for gate axes of strides S_out > Sin within limit amplitudes per array
bool plan_tile(Tile& t, int64_t N, int D, int64_t S_out, int64_t Sin, int64_t limit) {
  t.slab = (int64_t)D * S_out;
  t.S_out = S_out;
  t.Sin = (int)Sin;
  t.RW = (int)(D * Sin);
  const int64_t mid = S_out / (D * Sin);
  if (mid < 1 || t.slab <= 0 || N % t.slab != 0) return false;
  t.Mid = (int)mid;
  t.nslab = N / t.slab;
  if (t.slab <= limit) {
    t.G = (int)(limit / t.slab);
    if (t.G > t.nslab) t.G = (int)t.nslab;
    t.Mt = t.Mid;
    t.nmc = 1;
  } else {
    t.G = 1;
    const int64_t mt = limit / ((int64_t)D * t.RW);
    if (mt < 1) return false;
    t.Mt = (int)mt;
    t.nmc = (t.Mid + t.Mt - 1) / t.Mt;
  }
  t.cap = t.G * D * t.Mt * t.RW;
  t.ntile = (t.nslab + t.G - 1) / t.G * t.nmc;
  return true;
}

// matrix-core tiles: 8192 amplitudes per array (a whole slab of a DM mode pair
// at cutoff 20, so each block operator serves twice the fibers of a half slab),
// whole fiber blocks of 16 when a tile holds >= 16 slabs.  Operator entries are
// read from the gate table in global memory (L1/L2-resident); staging it in LDS
// cost more occupancy than it saved.
constexpr int64_t kMfmaTile = 8192;
// Forward tiles are halved (32 KB: more workgroups per CU to overlap one
// tile's copies with another's products) when the half still holds 16 whole
// slabs, i.e. a full 16-fiber chain per block: the ket beamsplitter on the
// innermost pair at cutoff 16 went from 38% to 42% of the HBM peak, while
// halving the cutoff-20 DM tiles (which splits their slabs) was slower (DM
// loss on the innermost mode 38% -> 28%).
inline int64_t mfma_fwd_tile(int D, int64_t S_out) {
  return 16 * (int64_t)D * S_out <= kMfmaTile / 2 ? kMfmaTile / 2 : kMfmaTile;
}
bool plan_tile_mfma(Tile& t, int64_t N, int D, int64_t S_out, int64_t Sin, int64_t limit) {
  if (!plan_tile(t, N, D, S_out, Sin, limit)) return false;
  if (t.G >= 16) {
    t.G -= t.G % 16;
    t.cap = t.G * D * t.Mt * t.RW;
    t.ntile = (t.nslab + t.G - 1) / t.G * t.nmc;
  }
  return true;
}

template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return 0;
  return (int)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

int slab_grid(const Tile& t) { return (int)(t.ntile < kMaxGrid ? t.ntile : kMaxGrid); }

}  // namespace

namespace sfamd {

int64_t slab_limit(int dtype, int adjoint) { return adjoint ? 4096 : 8192; }

bool slab_supported(int64_t N, int D, int64_t S_out, int64_t Sin, int dtype, int adjoint) {
  if (D < 2 || D > kSlabMaxD) return false;
  Tile t;
  if (!plan_tile(t, N, D, S_out, Sin, slab_limit(dtype, adjoint))) return false;
  const int64_t esz = dtype == SFAMD_C64 ? 8 : 16;
  return (adjoint ? 2 : 1) * (int64_t)t.cap * esz <= kSlabLds;
}

#define CP(T, p) ((const C<T>*)(p))
#define MP(T, p) ((C<T>*)(p))

// LDS bytes of a launch: tile arrays plus the staged coefficients when they fit
static size_t slab_lds(const Tile& t, int arrays, int64_t esz, int64_t coef_bytes, bool* ml) {
  const int64_t base = arrays * (int64_t)t.cap * esz;
  *ml = base + coef_bytes <= kSlabLds;
  return (size_t)(base + (*ml ? coef_bytes : 0));
}

template <typename T, int D>
static int bs_slab_launch(const void* in, void* out, const Tile& t, int p_outer, const void* blocks, int conj,
                          hipStream_t s) {
  bool ml;
  const size_t lds = slab_lds(t, 1, sizeof(C<T>), bs_nb(D) * (int64_t)sizeof(C<T>), &ml);
  if (ml) {
    RET_IF(allow_lds(k_bs_slab<T, D, true>, lds));
    hipLaunchKernelGGL((k_bs_slab<T, D, true>), dim3(slab_grid(t)), dim3(kSlabBlock), lds, s, CP(T, in), MP(T, out),
                       t, p_outer, CP(T, blocks), conj);
  } else {
    RET_IF(allow_lds(k_bs_slab<T, D, false>, lds));
    hipLaunchKernelGGL((k_bs_slab<T, D, false>), dim3(slab_grid(t)), dim3(kSlabBlock), lds, s, CP(T, in),
                       MP(T, out), t, p_outer, CP(T, blocks), conj);
  }
  return (int)hipGetLastError();
}

template <typename T, int D>
static int bs_slab_adj_launch(const void* psi, const void* lo, void* li, const Tile& t, int p_outer,
                              const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial,
                              int nwg, int zp1, hipStream_t s) {
  bool ml;
  const size_t lds = slab_lds(t, 2, sizeof(C<T>), 2 * bs_nb(D) * (int64_t)sizeof(C<T>), &ml);
  if (ml) {
    RET_IF(allow_lds(k_bs_slab_adj<T, D, true>, lds));
    hipLaunchKernelGGL((k_bs_slab_adj<T, D, true>), dim3(nwg), dim3(kSlabBlock), lds, s, CP(T, psi), CP(T, lo),
                       MP(T, li), t, p_outer, CP(T, blocks), CP(T, dblocks), conj, wa, wb, partial, nwg,
                       zp1);
  } else {
    RET_IF(allow_lds(k_bs_slab_adj<T, D, false>, lds));
    hipLaunchKernelGGL((k_bs_slab_adj<T, D, false>), dim3(nwg), dim3(kSlabBlock), lds, s, CP(T, psi), CP(T, lo),
                       MP(T, li), t, p_outer, CP(T, blocks), CP(T, dblocks), conj, wa, wb, partial, nwg,
                       zp1);
  }
  return (int)hipGetLastError();
}

template <typename T, int D>
static int loss_slab_launch(const void* in, void* out, const Tile& t, const void* coef, hipStream_t s) {
  bool ml;
  const size_t lds = slab_lds(t, 1, sizeof(C<T>), (int64_t)D * D * D * sizeof(T), &ml);
  if (ml) {
    RET_IF(allow_lds(k_loss_slab<T, D, true>, lds));
    hipLaunchKernelGGL((k_loss_slab<T, D, true>), dim3(slab_grid(t)), dim3(kSlabBlock), lds, s, CP(T, in),
                       MP(T, out), t, (const T*)coef);
  } else {
    RET_IF(allow_lds(k_loss_slab<T, D, false>, lds));
    hipLaunchKernelGGL((k_loss_slab<T, D, false>), dim3(slab_grid(t)), dim3(kSlabBlock), lds, s, CP(T, in),
                       MP(T, out), t, (const T*)coef);
  }
  return (int)hipGetLastError();
}

template <typename T, int D>
static int loss_slab_adj_launch(const void* psi, const void* lo, void* li, const Tile& t, const void* coef,
                                const void* dcoef, double* partial, int nwg, int zp1, hipStream_t s) {
  bool ml;
  const size_t lds = slab_lds(t, 2, sizeof(C<T>), 2 * (int64_t)D * D * D * sizeof(T), &ml);
  if (ml) {
    RET_IF(allow_lds(k_loss_slab_adj<T, D, true>, lds));
    hipLaunchKernelGGL((k_loss_slab_adj<T, D, true>), dim3(nwg), dim3(kSlabBlock), lds, s, CP(T, psi), CP(T, lo),
                       MP(T, li), t, (const T*)coef, (const T*)dcoef, partial, nwg, zp1);
  } else {
    RET_IF(allow_lds(k_loss_slab_adj<T, D, false>, lds));
    hipLaunchKernelGGL((k_loss_slab_adj<T, D, false>), dim3(nwg), dim3(kSlabBlock), lds, s, CP(T, psi), CP(T, lo),
                       MP(T, li), t, (const T*)coef, (const T*)dcoef, partial, nwg, zp1);
  }
  return (int)hipGetLastError();
}

// ---- complex64: matrix-core slab kernels
template <int D, class P>
static int slab_mfma_launch(const void* in, void* out, const Tile& t, const P& pol, hipStream_t s) {
  const size_t lds = (size_t)t.cap * sizeof(Cf);
  RET_IF(allow_lds(k_slab_mfma<D, P>, lds));
  hipLaunchKernelGGL((k_slab_mfma<D, P>), dim3(slab_grid(t)), dim3(256), lds, s, (const Cf*)in, (Cf*)out, t, pol);
  return (int)hipGetLastError();
}

template <int D, class P>
static int slab_mfma_adj_launch(const void* psi, const void* lo, void* li, const Tile& t, const P& pol,
                                double* partial, int nwg, int zp1, hipStream_t s) {
  const size_t lds = 2 * (size_t)t.cap * sizeof(Cf);
  RET_IF(allow_lds(k_slab_mfma_adj<D, P>, lds));
  hipLaunchKernelGGL((k_slab_mfma_adj<D, P>), dim3(nwg), dim3(512), lds, s, (const Cf*)psi, (const Cf*)lo, (Cf*)li,
                     t, pol, partial, nwg, zp1);
  return (int)hipGetLastError();
}

static int bs_slab_mfma(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks,
                        int conj, int diff, hipStream_t s) {
  Tile t;
  const int p_outer = sp > sq;
  if (!plan_tile_mfma(t, N, D, p_outer ? sp : sq, p_outer ? sq : sp, mfma_fwd_tile(D, p_outer ? sp : sq)))
    return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                                     \
  case d: {                                                         \
    BsPol<d> pol{(const Cf*)blocks, nullptr, conj, p_outer, 0, 0, diff}; \
    return slab_mfma_launch<d>(in, out, t, pol, s);                 \
  }
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

static int bs_slab_mfma_adj(const void* psi, const void* lo, void* li, int64_t N, int D, int64_t sp, int64_t sq,
                            const void* blocks, const void* dblocks, int conj, int wa, int wb, int diff,
                            double* partial, int nwg, int zp1, hipStream_t s) {
  Tile t;
  const int p_outer = sp > sq;
  if (!plan_tile_mfma(t, N, D, p_outer ? sp : sq, p_outer ? sq : sp, kMfmaTile)) return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                                                    \
  case d: {                                                                        \
    BsPol<d> pol{(const Cf*)blocks, (const Cf*)dblocks, conj, p_outer, wa, wb, diff}; \
    return slab_mfma_adj_launch<d>(psi, lo, li, t, pol, partial, nwg, zp1, s);     \
  }
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

static int loss_slab_mfma(const void* in, void* out, int64_t N, int D, int64_t sb, int64_t sk, const void* coef,
                          hipStream_t s) {
  Tile t;
  if (!plan_tile_mfma(t, N, D, sb, sk, mfma_fwd_tile(D, sb))) return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                             \
  case d: {                                                 \
    LossPol<d> pol{(const float*)coef, nullptr};            \
    return slab_mfma_launch<d>(in, out, t, pol, s);         \
  }
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

static int loss_slab_mfma_adj(const void* psi, const void* lo, void* li, int64_t N, int D, int64_t sb, int64_t sk,
                              const void* coef, const void* dcoef, double* partial, int nwg, int zp1, hipStream_t s) {
  Tile t;
  if (!plan_tile_mfma(t, N, D, sb, sk, kMfmaTile)) return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                                                    \
  case d: {                                                                        \
    LossPol<d> pol{(const float*)coef, (const float*)dcoef};                       \
    return slab_mfma_adj_launch<d>(psi, lo, li, t, pol, partial, nwg, zp1, s);     \
  }
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

int pair_bs_slab(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
                 int dtype, hipStream_t s, int diff) {
  if (dtype == SFAMD_C64) return bs_slab_mfma(in, out, N, D, sp, sq, blocks, conj, diff, s);
  if (diff) return SFAMD_E_ARG;  // difference blocks: matrix-core variant only
  Tile t;
  const int p_outer = sp > sq;
  if (!plan_tile(t, N, D, p_outer ? sp : sq, p_outer ? sq : sp, slab_limit(dtype, 0))) return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                                                                 \
  case d:                                                                                       \
    return dtype == SFAMD_C64 ? bs_slab_launch<float, d>(in, out, t, p_outer, blocks, conj, s)  \
                              : bs_slab_launch<double, d>(in, out, t, p_outer, blocks, conj, s);
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

int pair_bs_adj_slab(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                     const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial, int nwg,
                     int zp1, int dtype, hipStream_t s, int diff) {
  if (dtype == SFAMD_C64)
    return bs_slab_mfma_adj(psi, lam_out, lam_in, N, D, sp, sq, blocks, dblocks, conj, wa, wb, diff, partial, nwg, zp1,
                            s);
  if (diff) return SFAMD_E_ARG;
  Tile t;
  const int p_outer = sp > sq;
  if (!plan_tile(t, N, D, p_outer ? sp : sq, p_outer ? sq : sp, slab_limit(dtype, 1))) return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                                                                                 \
  case d:                                                                                                       \
    return dtype == SFAMD_C64 ? bs_slab_adj_launch<float, d>(psi, lam_out, lam_in, t, p_outer, blocks, dblocks, \
                                                             conj, wa, wb, partial, nwg, zp1, s)               \
                              : bs_slab_adj_launch<double, d>(psi, lam_out, lam_in, t, p_outer, blocks,        \
                                                              dblocks, conj, wa, wb, partial, nwg, zp1, s);
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

int loss_slab(const void* in, void* out, int64_t N, int D, int64_t sb, int64_t sk, const void* coef, int dtype,
              hipStream_t s) {
  if (dtype == SFAMD_C64) return loss_slab_mfma(in, out, N, D, sb, sk, coef, s);
  Tile t;
  if (!plan_tile(t, N, D, sb, sk, slab_limit(dtype, 0))) return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                                                   \
  case d:                                                                         \
    return dtype == SFAMD_C64 ? loss_slab_launch<float, d>(in, out, t, coef, s)   \
                              : loss_slab_launch<double, d>(in, out, t, coef, s);
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

int loss_adj_slab(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sb, int64_t sk,
                  const void* coef, const void* dcoef, double* partial, int nwg, int zp1, int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64) return loss_slab_mfma_adj(psi, lam_out, lam_in, N, D, sb, sk, coef, dcoef, partial, nwg, zp1, s);
  Tile t;
  if (!plan_tile(t, N, D, sb, sk, slab_limit(dtype, 1))) return SFAMD_E_ARG;
  switch (D) {
#define CASE(d)                                                                                                   \
  case d:                                                                                                         \
    return dtype == SFAMD_C64                                                                                     \
               ? loss_slab_adj_launch<float, d>(psi, lam_out, lam_in, t, coef, dcoef, partial, nwg, zp1, s)       \
               : loss_slab_adj_launch<double, d>(psi, lam_out, lam_in, t, coef, dcoef, partial, nwg, zp1, s);
    SFAMD_SLAB_D(CASE)
#undef CASE
    default:
      return SFAMD_E_CUTOFF;
  }
}

}  // namespace sfamd
