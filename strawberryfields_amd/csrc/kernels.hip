// Element-wise gate application, adjoint and readout kernels.
//
// A state is a tensor of R axes of extent D (R = n for a ket, 2n for a density
// matrix with interleaved axes [i0, j0, i1, j1, ...]).  Each thread owns one
// output element e; the element's digit on a gate axis of stride B is
// (e / B) % D, and the D elements it mixes with sit at base + k*B.  Adjacent
// threads therefore read adjacent addresses for every k, which keeps every
// load coalesced whatever the axis; the D-fold reuse of an input is served by
// L1/L2.  The cutoff D is a runtime argument, so any cutoff works here.
//
// Reference semantics:
//   single-mode gate  fockbackend/circuit.py:118-217 (apply_gate_BLAS)
//   beamsplitter      fockbackend/circuit.py:219-335 (selection rule n1+n2 conserved)
//   loss channel      fockbackend/circuit.py:65-87 + ops.py:471-490
//   pure -> mixed     fockbackend/ops.py:110-120
//   readout           backends/states.py:556-663,947-980
// Adjoint convention (PyTorch / TF): for a state cotangent g = dL/dRe + i dL/dIm,
// a linear map y = G x pulls back as g_x = G^H g_y and a real parameter p of G
// receives Re sum conj(g_y) (dG/dp x)  (tfbackend/ops.py:197-202).
#include "common.h"

namespace {

template <typename T>
using C = cpx<T>;

// ------------------------------------------------------------------ forward
template <typename T>
__global__ void __launch_bounds__(256)
k_axis_dense(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, int64_t B,
             const C<T>* __restrict__ M, int conj) {
  GRID_STRIDE(e, N) {
    const int i = (int)((e / B) % D);
    const int64_t base = e - (int64_t)i * B;
    const C<T>* row = M + (int64_t)i * D;
    C<T> acc = mk<T>(0, 0);
    if (conj) {
      for (int k = 0; k < D; ++k) cmac_conj(acc, row[k], in[base + (int64_t)k * B]);
    } else {
      for (int k = 0; k < D; ++k) cmac(acc, row[k], in[base + (int64_t)k * B]);
    }
    out[e] = acc;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_axis_diag(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, int64_t B,
            const C<T>* __restrict__ d, int conj) {
  GRID_STRIDE(e, N) {
    const int i = (int)((e / B) % D);
    C<T> f = d[i];
    if (conj) f.im = -f.im;
    out[e] = cmul(f, in[e]);
  }
}

// ---- complex64 diagonal gates, 4 amplitudes per lane and step through
// 16-byte nontemporal accesses (states are streamed once per gate).  The
// axis positions i (stride Bi) and, for the two-sided DM form, j (stride Bj)
// are found once per step and advanced by one amplitude at a time.
typedef float f4v __attribute__((ext_vector_type(4)));

struct AxisPos {
  int i;
  int64_t r;  // offset within the current stride
};

__device__ __forceinline__ AxisPos axis_pos(int64_t e, int D, int64_t B) {
  const int64_t q = e / B;
  return AxisPos{(int)(q % D), e - q * B};
}

__device__ __forceinline__ void axis_next(AxisPos& p, int D, int64_t B) {
  if (++p.r == B) {
    p.r = 0;
    if (++p.i == D) p.i = 0;
  }
}

// factor of amplitude (i, j): d_i (conj: conj(d_i)) on one axis, d_i conj(d_j) on a DM mode
__device__ __forceinline__ C<float> diag_factor(const C<float>* __restrict__ d, int i, int j, int pair, int conj) {
  C<float> f = d[i];
  if (pair) return cmul(f, cconj(d[j]));
  if (conj) f.im = -f.im;
  return f;
}

__global__ void __launch_bounds__(256)
k_diag_c64(const f4v* __restrict__ in, f4v* __restrict__ out, int64_t N4, int D, int64_t Bi, int64_t Bj,
           const C<float>* __restrict__ d, int pair, int conj) {
  GRID_STRIDE(t, N4) {
    AxisPos pi = axis_pos(4 * t, D, Bi), pj = pair ? axis_pos(4 * t, D, Bj) : AxisPos{0, 0};
    const f4v a = __builtin_nontemporal_load(in + 2 * t), b = __builtin_nontemporal_load(in + 2 * t + 1);
    C<float> x[4] = {mk<float>(a.x, a.y), mk<float>(a.z, a.w), mk<float>(b.x, b.y), mk<float>(b.z, b.w)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      x[k] = cmul(diag_factor(d, pi.i, pj.i, pair, conj), x[k]);
      axis_next(pi, D, Bi);
      if (pair) axis_next(pj, D, Bj);
    }
    __builtin_nontemporal_store(f4v{x[0].re, x[0].im, x[1].re, x[1].im}, out + 2 * t);
    __builtin_nontemporal_store(f4v{x[2].re, x[2].im, x[3].re, x[3].im}, out + 2 * t + 1);
  }
}

// two-sided diagonal gate on a DM: rho'[i, j] = d_i conj(d_j) rho[i, j]
template <typename T>
__global__ void __launch_bounds__(256)
k_pair_diag(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, int64_t Bbra,
            int64_t Bket, const C<T>* __restrict__ d) {
  GRID_STRIDE(e, N) {
    const int i = (int)((e / Bbra) % D);
    const int j = (int)((e / Bket) % D);
    const C<T> f = cmul(d[i], cconj(d[j]));
    out[e] = cmul(f, in[e]);
  }
}

// beamsplitter on axes p (stride sp, first pair of the matrix) and q (stride
// sq): out[i, j] = sum_k B_s[i][k] in[k, s-k], s = i + j
template <typename T>
__global__ void __launch_bounds__(256)
k_pair_bs(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, int64_t sp, int64_t sq,
          const C<T>* __restrict__ blk, int conj) {
  GRID_STRIDE(e, N) {
    const int i = (int)((e / sp) % D);
    const int j = (int)((e / sq) % D);
    const int s = i + j;
    const int lo = bs_lo(s, D), hi = bs_hi(s, D), c = hi - lo + 1;
    const int64_t base = e - (int64_t)i * sp - (int64_t)j * sq;
    const C<T>* row = blk + bs_block_off(s, D) + (int64_t)(i - lo) * c;
    C<T> acc = mk<T>(0, 0);
    for (int k = lo; k <= hi; ++k) {
      const C<T> x = in[base + (int64_t)k * sp + (int64_t)(s - k) * sq];
      if (conj)
        cmac_conj(acc, row[k - lo], x);
      else
        cmac(acc, row[k - lo], x);
    }
    out[e] = acc;
  }
}

// two-mode squeezing on axes p (stride sp, first pair of the matrix) and q:
// out1 - out2 = in1 - in2 = d, so out[i, j] = sum_k G_d[i][k] in[k, k-d] with
// d = i - j; block d is stored as block s = D - 1 + d of the compact layout
// (first index i = out1, second k = in1, both from bs_lo(s) to bs_hi(s)).
template <typename T>
__global__ void __launch_bounds__(256)
k_pair_s2(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, int64_t sp, int64_t sq,
          const C<T>* __restrict__ blk, int conj) {
  GRID_STRIDE(e, N) {
    const int i = (int)((e / sp) % D);
    const int j = (int)((e / sq) % D);
    const int d = i - j, s = D - 1 + d;
    const int lo = bs_lo(s, D), hi = bs_hi(s, D), c = hi - lo + 1;
    const int64_t base = e - (int64_t)i * sp - (int64_t)j * sq;
    const C<T>* row = blk + bs_block_off(s, D) + (int64_t)(i - lo) * c;
    C<T> acc = mk<T>(0, 0);
    for (int k = lo; k <= hi; ++k) {
      const C<T> x = in[base + (int64_t)k * sp + (int64_t)(k - d) * sq];
      if (conj)
        cmac_conj(acc, row[k - lo], x);
      else
        cmac(acc, row[k - lo], x);
    }
    out[e] = acc;
  }
}

// loss on the (bra, ket) axes of one mode: out[p, q] = sum_l c[p][q][l] in[p+l, q+l]
template <typename T>
__global__ void __launch_bounds__(256)
k_loss(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, int64_t sp, int64_t sq,
       const T* __restrict__ coef) {
  GRID_STRIDE(e, N) {
    const int p = (int)((e / sp) % D);
    const int q = (int)((e / sq) % D);
    const int64_t base = e - (int64_t)p * sp - (int64_t)q * sq;
    const int lmax = D - (p > q ? p : q);
    const T* c = coef + ((int64_t)p * D + q) * D;
    C<T> acc = mk<T>(0, 0);
    for (int l = 0; l < lmax; ++l) {
      const C<T> x = in[base + (int64_t)(p + l) * sp + (int64_t)(q + l) * sq];
      acc.re += c[l] * x.re;
      acc.im += c[l] * x.im;
    }
    out[e] = acc;
  }
}

// rho[i0, j0, i1, j1, ...] = psi[i0, i1, ...] conj(psi[j0, j1, ...])
template <typename T>
__global__ void __launch_bounds__(256)
k_mix(const C<T>* __restrict__ psi, C<T>* __restrict__ rho, int n, int D, int64_t N) {
  GRID_STRIDE(e, N) {
    int64_t r = e, I = 0, J = 0, w = 1;
    for (int m = n - 1; m >= 0; --m) {
      const int j = (int)(r % D);
      r /= D;
      const int i = (int)(r % D);
      r /= D;
      I += i * w;
      J += j * w;
      w *= D;
    }
    rho[e] = cmul(psi[I], cconj(psi[J]));
  }
}

// rho'[.., i, j, ..] = sigma[i][j] sum_t rho[.., t, t, ..]
template <typename T>
__global__ void __launch_bounds__(256)
k_replace(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, int64_t sp, int64_t sq,
          const C<T>* __restrict__ sigma) {
  GRID_STRIDE(e, N) {
    const int i = (int)((e / sp) % D);
    const int j = (int)((e / sq) % D);
    const int64_t base = e - (int64_t)i * sp - (int64_t)j * sq;
    C<T> tr = mk<T>(0, 0);
    for (int t = 0; t < D; ++t) tr = cadd(tr, in[base + (int64_t)t * (sp + sq)]);
    out[e] = cmul(sigma[(int64_t)i * D + j], tr);
  }
}

__global__ void __launch_bounds__(256) k_fill_zero(uint32_t* p, int64_t words) {
  GRID_STRIDE(e, words) p[e] = 0u;
}

// ------------------------------------------------------------------ adjoint
// Each adjoint launch runs exactly nwg workgroups and writes its workgroup
// partial sums of the two parameter slots to partial[slot * nwg + blockIdx.x].
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

template <typename T>
__global__ void __launch_bounds__(256)
k_axis_dense_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li,
                 int64_t N, int D, int64_t B, const C<T>* __restrict__ M, const C<T>* __restrict__ dM0,
                 const C<T>* __restrict__ dM1, int conj, double* partial, int nwg, int zp1) {
  double g0 = 0.0, g1 = 0.0;
  GRID_STRIDE(e, N) {
    const int i = (int)((e / B) % D);
    const int64_t base = e - (int64_t)i * B;
    // lam_in[e] = sum_r conj(M'[r][i]) lam_out[r]   (M' = conj(M) if conj)
    C<T> acc = mk<T>(0, 0);
    for (int r = 0; r < D; ++r) {
      const C<T> m = M[(int64_t)r * D + i];
      const C<T> l = lo_[base + (int64_t)r * B];
      if (conj)
        cmac(acc, m, l);
      else
        cmac_conj(acc, m, l);
    }
    li[e] = acc;
    // grad: conj(lam_out[e]) * sum_k dM'[i][k] psi[k]
    C<T> u0 = mk<T>(0, 0), u1 = mk<T>(0, 0);
    const C<T>* r0 = dM0 + (int64_t)i * D;
    const C<T>* r1 = dM1 + (int64_t)i * D;
    for (int k = 0; k < D; ++k) {
      const C<T> x = psi[base + (int64_t)k * B];
      if (conj) {
        cmac_conj(u0, r0[k], x);
        cmac_conj(u1, r1[k], x);
      } else {
        cmac(u0, r0[k], x);
        cmac(u1, r1[k], x);
      }
    }
    const C<T> l = lo_[e];
    g0 += (double)re_conj_mul(l, u0);
    g1 += (double)re_conj_mul(l, u1);
  }
  write_partials(g0, g1, partial, nwg, zp1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_axis_diag_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t N,
                int D, int64_t B, const C<T>* __restrict__ d, int kerr, int conj, double* partial, int nwg, int zp1) {
  double g0 = 0.0;
  GRID_STRIDE(e, N) {
    const int i = (int)((e / B) % D);
    C<T> f = d[i];
    if (conj) f.im = -f.im;
    const C<T> l = lo_[e];
    li[e] = cmul(cconj(f), l);
    // d/dp f = i * sigma * w_i * f
    const T w = (T)(kerr ? i * i : i) * (conj ? (T)-1 : (T)1);
    const C<T> fx = cmul(f, psi[e]);
    const C<T> u = mk<T>(-w * fx.im, w * fx.re);
    g0 += (double)re_conj_mul(l, u);
  }
  write_partials(g0, 0.0, partial, nwg, zp1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_pair_diag_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t N,
                int D, int64_t Bbra, int64_t Bket, const C<T>* __restrict__ d, int kerr, double* partial,
                int nwg, int zp1) {
  double g0 = 0.0;
  GRID_STRIDE(e, N) {
    const int i = (int)((e / Bbra) % D);
    const int j = (int)((e / Bket) % D);
    const C<T> f = cmul(d[i], cconj(d[j]));
    const C<T> l = lo_[e];
    li[e] = cmul(cconj(f), l);
    const T w = kerr ? (T)(i * i - j * j) : (T)(i - j);
    const C<T> fx = cmul(f, psi[e]);
    const C<T> u = mk<T>(-w * fx.im, w * fx.re);
    g0 += (double)re_conj_mul(l, u);
  }
  write_partials(g0, 0.0, partial, nwg, zp1);
}

// adjoint: lam_in = conj(f) lam_out;  g += Re conj(lam_out) (i w f psi) with
// w = sigma i (R), sigma i^2 (K) on one axis, i - j or i^2 - j^2 on a DM mode
__global__ void __launch_bounds__(256)
k_diag_adj_c64(const f4v* __restrict__ psi, const f4v* __restrict__ lo_, f4v* __restrict__ li, int64_t N4, int D,
               int64_t Bi, int64_t Bj, const C<float>* __restrict__ d, int pair, int kerr, int conj, double* partial,
               int nwg, int zp1) {
  double g0 = 0.0;
  GRID_STRIDE(t, N4) {
    AxisPos pi = axis_pos(4 * t, D, Bi), pj = pair ? axis_pos(4 * t, D, Bj) : AxisPos{0, 0};
    const f4v xa = __builtin_nontemporal_load(psi + 2 * t), xb = __builtin_nontemporal_load(psi + 2 * t + 1);
    const f4v la = __builtin_nontemporal_load(lo_ + 2 * t), lb = __builtin_nontemporal_load(lo_ + 2 * t + 1);
    const C<float> x[4] = {mk<float>(xa.x, xa.y), mk<float>(xa.z, xa.w), mk<float>(xb.x, xb.y), mk<float>(xb.z, xb.w)};
    C<float> l[4] = {mk<float>(la.x, la.y), mk<float>(la.z, la.w), mk<float>(lb.x, lb.y), mk<float>(lb.z, lb.w)};
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const C<float> f = diag_factor(d, pi.i, pj.i, pair, conj);
      float w;
      if (pair)
        w = kerr ? (float)(pi.i * pi.i - pj.i * pj.i) : (float)(pi.i - pj.i);
      else
        w = (float)(kerr ? pi.i * pi.i : pi.i) * (conj ? -1.f : 1.f);
      const C<float> fx = cmul(f, x[k]);
      s += re_conj_mul(l[k], mk<float>(-w * fx.im, w * fx.re));
      l[k] = cmul(cconj(f), l[k]);
      axis_next(pi, D, Bi);
      if (pair) axis_next(pj, D, Bj);
    }
    g0 += (double)s;
    __builtin_nontemporal_store(f4v{l[0].re, l[0].im, l[1].re, l[1].im}, li + 2 * t);
    __builtin_nontemporal_store(f4v{l[2].re, l[2].im, l[3].re, l[3].im}, li + 2 * t + 1);
  }
  write_partials(g0, 0.0, partial, nwg, zp1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_pair_bs_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t N,
              int D, int64_t sp, int64_t sq, const C<T>* __restrict__ blk, const C<T>* __restrict__ dblk,
              int conj, int wa, int wb, double* partial, int nwg, int zp1) {
  double g0 = 0.0, g1 = 0.0;
  GRID_STRIDE(e, N) {
    const int i = (int)((e / sp) % D);
    const int j = (int)((e / sq) % D);
    const int s = i + j;
    const int lo = bs_lo(s, D), hi = bs_hi(s, D), c = hi - lo + 1;
    const int64_t base = e - (int64_t)i * sp - (int64_t)j * sq;
    const int64_t off = bs_block_off(s, D);
    // lam_in[(i, s-i)] = sum_N conj(B'[N][i]) lam_out[(N, s-N)]
    C<T> acc = mk<T>(0, 0);
    for (int Nn = lo; Nn <= hi; ++Nn) {
      const C<T> b = blk[off + (int64_t)(Nn - lo) * c + (i - lo)];
      const C<T> l = lo_[base + (int64_t)Nn * sp + (int64_t)(s - Nn) * sq];
      if (conj)
        cmac(acc, b, l);
      else
        cmac_conj(acc, b, l);
    }
    li[e] = acc;
    // grads: d/dp0 from the derivative blocks, d/dp1 B'[i][k] = i sigma (wa k + wb i) B'[i][k]
    C<T> ut = mk<T>(0, 0), up = mk<T>(0, 0);
    const C<T>* row = blk + off + (int64_t)(i - lo) * c;
    const C<T>* drow = dblk + off + (int64_t)(i - lo) * c;
    for (int k = lo; k <= hi; ++k) {
      const C<T> x = psi[base + (int64_t)k * sp + (int64_t)(s - k) * sq];
      C<T> b = row[k - lo], db = drow[k - lo];
      if (conj) {
        b.im = -b.im;
        db.im = -db.im;
      }
      cmac(ut, db, x);
      const T w = (T)(wa * k + wb * i) * (conj ? (T)-1 : (T)1);
      const C<T> bx = cmul(b, x);
      up.re += -w * bx.im;
      up.im += w * bx.re;
    }
    const C<T> l = lo_[e];
    g0 += (double)re_conj_mul(l, ut);
    g1 += (double)re_conj_mul(l, up);
  }
  write_partials(g0, g1, partial, nwg, zp1);
}

// adjoint of k_pair_s2: lam_in[(i, i-d)] = sum_N conj(G'[N][i]) lam_out[(N, N-d)];
// gradients: d/dr from the derivative blocks, d/dphi G'[N][k] = i sigma (N - k) G'[N][k]
template <typename T>
__global__ void __launch_bounds__(256)
k_pair_s2_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t N, int D,
              int64_t sp, int64_t sq, const C<T>* __restrict__ blk, const C<T>* __restrict__ dblk, int conj,
              double* partial, int nwg, int zp1) {
  double g0 = 0.0, g1 = 0.0;
  GRID_STRIDE(e, N) {
    const int i = (int)((e / sp) % D);
    const int j = (int)((e / sq) % D);
    const int d = i - j, s = D - 1 + d;
    const int lo = bs_lo(s, D), hi = bs_hi(s, D), c = hi - lo + 1;
    const int64_t base = e - (int64_t)i * sp - (int64_t)j * sq;
    const int64_t off = bs_block_off(s, D);
    C<T> acc = mk<T>(0, 0);
    for (int Nn = lo; Nn <= hi; ++Nn) {
      const C<T> b = blk[off + (int64_t)(Nn - lo) * c + (i - lo)];
      const C<T> l = lo_[base + (int64_t)Nn * sp + (int64_t)(Nn - d) * sq];
      if (conj)
        cmac(acc, b, l);
      else
        cmac_conj(acc, b, l);
    }
    li[e] = acc;
    C<T> ut = mk<T>(0, 0), up = mk<T>(0, 0);
    const C<T>* row = blk + off + (int64_t)(i - lo) * c;
    const C<T>* drow = dblk + off + (int64_t)(i - lo) * c;
    for (int k = lo; k <= hi; ++k) {
      const C<T> x = psi[base + (int64_t)k * sp + (int64_t)(k - d) * sq];
      C<T> b = row[k - lo], db = drow[k - lo];
      if (conj) {
        b.im = -b.im;
        db.im = -db.im;
      }
      cmac(ut, db, x);
      const T w = (T)(i - k) * (conj ? (T)-1 : (T)1);
      const C<T> bx = cmul(b, x);
      up.re += -w * bx.im;
      up.im += w * bx.re;
    }
    const C<T> l = lo_[e];
    g0 += (double)re_conj_mul(l, ut);
    g1 += (double)re_conj_mul(l, up);
  }
  write_partials(g0, g1, partial, nwg, zp1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_loss_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ lo_, C<T>* __restrict__ li, int64_t N, int D,
           int64_t sp, int64_t sq, const T* __restrict__ coef, const T* __restrict__ dcoef, double* partial,
           int nwg, int zp1) {
  double g0 = 0.0;
  GRID_STRIDE(e, N) {
    const int p = (int)((e / sp) % D);
    const int q = (int)((e / sq) % D);
    const int64_t base = e - (int64_t)p * sp - (int64_t)q * sq;
    // lam_in[p, q] = sum_l c[p-l][q-l][l] lam_out[p-l, q-l]
    const int lm = p < q ? p : q;
    C<T> acc = mk<T>(0, 0);
    for (int l = 0; l <= lm; ++l) {
      const T c = coef[((int64_t)(p - l) * D + (q - l)) * D + l];
      const C<T> x = lo_[base + (int64_t)(p - l) * sp + (int64_t)(q - l) * sq];
      acc.re += c * x.re;
      acc.im += c * x.im;
    }
    li[e] = acc;
    const int lmax = D - (p > q ? p : q);
    const T* dc = dcoef + ((int64_t)p * D + q) * D;
    C<T> u = mk<T>(0, 0);
    for (int l = 0; l < lmax; ++l) {
      const C<T> x = psi[base + (int64_t)(p + l) * sp + (int64_t)(q + l) * sq];
      u.re += dc[l] * x.re;
      u.im += dc[l] * x.im;
    }
    g0 += (double)re_conj_mul(lo_[e], u);
  }
  write_partials(g0, 0.0, partial, nwg, zp1);
}

// g_psi[I] = sum_J g[I, J] psi[J] + conj(g[J, I]) psi[J]  (rho = psi psi^dagger)
template <typename T>
__global__ void __launch_bounds__(256)
k_mix_adj(const C<T>* __restrict__ psi, const C<T>* __restrict__ g, C<T>* __restrict__ gk, int n, int D,
          int64_t NK) {
  GRID_STRIDE(I, NK) {
    int64_t eIb = 0, eIk = 0;
    {
      int64_t r = I, sk = 1;
      for (int m = n - 1; m >= 0; --m) {
        const int dgt = (int)(r % D);
        r /= D;
        eIk += dgt * sk;
        eIb += dgt * sk * D;
        sk *= (int64_t)D * D;
      }
    }
    C<T> acc = mk<T>(0, 0);
    // odometer over J
    int dig[32];
    for (int m = 0; m < n; ++m) dig[m] = 0;
    int64_t eJb = 0, eJk = 0;
    for (int64_t J = 0; J < NK; ++J) {
      const C<T> x = psi[J];
      cmac(acc, g[eIb + eJk], x);
      cmac_conj(acc, g[eJb + eIk], x);
      // increment (last mode fastest)
      int64_t sk = 1;
      for (int m = n - 1; m >= 0; --m) {
        if (dig[m] + 1 < D) {
          dig[m] += 1;
          eJk += sk;
          eJb += sk * D;
          break;
        }
        eJk -= (int64_t)(D - 1) * sk;
        eJb -= (int64_t)(D - 1) * sk * D;
        dig[m] = 0;
        sk *= (int64_t)D * D;
      }
    }
    gk[I] = acc;
  }
}

__global__ void __launch_bounds__(256)
k_reduce_grads(const double* __restrict__ partial, double* __restrict__ grads, sfamd::OpList ol, int nwg) {
  __shared__ double sh[8];
  const int op = ol.ops[blockIdx.x / 2], slot = blockIdx.x % 2;
  const unsigned c = ol.cols[blockIdx.x / 2];
  const int ncol = (int)(c & ~sfamd::kOpColsOnePass), npass = (c & sfamd::kOpColsOnePass) ? 1 : 2;
  double v = 0.0;
  for (int p = 0; p < npass; ++p) {
    const double* src = partial + (((int64_t)op * 2 + p) * 2 + slot) * nwg;
    for (int w = threadIdx.x; w < ncol; w += blockDim.x) v += src[w];
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) grads[2 * op + slot] = s;
}

// ------------------------------------------------------------------ readout

// DM element index of the diagonal entry (I, I)
__device__ __forceinline__ int64_t diag_index(int64_t I, int n, int D) {
  int64_t e = 0, w = 1;
  for (int m = n - 1; m >= 0; --m) {
    const int dgt = (int)(I % D);
    I /= D;
    e += dgt * w * (D + 1);
    w *= (int64_t)D * D;
  }
  return e;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_norm_partial(const C<T>* __restrict__ st, int n, int D, int mixed, int64_t NK, double* partial) {
  __shared__ double sh[8];
  double v = 0.0;
  GRID_STRIDE(I, NK) {
    if (mixed) {
      v += (double)st[diag_index(I, n, D)].re;
    } else {
      const C<T> x = st[I];
      v += (double)x.re * x.re + (double)x.im * x.im;
    }
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_sum_partials(const double* partial, int cnt, double* out) {
  __shared__ double sh[8];
  double v = 0.0;
  for (int w = threadIdx.x; w < cnt; w += blockDim.x) v += partial[w];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) out[0] = s;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_probs(const C<T>* __restrict__ st, T* __restrict__ p, int n, int D, int mixed, int64_t NK) {
  GRID_STRIDE(I, NK) {
    if (mixed) {
      p[I] = st[diag_index(I, n, D)].re;
    } else {
      const C<T> x = st[I];
      p[I] = x.re * x.re + x.im * x.im;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_probs_vjp(const C<T>* __restrict__ st, const T* __restrict__ gp, C<T>* __restrict__ g, int n, int D, int mixed,
            int64_t N) {
  GRID_STRIDE(e, N) {
    if (mixed) {
      // diagonal iff bra digit == ket digit for every mode
      int64_t r = e, I = 0, w = 1;
      bool diag = true;
      for (int m = n - 1; m >= 0; --m) {
        const int j = (int)(r % D);
        r /= D;
        const int i = (int)(r % D);
        r /= D;
        diag = diag && (i == j);
        I += i * w;
        w *= D;
      }
      g[e] = diag ? mk<T>(gp[I], (T)0) : mk<T>((T)0, (T)0);
    } else {
      const C<T> x = st[e];
      const T s = (T)2 * gp[e];
      g[e] = mk<T>(s * x.re, s * x.im);
    }
  }
}

// reduced DM: grid.x = output element (interleaved kept axes), grid.y = chunk
// of the traced index range
template <typename T>
__global__ void __launch_bounds__(256)
k_reduced_dm(const C<T>* __restrict__ st, ModeList ml, int D, int mixed, int64_t NR, int64_t chunk,
             double* partial) {
  __shared__ double sh[8];
  const int n = ml.n, k = ml.k;
  // strides of each mode: ket layout D^(n-1-m); DM bra D^(2(n-1-m)+1), ket D^(2(n-1-m))
  int64_t offA = 0, offB = 0;
  {
    int64_t o = blockIdx.x;
    for (int t = k - 1; t >= 0; --t) {
      const int b = (int)(o % D);
      o /= D;
      const int a = (int)(o % D);
      o /= D;
      const int m = ml.kept[t];
      if (mixed) {
        const int64_t sk = ipow(D, 2 * (n - 1 - m));
        offA += a * sk * D + b * sk;
      } else {
        const int64_t s = ipow(D, n - 1 - m);
        offA += a * s;
        offB += b * s;
      }
    }
  }
  double vr = 0.0, vi = 0.0;
  const int64_t r0 = (int64_t)blockIdx.y * chunk;
  const int64_t r1 = (r0 + chunk < NR) ? r0 + chunk : NR;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    int64_t rr = r, off = 0;
    for (int u = n - k - 1; u >= 0; --u) {
      const int dgt = (int)(rr % D);
      rr /= D;
      const int m = ml.traced[u];
      if (mixed) {
        const int64_t sk = ipow(D, 2 * (n - 1 - m));
        off += dgt * (sk * D + sk);
      } else {
        off += dgt * ipow(D, n - 1 - m);
      }
    }
    if (mixed) {
      const C<T> x = st[offA + off];
      vr += x.re;
      vi += x.im;
    } else {
      const C<T> a = st[offA + off], b = st[offB + off];
      // a * conj(b)
      vr += (double)a.re * b.re + (double)a.im * b.im;
      vi += (double)a.im * b.re - (double)a.re * b.im;
    }
  }
  const double sr = block_sum(vr, sh);
  const double si = block_sum(vi, sh);
  if (threadIdx.x == 0) {
    const int64_t idx = ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) * 2;
    partial[idx] = sr;
    partial[idx + 1] = si;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_reduced_dm_final(const double* __restrict__ partial, C<T>* __restrict__ out, int64_t NO, int chunks) {
  GRID_STRIDE(o, NO) {
    double vr = 0.0, vi = 0.0;
    for (int c = 0; c < chunks; ++c) {
      vr += partial[(o * chunks + c) * 2];
      vi += partial[(o * chunks + c) * 2 + 1];
    }
    out[o] = mk<T>((T)vr, (T)vi);
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_diag_expect(const C<T>* __restrict__ st, ModeList ml, int D, int mixed, int64_t NK,
              const double* __restrict__ values, double* partial) {
  __shared__ double sh[8];
  const int n = ml.n;
  double v = 0.0;
  GRID_STRIDE(I, NK) {
    double p;
    if (mixed) {
      p = (double)st[diag_index(I, n, D)].re;
    } else {
      const C<T> x = st[I];
      p = (double)x.re * x.re + (double)x.im * x.im;
    }
    double w = 1.0;
    for (int t = 0; t < ml.k; ++t) {
      const int m = ml.kept[t];
      const int dgt = (int)((I / ipow(D, n - 1 - m)) % D);
      w *= values[dgt];
    }
    v += p * w;
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// <a|b> with FP64 accumulation: per-block partial (re, im) pairs
template <typename T>
__global__ void __launch_bounds__(256)
k_overlap_partial(const C<T>* __restrict__ a, const C<T>* __restrict__ b, int64_t N, double* partial) {
  __shared__ double sh[8];
  double vr = 0.0, vi = 0.0;
  GRID_STRIDE(e, N) {
    const C<T> x = a[e], y = b[e];
    vr += (double)x.re * y.re + (double)x.im * y.im;
    vi += (double)x.re * y.im - (double)x.im * y.re;
  }
  const double sr = block_sum(vr, sh);
  const double si = block_sum(vi, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr;
    partial[2 * blockIdx.x + 1] = si;
  }
}

// out[0..1] = <a|b>; out[2] = 1 - |<a|b>|^2
__global__ void __launch_bounds__(256) k_overlap_final(const double* partial, int cnt, double* out) {
  __shared__ double sh[8];
  double vr = 0.0, vi = 0.0;
  for (int w = threadIdx.x; w < cnt; w += blockDim.x) {
    vr += partial[2 * w];
    vi += partial[2 * w + 1];
  }
  const double sr = block_sum(vr, sh);
  const double si = block_sum(vi, sh);
  if (threadIdx.x == 0) {
    out[0] = sr;
    out[1] = si;
    out[2] = 1.0 - sr * sr - si * si;
  }
}

// cotangent of L = 1 - |<t|psi>|^2 with respect to psi: g = -2 <t|psi> t
template <typename T>
__global__ void __launch_bounds__(256)
k_fidelity_grad(const C<T>* __restrict__ t, const double* __restrict__ ov, C<T>* __restrict__ g, int64_t N) {
  const T cr = (T)(-2.0 * ov[0]), ci = (T)(-2.0 * ov[1]);
  GRID_STRIDE(e, N) {
    const C<T> x = t[e];
    g[e] = mk<T>(cr * x.re - ci * x.im, cr * x.im + ci * x.re);
  }
}

// L = sum_I p(I) * (sum_m I_m) (total mean photon number of a ket) and its
// cotangent g = 2 (sum_m I_m) psi; per-block FP64 partials
template <typename T>
__global__ void __launch_bounds__(256)
k_number_loss(const C<T>* __restrict__ psi, C<T>* __restrict__ g, int n, int D, int64_t N, double* partial) {
  __shared__ double sh[8];
  double v = 0.0;
  GRID_STRIDE(e, N) {
    int64_t r = e;
    int tot = 0;
    for (int m = 0; m < n; ++m) {
      tot += (int)(r % D);
      r /= D;
    }
    const C<T> x = psi[e];
    v += (double)tot * ((double)x.re * x.re + (double)x.im * x.im);
    if (g) g[e] = mk<T>((T)(2 * tot) * x.re, (T)(2 * tot) * x.im);
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ------------------------------------------------------------------ helpers
#define CP(T, p) ((const C<T>*)(p))
#define MP(T, p) ((C<T>*)(p))

int64_t reduce_grid(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  return g;
}

int64_t rdm_chunks(int64_t NR) {
  int64_t c = (NR + 4095) / 4096;
  if (c < 1) c = 1;
  if (c > 256) c = 256;
  return c;
}

}  // namespace

namespace sfamd {

int axis_dense(const void* in, void* out, int64_t N, int D, int64_t B, const void* M, int conj, int dtype,
               hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_axis_dense<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out), N,
                       D, B, CP(float, M), conj);
  else
    hipLaunchKernelGGL(k_axis_dense<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in), MP(double, out),
                       N, D, B, CP(double, M), conj);
  return (int)hipGetLastError();
}

// grid for the 4-amplitude complex64 diagonal kernels
static bool diag4_ok(int64_t N, const void* a, const void* b, const void* c) {
  return N % 4 == 0 && (((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) & 15) == 0;
}

static int diag4_grid(int64_t N4) {
  int64_t g = (N4 + kBlock - 1) / kBlock;
  if (g > 4 * kMaxGrid) g = 4 * kMaxGrid;
  return (int)(g < 1 ? 1 : g);
}

int axis_diag(const void* in, void* out, int64_t N, int D, int64_t B, const void* d, int conj, int dtype,
              hipStream_t s) {
  if (dtype == SFAMD_C64 && diag4_ok(N, in, out, out)) {
    hipLaunchKernelGGL(k_diag_c64, dim3(diag4_grid(N / 4)), dim3(kBlock), 0, s, (const f4v*)in, (f4v*)out, N / 4, D,
                       B, B, CP(float, d), 0, conj);
    return (int)hipGetLastError();
  }
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_axis_diag<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out), N,
                       D, B, CP(float, d), conj);
  else
    hipLaunchKernelGGL(k_axis_diag<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in), MP(double, out),
                       N, D, B, CP(double, d), conj);
  return (int)hipGetLastError();
}

int pair_diag(const void* in, void* out, int64_t N, int D, int64_t Bbra, int64_t Bket, const void* d, int dtype,
              hipStream_t s) {
  if (dtype == SFAMD_C64 && diag4_ok(N, in, out, out)) {
    hipLaunchKernelGGL(k_diag_c64, dim3(diag4_grid(N / 4)), dim3(kBlock), 0, s, (const f4v*)in, (f4v*)out, N / 4, D,
                       Bbra, Bket, CP(float, d), 1, 0);
    return (int)hipGetLastError();
  }
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_pair_diag<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out), N,
                       D, Bbra, Bket, CP(float, d));
  else
    hipLaunchKernelGGL(k_pair_diag<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in), MP(double, out),
                       N, D, Bbra, Bket, CP(double, d));
  return (int)hipGetLastError();
}

int pair_bs(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
            int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_pair_bs<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out), N, D,
                       sp, sq, CP(float, blocks), conj);
  else
    hipLaunchKernelGGL(k_pair_bs<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in), MP(double, out), N,
                       D, sp, sq, CP(double, blocks), conj);
  return (int)hipGetLastError();
}

int pair_s2(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* blocks, int conj,
            int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_pair_s2<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out), N, D,
                       sp, sq, CP(float, blocks), conj);
  else
    hipLaunchKernelGGL(k_pair_s2<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in), MP(double, out), N,
                       D, sp, sq, CP(double, blocks), conj);
  return (int)hipGetLastError();
}

int pair_s2_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                const void* blocks, const void* dblocks, int conj, double* partial, int nwg, int zp1, int dtype,
                hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_pair_s2_adj<float>, dim3(nwg), dim3(kBlock), 0, s, CP(float, psi), CP(float, lam_out),
                       MP(float, lam_in), N, D, sp, sq, CP(float, blocks), CP(float, dblocks), conj, partial, nwg, zp1);
  else
    hipLaunchKernelGGL(k_pair_s2_adj<double>, dim3(nwg), dim3(kBlock), 0, s, CP(double, psi), CP(double, lam_out),
                       MP(double, lam_in), N, D, sp, sq, CP(double, blocks), CP(double, dblocks), conj, partial, nwg,
                       zp1);
  return (int)hipGetLastError();
}

int loss_fwd(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* coef, int dtype,
             hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_loss<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out), N, D, sp,
                       sq, (const float*)coef);
  else
    hipLaunchKernelGGL(k_loss<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in), MP(double, out), N, D,
                       sp, sq, (const double*)coef);
  return (int)hipGetLastError();
}

int mix_fwd(const void* ket, void* dm, int n, int D, int dtype, hipStream_t s) {
  const int64_t N = ipow(D, 2 * n);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_mix<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, ket), MP(float, dm), n, D, N);
  else
    hipLaunchKernelGGL(k_mix<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, ket), MP(double, dm), n, D,
                       N);
  return (int)hipGetLastError();
}

int replace_mode(const void* in, void* out, int64_t N, int D, int64_t sp, int64_t sq, const void* sigma, int dtype,
                 hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_replace<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out), N, D,
                       sp, sq, CP(float, sigma));
  else
    hipLaunchKernelGGL(k_replace<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in), MP(double, out), N,
                       D, sp, sq, CP(double, sigma));
  return (int)hipGetLastError();
}

int fill_zero(void* p, int64_t bytes, hipStream_t s) {
  const int64_t words = bytes / 4;
  if (words <= 0) return 0;
  hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(words)), dim3(kBlock), 0, s, (uint32_t*)p, words);
  return (int)hipGetLastError();
}

int axis_dense_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t B, const void* M,
                   const void* dM0, const void* dM1, int conj, double* partial, int nwg, int zp1, int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_axis_dense_adj<float>, dim3(nwg), dim3(kBlock), 0, s, CP(float, psi), CP(float, lam_out),
                       MP(float, lam_in), N, D, B, CP(float, M), CP(float, dM0), CP(float, dM1), conj, partial, nwg, zp1);
  else
    hipLaunchKernelGGL(k_axis_dense_adj<double>, dim3(nwg), dim3(kBlock), 0, s, CP(double, psi), CP(double, lam_out),
                       MP(double, lam_in), N, D, B, CP(double, M), CP(double, dM0), CP(double, dM1), conj, partial,
                       nwg, zp1);
  return (int)hipGetLastError();
}

int axis_diag_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t B, const void* d,
                  int kerr, int conj, double* partial, int nwg, int zp1, int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64 && diag4_ok(N, psi, lam_out, lam_in)) {
    hipLaunchKernelGGL(k_diag_adj_c64, dim3(nwg), dim3(kBlock), 0, s, (const f4v*)psi, (const f4v*)lam_out,
                       (f4v*)lam_in, N / 4, D, B, B, CP(float, d), 0, kerr, conj, partial, nwg, zp1);
    return (int)hipGetLastError();
  }
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_axis_diag_adj<float>, dim3(nwg), dim3(kBlock), 0, s, CP(float, psi), CP(float, lam_out),
                       MP(float, lam_in), N, D, B, CP(float, d), kerr, conj, partial, nwg, zp1);
  else
    hipLaunchKernelGGL(k_axis_diag_adj<double>, dim3(nwg), dim3(kBlock), 0, s, CP(double, psi), CP(double, lam_out),
                       MP(double, lam_in), N, D, B, CP(double, d), kerr, conj, partial, nwg, zp1);
  return (int)hipGetLastError();
}

int pair_diag_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t Bbra, int64_t Bket,
                  const void* d, int kerr, double* partial, int nwg, int zp1, int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64 && diag4_ok(N, psi, lam_out, lam_in)) {
    hipLaunchKernelGGL(k_diag_adj_c64, dim3(nwg), dim3(kBlock), 0, s, (const f4v*)psi, (const f4v*)lam_out,
                       (f4v*)lam_in, N / 4, D, Bbra, Bket, CP(float, d), 1, kerr, 0, partial, nwg, zp1);
    return (int)hipGetLastError();
  }
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_pair_diag_adj<float>, dim3(nwg), dim3(kBlock), 0, s, CP(float, psi), CP(float, lam_out),
                       MP(float, lam_in), N, D, Bbra, Bket, CP(float, d), kerr, partial, nwg, zp1);
  else
    hipLaunchKernelGGL(k_pair_diag_adj<double>, dim3(nwg), dim3(kBlock), 0, s, CP(double, psi), CP(double, lam_out),
                       MP(double, lam_in), N, D, Bbra, Bket, CP(double, d), kerr, partial, nwg, zp1);
  return (int)hipGetLastError();
}

int pair_bs_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
                const void* blocks, const void* dblocks, int conj, int wa, int wb, double* partial, int nwg, int zp1,
                int dtype,
                hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_pair_bs_adj<float>, dim3(nwg), dim3(kBlock), 0, s, CP(float, psi), CP(float, lam_out),
                       MP(float, lam_in), N, D, sp, sq, CP(float, blocks), CP(float, dblocks), conj, wa, wb, partial, nwg,
                       zp1);
  else
    hipLaunchKernelGGL(k_pair_bs_adj<double>, dim3(nwg), dim3(kBlock), 0, s, CP(double, psi), CP(double, lam_out),
                       MP(double, lam_in), N, D, sp, sq, CP(double, blocks), CP(double, dblocks), conj, wa, wb, partial,
                       nwg, zp1);
  return (int)hipGetLastError();
}

int loss_adj(const void* psi, const void* lam_out, void* lam_in, int64_t N, int D, int64_t sp, int64_t sq,
             const void* coef, const void* dcoef, double* partial, int nwg, int zp1, int dtype, hipStream_t s) {
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_loss_adj<float>, dim3(nwg), dim3(kBlock), 0, s, CP(float, psi), CP(float, lam_out),
                       MP(float, lam_in), N, D, sp, sq, (const float*)coef, (const float*)dcoef, partial, nwg, zp1);
  else
    hipLaunchKernelGGL(k_loss_adj<double>, dim3(nwg), dim3(kBlock), 0, s, CP(double, psi), CP(double, lam_out),
                       MP(double, lam_in), N, D, sp, sq, (const double*)coef, (const double*)dcoef, partial, nwg, zp1);
  return (int)hipGetLastError();
}

int mix_adj(const void* ket, const void* g_dm, void* g_ket, int n, int D, int dtype, hipStream_t s) {
  const int64_t NK = ipow(D, n);
  if (n > 32) return SFAMD_E_ARG;
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_mix_adj<float>, dim3(grid_for(NK)), dim3(kBlock), 0, s, CP(float, ket), CP(float, g_dm),
                       MP(float, g_ket), n, D, NK);
  else
    hipLaunchKernelGGL(k_mix_adj<double>, dim3(grid_for(NK)), dim3(kBlock), 0, s, CP(double, ket), CP(double, g_dm),
                       MP(double, g_ket), n, D, NK);
  return (int)hipGetLastError();
}

int reduce_grads(const double* partial, double* grads, const OpList& ol, int nwg, hipStream_t s) {
  if (ol.n <= 0) return 0;
  hipLaunchKernelGGL(k_reduce_grads, dim3(2 * ol.n), dim3(kBlock), 0, s, partial, grads, ol, nwg);
  return (int)hipGetLastError();
}

}  // namespace sfamd

// ====================================================================== C ABI
using namespace sfamd;

extern "C" int64_t sfamd_reduce_partials(int n, int D, int mixed) {
  (void)mixed;
  return reduce_grid(ipow(D, n));
}

extern "C" int sfamd_norm(const void* state, int n, int D, int mixed, double* partial, double* out_dev, int dtype,
                          void* stream) {
  CHECK_ARG(state && partial && out_dev && n >= 0 && D >= 1);
  hipStream_t s = (hipStream_t)stream;
  const int64_t NK = ipow(D, n);
  const int g = (int)reduce_grid(NK);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_norm_partial<float>, dim3(g), dim3(kBlock), 0, s, CP(float, state), n, D, mixed, NK,
                       partial);
  else
    hipLaunchKernelGGL(k_norm_partial<double>, dim3(g), dim3(kBlock), 0, s, CP(double, state), n, D, mixed, NK,
                       partial);
  HIP_RET(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, partial, g, out_dev);
  return (int)hipGetLastError();
}

extern "C" int sfamd_all_fock_probs(const void* state, void* probs, int n, int D, int mixed, int dtype,
                                    void* stream) {
  CHECK_ARG(state && probs && n >= 0 && D >= 1);
  hipStream_t s = (hipStream_t)stream;
  const int64_t NK = ipow(D, n);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_probs<float>, dim3(grid_for(NK)), dim3(kBlock), 0, s, CP(float, state), (float*)probs, n, D,
                       mixed, NK);
  else
    hipLaunchKernelGGL(k_probs<double>, dim3(grid_for(NK)), dim3(kBlock), 0, s, CP(double, state), (double*)probs, n,
                       D, mixed, NK);
  return (int)hipGetLastError();
}

extern "C" int sfamd_all_fock_probs_vjp(const void* state, const void* g_probs, void* g_state, int n, int D,
                                        int mixed, int dtype, void* stream) {
  CHECK_ARG(state && g_probs && g_state && n >= 0 && D >= 1);
  hipStream_t s = (hipStream_t)stream;
  const int64_t N = ipow(D, mixed ? 2 * n : n);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_probs_vjp<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, state),
                       (const float*)g_probs, MP(float, g_state), n, D, mixed, N);
  else
    hipLaunchKernelGGL(k_probs_vjp<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, state),
                       (const double*)g_probs, MP(double, g_state), n, D, mixed, N);
  return (int)hipGetLastError();
}

extern "C" int64_t sfamd_reduced_dm_partials(int n, int D, int k, int mixed) {
  (void)mixed;
  return ipow(D, 2 * k) * rdm_chunks(ipow(D, n - k)) * 2;
}

extern "C" int sfamd_reduced_dm(const void* state, void* out, int n, int D, int mixed, const int* modes, int k,
                                double* partial, int dtype, void* stream) {
  CHECK_ARG(state && out && partial && modes && D >= 1);
  ModeList ml;
  RET_IF(make_modelist(n, modes, k, &ml));
  hipStream_t s = (hipStream_t)stream;
  const int64_t NO = ipow(D, 2 * k), NR = ipow(D, n - k);
  const int64_t chunks = rdm_chunks(NR), chunk = (NR + chunks - 1) / chunks;
  CHECK_ARG(NO < (int64_t)1 << 31);
  dim3 grid((unsigned)NO, (unsigned)chunks);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_reduced_dm<float>, grid, dim3(kBlock), 0, s, CP(float, state), ml, D, mixed, NR, chunk,
                       partial);
  else
    hipLaunchKernelGGL(k_reduced_dm<double>, grid, dim3(kBlock), 0, s, CP(double, state), ml, D, mixed, NR, chunk,
                       partial);
  HIP_RET(hipGetLastError());
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_reduced_dm_final<float>, dim3(grid_for(NO)), dim3(kBlock), 0, s, partial, MP(float, out), NO,
                       (int)chunks);
  else
    hipLaunchKernelGGL(k_reduced_dm_final<double>, dim3(grid_for(NO)), dim3(kBlock), 0, s, partial, MP(double, out),
                       NO, (int)chunks);
  return (int)hipGetLastError();
}

extern "C" int sfamd_diagonal_expectation(const void* state, int n, int D, int mixed, const int* modes, int k,
                                          const double* values_dev, double* partial, double* out_dev, int dtype,
                                          void* stream) {
  CHECK_ARG(state && partial && out_dev && values_dev && modes && D >= 1);
  ModeList ml;
  // modes need not be sorted here; only duplicates are rejected
  if (n > 32 || k < 0 || k > n) return SFAMD_E_ARG;
  ml.n = n;
  ml.k = k;
  for (int t = 0; t < k; ++t) {
    if (modes[t] < 0 || modes[t] >= n) return SFAMD_E_ARG;
    for (int u = 0; u < t; ++u)
      if (modes[u] == modes[t]) return SFAMD_E_ARG;
    ml.kept[t] = modes[t];
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t NK = ipow(D, n);
  const int g = (int)reduce_grid(NK);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_diag_expect<float>, dim3(g), dim3(kBlock), 0, s, CP(float, state), ml, D, mixed, NK,
                       values_dev, partial);
  else
    hipLaunchKernelGGL(k_diag_expect<double>, dim3(g), dim3(kBlock), 0, s, CP(double, state), ml, D, mixed, NK,
                       values_dev, partial);
  HIP_RET(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, partial, g, out_dev);
  return (int)hipGetLastError();
}

extern "C" int64_t sfamd_overlap_partials(int64_t n_elems) { return 2 * reduce_grid(n_elems); }

extern "C" int sfamd_overlap(const void* a, const void* b, int64_t n_elems, double* partial, double* out_dev,
                             int dtype, void* stream) {
  CHECK_ARG(a && b && partial && out_dev && n_elems >= 0);
  hipStream_t s = (hipStream_t)stream;
  const int g = (int)reduce_grid(n_elems);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_overlap_partial<float>, dim3(g), dim3(kBlock), 0, s, CP(float, a), CP(float, b), n_elems,
                       partial);
  else
    hipLaunchKernelGGL(k_overlap_partial<double>, dim3(g), dim3(kBlock), 0, s, CP(double, a), CP(double, b), n_elems,
                       partial);
  HIP_RET(hipGetLastError());
  hipLaunchKernelGGL(k_overlap_final, dim3(1), dim3(kBlock), 0, s, partial, g, out_dev);
  return (int)hipGetLastError();
}

extern "C" int sfamd_fidelity_loss(const void* target, const void* psi, int64_t n_elems, double* partial,
                                   double* out_dev, void* g_psi, int dtype, void* stream) {
  RET_IF(sfamd_overlap(target, psi, n_elems, partial, out_dev, dtype, stream));
  if (!g_psi) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_fidelity_grad<float>, dim3(grid_for(n_elems)), dim3(kBlock), 0, s, CP(float, target),
                       out_dev, MP(float, g_psi), n_elems);
  else
    hipLaunchKernelGGL(k_fidelity_grad<double>, dim3(grid_for(n_elems)), dim3(kBlock), 0, s, CP(double, target),
                       out_dev, MP(double, g_psi), n_elems);
  return (int)hipGetLastError();
}

extern "C" int sfamd_number_loss(const void* psi, void* g_psi, int n, int D, double* partial, double* out_dev,
                                 int dtype, void* stream) {
  CHECK_ARG(psi && partial && out_dev && n >= 1 && D >= 1);
  hipStream_t s = (hipStream_t)stream;
  const int64_t N = ipow(D, n);
  const int g = (int)reduce_grid(N);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_number_loss<float>, dim3(g), dim3(kBlock), 0, s, CP(float, psi), MP(float, g_psi), n, D, N,
                       partial);
  else
    hipLaunchKernelGGL(k_number_loss<double>, dim3(g), dim3(kBlock), 0, s, CP(double, psi), MP(double, g_psi), n, D,
                       N, partial);
  HIP_RET(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, partial, g, out_dev);
  return (int)hipGetLastError();
}
