// Gate-table builder: Fock-basis matrices of R, K, D, S, BS and the loss
// channel, and their derivatives with respect to the gate parameters.
//
// All recurrences run in FP64 in LDS and each entry is rounded once to the
// state precision when it is written to the table.  One workgroup builds one
// operation's table entry.
//
// The recurrences follow from the Heisenberg action of each gate on the
// ladder operators (each step uses only smaller indices, so the truncated
// matrix is exact: nothing is renormalised, as in
// strawberryfields/backends/fockbackend/ops.py:219-254,318-329, which call
// thewalrus.fock_gradients):
//   D(a), a = r e^{i phi}:  D a^dag = (a^dag - a*) D
//       <m|D|0> = a/sqrt(m) <m-1|D|0>,  <0|D|0> = exp(-r^2/2)
//       <m|D|n> = (sqrt(m) <m-1|D|n-1> - a* <m|D|n-1>) / sqrt(n)
//   S(z), z = r e^{i phi}:  S a^dag S^dag = a^dag cosh r + a e^{-i phi} sinh r
//       <m|S|0> = -e^{i phi} tanh r sqrt((m-1)/m) <m-2|S|0>,  <0|S|0> = sqrt(sech r)
//       <m|S|n> = (sqrt(n-1) e^{-i phi} tanh r <m|S|n-2> + sqrt(m) sech r <m-1|S|n-1>) / sqrt(n)
//   BS(theta, phi): BS a2^dag = (a2^dag cos - e^{-i phi} sin a1^dag) BS,
//                   BS a1^dag = (a1^dag cos + e^{i phi} sin a2^dag) BS
//       which links photon-number block s to block s-1.
// Derivatives are carried through the same recurrences (forward mode), so they
// are the exact derivatives of the truncated matrices.
#include "common.h"
#include "fused.h"

namespace {

using cd = cpx<double>;

__device__ __forceinline__ cd cexpi(double ph) { return mk<double>(cos(ph), sin(ph)); }

template <typename T>
__device__ __forceinline__ void put(cpx<T>* dst, int64_t i, cd v) {
  dst[i] = mk<T>((T)v.re, (T)v.im);
}

// ---------------------------------------------------------------------------
template <typename T>
__device__ void build_diag(int kind, double p, int D, cpx<T>* out) {
  for (int k = threadIdx.x; k < D; k += blockDim.x) {
    const double w = (kind == SFAMD_OP_KERR) ? (double)k * k : (double)k;
    put(out, k, cexpi(p * w));
  }
}

// displacement: out[0..D^2) = U, then dU/dr, then dU/dphi (row = out index)
template <typename T>
__device__ void build_disp(double r, double phi, int D, cpx<T>* out, cd (*col)[2][SFAMD_MAX_CUTOFF]) {
  const int64_t DD = (int64_t)D * D;
  const cd a = mk<double>(r * cos(phi), r * sin(phi));
  const cd da[2] = {mk<double>(cos(phi), sin(phi)), mk<double>(-a.im, a.re)};  // d/dr, d/dphi
  const cd ac = cconj(a);
  const cd dac[2] = {cconj(da[0]), cconj(da[1])};
  // column 0: sequential chain
  if (threadIdx.x == 0) {
    cd v = mk<double>(exp(-0.5 * r * r), 0.0);
    cd dv[2] = {cscale(v, -r), mk<double>(0.0, 0.0)};
    for (int m = 0; m < D; ++m) {
      if (m > 0) {
        const double is = 1.0 / sqrt((double)m);
        cd nv = cscale(cmul(a, v), is);
        for (int j = 0; j < 2; ++j) {
          cd t = cmul(da[j], v);
          cmac(t, a, dv[j]);
          dv[j] = cscale(t, is);
        }
        v = nv;
      }
      col[0][0][m] = v;
      col[1][0][m] = dv[0];
      col[2][0][m] = dv[1];
      put(out, (int64_t)m * D, v);
      put(out, DD + (int64_t)m * D, dv[0]);
      put(out, 2 * DD + (int64_t)m * D, dv[1]);
    }
  }
  __syncthreads();
  int cur = 0;
  for (int n = 1; n < D; ++n) {
    const int nxt = cur ^ 1;
    const double isn = 1.0 / sqrt((double)n);
    for (int m = threadIdx.x; m < D; m += blockDim.x) {
      const double sm = sqrt((double)m);
      // value
      cd v = mk<double>(0.0, 0.0);
      if (m > 0) v = cscale(col[0][cur][m - 1], sm);
      v = csub(v, cmul(ac, col[0][cur][m]));
      v = cscale(v, isn);
      col[0][nxt][m] = v;
      put(out, (int64_t)m * D + n, v);
      for (int j = 0; j < 2; ++j) {
        cd t = mk<double>(0.0, 0.0);
        if (m > 0) t = cscale(col[1 + j][cur][m - 1], sm);
        t = csub(t, cmul(dac[j], col[0][cur][m]));
        t = csub(t, cmul(ac, col[1 + j][cur][m]));
        t = cscale(t, isn);
        col[1 + j][nxt][m] = t;
        put(out, (j + 1) * DD + (int64_t)m * D + n, t);
      }
    }
    __syncthreads();
    cur = nxt;
  }
}

// squeezing: out[0..D^2) = U, dU/dr, dU/dphi; zero unless m + n even
template <typename T>
__device__ void build_sq(double r, double phi, int D, cpx<T>* out, cd (*col3)[3][SFAMD_MAX_CUTOFF]) {
  const int64_t DD = (int64_t)D * D;
  const double th = tanh(r), sech = 1.0 / cosh(r);
  const double dth = sech * sech, dsech = -sech * th;
  const cd e = mk<double>(cos(phi), sin(phi));
  const cd ie = mk<double>(-e.im, e.re);  // i e
  // R00 = -e th (column 0), R11 = conj(e) th, R01 = sech
  const cd R00 = cscale(e, -th);
  const cd dR00[2] = {cscale(e, -dth), cscale(ie, -th)};
  const cd R11 = cscale(cconj(e), th);
  const cd dR11[2] = {cscale(cconj(e), dth), cscale(cconj(ie), th)};  // d/dphi conj(e) = conj(i e)
  const double R01 = sech;
  const double dR01[2] = {dsech, 0.0};
  // column 0 (slot 0 of the rolling buffer holds column n, slot 1 n-1, slot 2 n-2)
  if (threadIdx.x == 0) {
    const double s0 = sqrt(sech);
    cd v = mk<double>(s0, 0.0);
    cd dv[2] = {mk<double>(-0.5 * s0 * th, 0.0), mk<double>(0.0, 0.0)};
    for (int m = 0; m < D; ++m) {
      for (int j = 0; j < 3; ++j) col3[j][0][m] = mk<double>(0.0, 0.0);
    }
    for (int m = 1; m < D; m += 2) {
      put(out, (int64_t)m * D, mk<double>(0.0, 0.0));
      put(out, DD + (int64_t)m * D, mk<double>(0.0, 0.0));
      put(out, 2 * DD + (int64_t)m * D, mk<double>(0.0, 0.0));
    }
    for (int m = 0; m < D; m += 2) {
      if (m > 0) {
        const double f = sqrt((double)(m - 1) / (double)m);
        cd nv = cscale(cmul(R00, v), f);
        for (int j = 0; j < 2; ++j) {
          cd t = cmul(dR00[j], v);
          cmac(t, R00, dv[j]);
          dv[j] = cscale(t, f);
        }
        v = nv;
      }
      col3[0][0][m] = v;
      col3[1][0][m] = dv[0];
      col3[2][0][m] = dv[1];
      put(out, (int64_t)m * D, v);
      put(out, DD + (int64_t)m * D, dv[0]);
      put(out, 2 * DD + (int64_t)m * D, dv[1]);
    }
  }
  __syncthreads();
  // rolling column slots: c0 = column n-1, c1 = column n-2, new column into c2
  int c_n1 = 0, c_n2 = 1, c_new = 2;
  for (int m = threadIdx.x; m < D; m += blockDim.x) {
    for (int j = 0; j < 3; ++j) col3[j][1][m] = mk<double>(0.0, 0.0);
  }
  __syncthreads();
  for (int n = 1; n < D; ++n) {
    const double isn = 1.0 / sqrt((double)n);
    const double sn1 = sqrt((double)(n - 1));
    for (int m = threadIdx.x; m < D; m += blockDim.x) {
      cd v = mk<double>(0.0, 0.0), dv[2] = {mk<double>(0.0, 0.0), mk<double>(0.0, 0.0)};
      if (((m + n) & 1) == 0) {
        const double sm = sqrt((double)m);
        if (n >= 2) {
          const cd x = col3[0][c_n2][m];
          cmac(v, R11, cscale(x, sn1));
          for (int j = 0; j < 2; ++j) {
            cmac(dv[j], dR11[j], cscale(x, sn1));
            cmac(dv[j], R11, cscale(col3[1 + j][c_n2][m], sn1));
          }
        }
        if (m > 0) {
          const cd y = col3[0][c_n1][m - 1];
          v = cadd(v, cscale(y, sm * R01));
          for (int j = 0; j < 2; ++j) {
            dv[j] = cadd(dv[j], cscale(y, sm * dR01[j]));
            dv[j] = cadd(dv[j], cscale(col3[1 + j][c_n1][m - 1], sm * R01));
          }
        }
        v = cscale(v, isn);
        dv[0] = cscale(dv[0], isn);
        dv[1] = cscale(dv[1], isn);
      }
      put(out, (int64_t)m * D + n, v);
      put(out, DD + (int64_t)m * D + n, dv[0]);
      put(out, 2 * DD + (int64_t)m * D + n, dv[1]);
      col3[0][c_new][m] = v;
      col3[1][c_new][m] = dv[0];
      col3[2][c_new][m] = dv[1];
    }
    __syncthreads();
    const int t = c_n2;
    c_n2 = c_n1;
    c_n1 = c_new;
    c_new = t;
  }
}

constexpr int kBSMax = SFAMD_MAX_CUTOFF_BS;
constexpr int kBSPerThread = (kBSMax * kBSMax + 255) / 256;
constexpr int kMax3 = (2 * kBSMax * kBSMax * 2 > 7 * SFAMD_MAX_CUTOFF_V * SFAMD_MAX_CUTOFF_V)
                          ? 2 * kBSMax * kBSMax * 2
                          : 7 * SFAMD_MAX_CUTOFF_V * SFAMD_MAX_CUTOFF_V;
constexpr int kBuildLdsDoubles = kMax3 > 3 * 3 * SFAMD_MAX_CUTOFF * 2 ? kMax3 : 3 * 3 * SFAMD_MAX_CUTOFF * 2;

// Passive two-mode gates (BS, MZ): the gate maps a_j^dag -> sum_i W[i][j] a_i^dag
// for a 2x2 unitary W (BS: W = [[cos, -e^{-i phi} sin], [e^{i phi} sin, cos]]),
// so G a2^dag = (W[0][1] a1^dag + W[1][1] a2^dag) G and
//    G a1^dag = (W[0][0] a1^dag + W[1][0] a2^dag) G,
// which links photon-number block s to block s - 1.  dW carries the derivative
// with respect to the first parameter through the same recurrence.
// Output: blocks of G, then blocks of dG/dp0.
template <typename T>
__device__ void build_passive(const cd (&W)[2][2], const cd (&dW)[2][2], int D, cpx<T>* out,
                              cd (*prev)[kBSMax * kBSMax]) {
  const int64_t nb = bs_nb(D);
  if (threadIdx.x == 0) {
    prev[0][0] = mk<double>(1.0, 0.0);
    prev[1][0] = mk<double>(0.0, 0.0);
    put(out, 0, mk<double>(1.0, 0.0));
    put(out, nb, mk<double>(0.0, 0.0));
  }
  __syncthreads();
  for (int s = 1; s <= 2 * D - 2; ++s) {
    const int lo = bs_lo(s, D), hi = bs_hi(s, D), c = hi - lo + 1;
    const int plo = bs_lo(s - 1, D), pc = bs_hi(s - 1, D) - plo + 1;
    const int64_t off = bs_block_off(s, D);
    cd v[kBSPerThread], dv[kBSPerThread];
    int cnt = 0;
    for (int idx = threadIdx.x; idx < c * c; idx += blockDim.x, ++cnt) {
      const int N = lo + idx / c, n = lo + idx % c;
      const int M = s - N, m = s - n;
      cd a = mk<double>(0.0, 0.0), da = mk<double>(0.0, 0.0);
      // input (n, m): raise mode 2 from (n, m-1) if m > 0, else mode 1 from (s-1, 0)
      const int pn = m > 0 ? n : s - 1;
      const cd w1 = m > 0 ? W[0][1] : W[0][0], w2 = m > 0 ? W[1][1] : W[1][0];
      const cd dw1 = m > 0 ? dW[0][1] : dW[0][0], dw2 = m > 0 ? dW[1][1] : dW[1][0];
      if (N >= 1) {  // w1 sqrt(N) P[N-1][pn]
        const int pi = (N - 1 - plo) * pc + (pn - plo);
        const double sN = sqrt((double)N);
        a = cadd(a, cscale(cmul(w1, prev[0][pi]), sN));
        da = cadd(da, cscale(cadd(cmul(dw1, prev[0][pi]), cmul(w1, prev[1][pi])), sN));
      }
      if (M >= 1) {  // w2 sqrt(M) P[N][pn]
        const int pi = (N - plo) * pc + (pn - plo);
        const double sM = sqrt((double)M);
        a = cadd(a, cscale(cmul(w2, prev[0][pi]), sM));
        da = cadd(da, cscale(cadd(cmul(dw2, prev[0][pi]), cmul(w2, prev[1][pi])), sM));
      }
      const double is = 1.0 / sqrt((double)(m > 0 ? m : s));
      v[cnt] = cscale(a, is);
      dv[cnt] = cscale(da, is);
    }
    __syncthreads();
    cnt = 0;
    for (int idx = threadIdx.x; idx < c * c; idx += blockDim.x, ++cnt) {
      prev[0][idx] = v[cnt];
      prev[1][idx] = dv[cnt];
      put(out, off + idx, v[cnt]);
      put(out, nb + off + idx, dv[cnt]);
    }
    __syncthreads();
  }
}

// beamsplitter BS(theta, phi): blocks of B then of dB/dtheta
template <typename T>
__device__ void build_bs(double theta, double phi, int D, cpx<T>* out, cd (*prev)[kBSMax * kBSMax]) {
  const double ct = cos(theta), st = sin(theta);
  const cd e = cexpi(phi), ec = cconj(e);
  const cd W[2][2] = {{mk<double>(ct, 0.0), cscale(ec, -st)}, {cscale(e, st), mk<double>(ct, 0.0)}};
  const cd dW[2][2] = {{mk<double>(-st, 0.0), cscale(ec, -ct)}, {cscale(e, ct), mk<double>(-st, 0.0)}};
  build_passive<T>(W, dW, D, out, prev);
}

// Mach-Zehnder MZ(phi_in, phi_ex) = BS(pi/4, pi/2) (R(phi_in) x 1) BS(pi/4, pi/2)
// (R(phi_ex) x 1) (strawberryfields/ops.py:1951-1991): W = B Rin B Rex with
// B = [[c, i c], [i c, c]], c = 1/sqrt 2.  Blocks of MZ, then dMZ/dphi_in.
template <typename T>
__device__ void build_mz(double phi_in, double phi_ex, int D, cpx<T>* out, cd (*prev)[kBSMax * kBSMax]) {
  const cd ei = cexpi(phi_in), ee = cexpi(phi_ex);
  const cd I = mk<double>(0.0, 1.0);
  // B diag(x, 1) B diag(ee, 1) = 1/2 [[(x - 1) ee, i (x + 1)], [i (x + 1) ee, 1 - x]]
  const cd one = mk<double>(1.0, 0.0);
  cd W[2][2], dW[2][2];
  W[0][0] = cscale(cmul(csub(ei, one), ee), 0.5);
  W[0][1] = cscale(cmul(I, cadd(ei, one)), 0.5);
  W[1][0] = cscale(cmul(cmul(I, cadd(ei, one)), ee), 0.5);
  W[1][1] = cscale(csub(one, ei), 0.5);
  const cd dx = cmul(I, ei);  // d(e^{i phi_in})/dphi_in
  dW[0][0] = cscale(cmul(dx, ee), 0.5);
  dW[0][1] = cscale(cmul(I, dx), 0.5);
  dW[1][0] = cscale(cmul(cmul(I, dx), ee), 0.5);
  dW[1][1] = cscale(dx, -0.5);
  build_passive<T>(W, dW, D, out, prev);
}

// cross-Kerr CK(kappa) = exp(i kappa n1 n2): diagonal blocks, then dCK/dkappa
template <typename T>
__device__ void build_ck(double kappa, int D, cpx<T>* out) {
  const int64_t nb = bs_nb(D);
  for (int s = 0; s <= 2 * D - 2; ++s) {
    const int lo = bs_lo(s, D), c = bs_hi(s, D) - lo + 1;
    const int64_t off = bs_block_off(s, D);
    for (int idx = threadIdx.x; idx < c * c; idx += blockDim.x) {
      const int N = lo + idx / c, n = lo + idx % c;
      cd v = mk<double>(0.0, 0.0), dv = v;
      if (N == n) {
        const double w = (double)N * (double)(s - N);
        v = cexpi(kappa * w);
        dv = mk<double>(-w * v.im, w * v.re);
      }
      put(out, off + idx, v);
      put(out, nb + off + idx, dv);
    }
  }
}

// two-mode squeezing S2(r, phi) = exp(r (e^{i phi} a1^dag a2^dag - e^{-i phi} a1 a2))
// (strawberryfields/ops.py:2036-2054).  From S2^dag a1 S2 = cosh r a1 + e^{i phi}
// sinh r a2^dag (and 1 <-> 2) the entries G[N,M,n,m] = <N,M|S2|n,m> obey
//   sqrt(N) G[N,M,n,m] = sech sqrt(n) G[N-1,M,n-1,m] + e^{i phi} tanh sqrt(M) G[N-1,M-1,n,m]
//   sqrt(M) G[0,M,n,m] = sech sqrt(m) G[0,M-1,n,m-1]
//   sqrt(n) G[0,0,n,n] = -e^{-i phi} tanh sqrt(n) G[0,0,n-1,n-1]
// with G[0,0,0,0] = sech r.  Every coefficient is at most 1 in magnitude (no
// cancelling alternating sums), and each right-hand side lies one level lower
// in lambda = N + m = M + n, so a level is computed from the previous one
// alone: level lambda is the grid of (N, M) with n = lambda - M, m = lambda - N
// in range, kept in LDS (value and d/dr) between levels.  Output: blocks
// s = D - 1 + d (d = N - M = n - m) in the BS layout (first index N, second n),
// then the blocks of dS2/dr.
template <typename T>
__device__ void build_s2(double r, double phi, int D, cpx<T>* out, cd (*prev)[kBSMax * kBSMax]) {
  const int64_t nb = bs_nb(D);
  const double sech = 1.0 / cosh(r), th = tanh(r);
  const cd e = cexpi(phi), ec = cconj(e);
  const int DD = D * D;
  if (threadIdx.x == 0) {
    prev[0][0] = mk<double>(sech, 0.0);
    prev[1][0] = mk<double>(-sech * th, 0.0);
    const int64_t off = bs_block_off(D - 1, D);  // d = 0, entry (0, 0)
    put(out, off, prev[0][0]);
    put(out, nb + off, prev[1][0]);
  }
  __syncthreads();
  for (int lam = 1; lam <= 2 * D - 2; ++lam) {
    cd v[kBSPerThread], dv[kBSPerThread];
    int cnt = 0;
    for (int idx = threadIdx.x; idx < DD; idx += blockDim.x, ++cnt) {
      const int N = idx / D, M = idx % D, n = lam - M, m = lam - N;
      v[cnt] = dv[cnt] = mk<double>(0.0, 0.0);
      if (n < 0 || n >= D || m < 0 || m >= D) continue;
      cd a = mk<double>(0.0, 0.0), da = a;
      if (N > 0) {
        if (n > 0) {  // sech sqrt(n) G[N-1, M, n-1, m]
          const cd g = prev[0][(N - 1) * D + M], dg = prev[1][(N - 1) * D + M];
          const double c = sech * sqrt((double)n);
          a = cadd(a, cscale(g, c));
          da = cadd(da, cscale(csub(dg, cscale(g, th)), c));
        }
        if (M > 0) {  // e tanh sqrt(M) G[N-1, M-1, n, m]
          const cd g = prev[0][(N - 1) * D + M - 1], dg = prev[1][(N - 1) * D + M - 1];
          const double sM = sqrt((double)M);
          a = cadd(a, cscale(cmul(e, g), th * sM));
          da = cadd(da, cscale(cmul(e, cadd(cscale(dg, th), cscale(g, sech * sech))), sM));
        }
        const double is = 1.0 / sqrt((double)N);
        a = cscale(a, is);
        da = cscale(da, is);
      } else if (M > 0) {  // N = 0, so m = lam >= M >= 1
        const cd g = prev[0][M - 1], dg = prev[1][M - 1];
        const double c = sech * sqrt((double)m) / sqrt((double)M);
        a = cscale(g, c);
        da = cscale(csub(dg, cscale(g, th)), c);
      } else {  // N = M = 0, n = m = lam
        const cd g = prev[0][0], dg = prev[1][0];
        a = cscale(cmul(ec, g), -th);
        da = cscale(cmul(ec, cadd(cscale(dg, th), cscale(g, sech * sech))), -1.0);
      }
      v[cnt] = a;
      dv[cnt] = da;
    }
    __syncthreads();
    cnt = 0;
    for (int idx = threadIdx.x; idx < DD; idx += blockDim.x, ++cnt) {
      const int N = idx / D, M = idx % D, n = lam - M, m = lam - N;
      if (n < 0 || n >= D || m < 0 || m >= D) continue;
      prev[0][idx] = v[cnt];
      prev[1][idx] = dv[cnt];
      const int s = D - 1 + N - M, lo = bs_lo(s, D), c = bs_hi(s, D) - lo + 1;
      const int64_t at = bs_block_off(s, D) + (int64_t)(N - lo) * c + (n - lo);
      put(out, at, v[cnt]);
      put(out, nb + at, dv[cnt]);
    }
    __syncthreads();
  }
}

// cubic phase V(gamma) = expm(i gamma/(3 hbar) x^3), hbar = 2, x = a + a^dag on
// the truncated basis (fockbackend/ops.py:297-306 builds x^3 from the truncated
// x and exponentiates it; the gate is therefore that matrix exponential, not a
// truncation of the exact operator).  H = (gamma / 6) x^3 is real symmetric;
// exp(iH) by scaling and squaring of a degree-20 Taylor polynomial in FP64, and
// dV/dgamma = (i x^3 / 6) V.  Output: V, dV/dgamma, zeros (row = out index).
template <typename T>
__device__ void build_cubic(double gamma, int D, cpx<T>* out, double* sm) {
  // LDS: X3 (real D*D), E, P, Q (complex D*D each)
  double* X3 = sm;
  cd* E = reinterpret_cast<cd*>(sm + D * D);
  cd* P = E + D * D;
  cd* Q = P + D * D;
  const int DD = D * D;
  // x^3: x[i][j] = sqrt(j) [i == j-1] + sqrt(i) [i == j+1];  x^3 = x (x x)
  for (int e = threadIdx.x; e < DD; e += blockDim.x) {
    const int i = e / D, j = e % D;
    double x2 = 0.0;  // (x x)[i][j]
    for (int k = 0; k < D; ++k) {
      const double xik = (k == i + 1 ? sqrt((double)k) : 0.0) + (k + 1 == i ? sqrt((double)i) : 0.0);
      const double xkj = (j == k + 1 ? sqrt((double)j) : 0.0) + (j + 1 == k ? sqrt((double)k) : 0.0);
      x2 += xik * xkj;
    }
    P[e] = mk<double>(x2, 0.0);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < DD; e += blockDim.x) {
    const int i = e / D, j = e % D;
    double v = 0.0;
    if (i > 0) v += sqrt((double)i) * P[(i - 1) * D + j].re;
    if (i + 1 < D) v += sqrt((double)(i + 1)) * P[(i + 1) * D + j].re;
    X3[e] = v;
  }
  __syncthreads();
  // scale: ||H||_1 / 2^k <= 1/2
  __shared__ double norm1;
  if (threadIdx.x == 0) {
    double mx = 0.0;
    for (int j = 0; j < D; ++j) {
      double cs = 0.0;
      for (int i = 0; i < D; ++i) cs += fabs(X3[i * D + j]);
      mx = cs > mx ? cs : mx;
    }
    norm1 = mx * fabs(gamma) / 6.0;
  }
  __syncthreads();
  int k = 0;  // bounded: a non-finite gamma stops at once and gives NaN entries
  while (k < 64 && norm1 / ldexp(1.0, k) > 0.5) ++k;
  const double h = gamma / 6.0 / ldexp(1.0, k);  // A = i h X3
  // E = sum_{m <= 20} A^m / m!  (P holds A^m / m!)
  for (int e = threadIdx.x; e < DD; e += blockDim.x) {
    const int i = e / D, j = e % D;
    E[e] = mk<double>(i == j ? 1.0 : 0.0, 0.0);
    P[e] = E[e];
  }
  __syncthreads();
  for (int m = 1; m <= 20; ++m) {
    for (int e = threadIdx.x; e < DD; e += blockDim.x) {
      const int i = e / D, j = e % D;
      cd acc = mk<double>(0.0, 0.0);
      for (int q = 0; q < D; ++q) acc = cadd(acc, cscale(P[i * D + q], X3[q * D + j]));
      // times i h / m
      Q[e] = mk<double>(-acc.im * h / m, acc.re * h / m);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < DD; e += blockDim.x) {
      P[e] = Q[e];
      E[e] = cadd(E[e], Q[e]);
    }
    __syncthreads();
  }
  for (int it = 0; it < k; ++it) {  // E <- E E
    for (int e = threadIdx.x; e < DD; e += blockDim.x) {
      const int i = e / D, j = e % D;
      cd acc = mk<double>(0.0, 0.0);
      for (int q = 0; q < D; ++q) acc = cadd(acc, cmul(E[i * D + q], E[q * D + j]));
      Q[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < DD; e += blockDim.x) E[e] = Q[e];
    __syncthreads();
  }
  for (int e = threadIdx.x; e < DD; e += blockDim.x) {
    const int i = e / D, j = e % D;
    cd acc = mk<double>(0.0, 0.0);  // (x^3 / 6) V
    for (int q = 0; q < D; ++q) acc = cadd(acc, cscale(E[q * D + j], X3[i * D + q] / 6.0));
    put(out, e, E[e]);
    put(out, DD + e, mk<double>(-acc.im, acc.re));
    put(out, 2 * DD + e, mk<double>(0.0, 0.0));
  }
}

// loss channel superoperator coefficients (real):
//   c[p][q][l] = (1-T)^l / l! * T^{(p+q)/2} * sqrt((p+l)!/p! * (q+l)!/q!)
// for p+l < D and q+l < D (0 elsewhere), then dc/dT.  This is the identity of
// strawberryfields/ops.py:1369-1376 and tfbackend/ops.py:288-315; it equals the
// Kraus sum of fockbackend/ops.py:471-490 (at T = 0 only l-sums onto |0><0| remain).
template <typename T>
__device__ void build_loss(double Tt, int D, cpx<T>* out) {
  T* c = reinterpret_cast<T*>(out);
  const int64_t D3 = (int64_t)D * D * D;
  for (int64_t idx = threadIdx.x; idx < D3; idx += blockDim.x) {
    const int l = (int)(idx % D);
    const int q = (int)((idx / D) % D);
    const int p = (int)(idx / ((int64_t)D * D));
    double v = 0.0, dv = 0.0;
    if (p + l < D && q + l < D) {
      double K = 1.0;
      for (int k = 1; k <= l; ++k) K *= sqrt((double)(p + k) * (double)(q + k)) / (double)k;
      const double h = 0.5 * (p + q);
      const double omt = 1.0 - Tt;
      const double pl = (l == 0) ? 1.0 : pow(omt, (double)l);
      const double th = (p + q == 0) ? 1.0 : pow(Tt, h);
      v = pl * th * K;
      double d = 0.0;
      if (l >= 1) d -= (double)l * ((l == 1) ? 1.0 : pow(omt, (double)(l - 1))) * th;
      if (p + q > 0) d += pl * h * pow(Tt, h - 1.0);
      dv = d * K;
    }
    c[idx] = (T)v;
    c[D3 + idx] = (T)dv;
  }
}

// One gate's region of a fused stage's matrix image: the head of its table
// entry (the matrix; for the adjoint also the first parameter's derivative),
// with beamsplitters up to kFullBsMaxD expanded to the per-fiber layout
// M[r][k][p] at (r D + k) D + p: row r (output first-mode digit), column k
// (input first-mode digit) of blockdiag(B_p, B_(p+D)), the gate on the
// modular anti-diagonal {(k, (p - k) mod D)}; entries joining the two photon-
// number blocks are zero.
template <typename T>
__device__ void write_stage_image(int kind, int D, const cpx<T>* __restrict__ entry, cpx<T>* __restrict__ dst,
                                  bool derivs) {
  const int parts = derivs ? 2 : 1;
  if (kind == SFAMD_OP_BEAMSPLITTER && D <= sfamd::kFullBsMaxD) {
    const int DD = D * D, D3 = DD * D;
    const int64_t nb = bs_nb(D);
    for (int l = threadIdx.x; l < parts * D3; l += blockDim.x) {
      const int part = l / D3, e = l - part * D3;
      const int r = e / DD, k = (e / D) % D, p = e % D;
      cpx<T> v = mk<T>(0, 0);
      if ((r <= p) == (k <= p)) {
        const int s = r <= p ? p : p + D;
        const int lo = bs_lo(s, D), c = bs_hi(s, D) - lo + 1;
        v = entry[part * nb + bs_block_off(s, D) + (int64_t)(r - lo) * c + (k - lo)];
      }
      dst[l] = v;
    }
    return;
  }
  int64_t cnt;
  switch (kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      cnt = D;
      break;
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE:
      cnt = (int64_t)parts * D * D;
      break;
    case SFAMD_OP_BEAMSPLITTER:
      cnt = parts * bs_nb(D);
      break;
    default:
      cnt = 0;
      break;
  }
  for (int64_t l = threadIdx.x; l < cnt; l += blockDim.x) dst[l] = entry[l];
}

template <typename T>
__global__ void __launch_bounds__(256) k_build(sfamd::BuildBatch b, int D, const double* params, cpx<T>* table) {
  // one LDS buffer serves every builder: passive-gate blocks, D/S columns, or
  // the cubic-phase matrices (x^3 real, three complex D x D)
  __shared__ __align__(16) double sm_raw[kBuildLdsDoubles];
  cd(*sm_bs)[kBSMax * kBSMax] = reinterpret_cast<cd(*)[kBSMax * kBSMax]>(sm_raw);
  cd(*sm_col)[3][SFAMD_MAX_CUTOFF] = reinterpret_cast<cd(*)[3][SFAMD_MAX_CUTOFF]>(sm_raw);
  const sfamd::BuildDesc d = b.d[blockIdx.x];
  double p0 = b.p0, p1 = b.p1;
  if (b.first >= 0) {
    const int64_t pidx = (int64_t)b.first + blockIdx.x;
    p0 = params[2 * pidx];
    p1 = params[2 * pidx + 1];
  }
  cpx<T>* out = table + d.off;
  switch (d.kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      build_diag<T>(d.kind, p0, D, out);
      break;
    case SFAMD_OP_DISPLACEMENT:
      build_disp<T>(p0, p1, D, out, reinterpret_cast<cd(*)[2][SFAMD_MAX_CUTOFF]>(&sm_col[0][0][0]));
      break;
    case SFAMD_OP_SQUEEZE:
      build_sq<T>(p0, p1, D, out, sm_col);
      break;
    case SFAMD_OP_BEAMSPLITTER:
      build_bs<T>(p0, p1, D, out, sm_bs);
      break;
    case SFAMD_OP_MZGATE:
      build_mz<T>(p0, p1, D, out, sm_bs);
      break;
    case SFAMD_OP_CROSS_KERR:
      build_ck<T>(p0, D, out);
      break;
    case SFAMD_OP_TWO_MODE_SQUEEZE:
      build_s2<T>(p0, p1, D, out, sm_bs);
      break;
    case SFAMD_OP_CUBIC_PHASE:
      build_cubic<T>(p0, D, out, sm_raw);
      break;
    case SFAMD_OP_LOSS:
      build_loss<T>(p0, D, out);
      break;
    default:
      break;
  }
  if (d.img_fwd >= 0 || d.img_adj >= 0) {
    // the entry was written by this workgroup: the barrier orders those
    // stores before the reads below (same compute unit)
    __syncthreads();
    if (d.img_fwd >= 0) write_stage_image<T>(d.kind, D, out, table + d.img_fwd, false);
    if (d.img_adj >= 0) write_stage_image<T>(d.kind, D, out, table + d.img_adj, true);
  }
}

}  // namespace

namespace sfamd {

int64_t table_size(int kind, int D) {
  switch (kind) {
    case SFAMD_OP_ROTATION:
    case SFAMD_OP_KERR:
      return D;
    case SFAMD_OP_DISPLACEMENT:
    case SFAMD_OP_SQUEEZE:
      return 3 * (int64_t)D * D;
    case SFAMD_OP_BEAMSPLITTER:
    case SFAMD_OP_TWO_MODE_SQUEEZE:
    case SFAMD_OP_MZGATE:
    case SFAMD_OP_CROSS_KERR:
      return 2 * bs_nb(D);
    case SFAMD_OP_CUBIC_PHASE:
      return 3 * (int64_t)D * D;
    case SFAMD_OP_LOSS:
      return (int64_t)D * D * D;
    case SFAMD_OP_MIX:
      return 0;
    default:
      return -1;
  }
}

int launch_build(const BuildBatch& b, int count, int D, const double* params, void* table, int dtype,
                 hipStream_t s) {
  if (count <= 0) return 0;
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_build<float>, dim3(count), dim3(256), 0, s, b, D, params, (cpx<float>*)table);
  else
    hipLaunchKernelGGL(k_build<double>, dim3(count), dim3(256), 0, s, b, D, params, (cpx<double>*)table);
  return (int)hipGetLastError();
}

}  // namespace sfamd
