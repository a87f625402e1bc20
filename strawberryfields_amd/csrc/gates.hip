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

// beamsplitter: photon-number blocks of B then of dB/dtheta
template <typename T>
__device__ void build_bs(double theta, double phi, int D, cpx<T>* out, cd (*prev)[kBSMax * kBSMax]) {
  const int64_t nb = bs_nb(D);
  const double ct = cos(theta), st = sin(theta);
  const cd e = mk<double>(cos(phi), sin(phi));
  const cd ec = cconj(e);
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
      if (m > 0) {
        // (ct sqrt(M) P[N][n] - ec st sqrt(N) P[N-1][n]) / sqrt(m)
        if (M >= 1) {
          const int pi = (N - plo) * pc + (n - plo);
          const double sM = sqrt((double)M);
          a = cadd(a, cscale(prev[0][pi], ct * sM));
          da = cadd(da, cscale(prev[0][pi], -st * sM));
          da = cadd(da, cscale(prev[1][pi], ct * sM));
        }
        if (N >= 1) {
          const int pi = (N - 1 - plo) * pc + (n - plo);
          const double sN = sqrt((double)N);
          a = csub(a, cscale(cmul(ec, prev[0][pi]), st * sN));
          da = csub(da, cscale(cmul(ec, prev[0][pi]), ct * sN));
          da = csub(da, cscale(cmul(ec, prev[1][pi]), st * sN));
        }
        const double is = 1.0 / sqrt((double)m);
        a = cscale(a, is);
        da = cscale(da, is);
      } else {
        // n == s: (ct sqrt(N) P[N-1][s-1] + e st sqrt(M) P[N][s-1]) / sqrt(s)
        if (N >= 1) {
          const int pi = (N - 1 - plo) * pc + (s - 1 - plo);
          const double sN = sqrt((double)N);
          a = cadd(a, cscale(prev[0][pi], ct * sN));
          da = cadd(da, cscale(prev[0][pi], -st * sN));
          da = cadd(da, cscale(prev[1][pi], ct * sN));
        }
        if (M >= 1) {
          const int pi = (N - plo) * pc + (s - 1 - plo);
          const double sM = sqrt((double)M);
          a = cadd(a, cscale(cmul(e, prev[0][pi]), st * sM));
          da = cadd(da, cscale(cmul(e, prev[0][pi]), ct * sM));
          da = cadd(da, cscale(cmul(e, prev[1][pi]), st * sM));
        }
        const double is = 1.0 / sqrt((double)s);
        a = cscale(a, is);
        da = cscale(da, is);
      }
      v[cnt] = a;
      dv[cnt] = da;
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

template <typename T>
__global__ void __launch_bounds__(256) k_build(sfamd::BuildBatch b, int D, const double* params, cpx<T>* table) {
  __shared__ cd sm_bs[2][kBSMax * kBSMax];
  __shared__ cd sm_col[3][3][SFAMD_MAX_CUTOFF];
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
    case SFAMD_OP_LOSS:
      build_loss<T>(p0, D, out);
      break;
    default:
      break;
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
      return 2 * bs_nb(D);
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
