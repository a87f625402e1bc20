// Measurement and preparation kernels: marginal Fock distributions, the
// projection |0><x| of measured modes, the homodyne density on a grid, and the
// embedding of a prepared multi-mode state into the register.
//
// Layouts as elsewhere in the library: a ket [D]^n in C order (mode 0 most
// significant, stride D^(n-1-m)); a DM [D]^2n with interleaved axes (mode m
// bra stride D^(2(n-1-m)+1), ket stride D^(2(n-1-m))).
#include "common.h"

namespace {

template <typename T>
using C = cpx<T>;

#define CP(T, p) reinterpret_cast<const C<T>*>(p)
#define MP(T, p) reinterpret_cast<C<T>*>(p)

__host__ __device__ __forceinline__ int64_t ket_stride(int n, int D, int m) { return ipow(D, n - 1 - m); }
__host__ __device__ __forceinline__ int64_t dm_ket_stride(int n, int D, int m) { return ipow(D, 2 * (n - 1 - m)); }

// marginal chunks: as for reduced DMs, about 4096 traced indices per workgroup
int64_t marginal_chunks(int64_t NR) {
  int64_t c = (NR + 4095) / 4096;
  if (c < 1) c = 1;
  if (c > 256) c = 256;
  return c;
}

// p(bin) = sum over the traced modes of |psi|^2 (ket) or rho diagonal (DM);
// grid.x = bin (kept-mode digits, first kept mode most significant), grid.y =
// chunk of the traced index range.  FP64 accumulation.  The strides of each
// digit (diagonal strides on a DM) come precomputed from the host.
struct Marg {
  int32_t k, nt;
  int64_t kstride[32];  // kept mode t: stride of its digit in the state
  int64_t tstride[32];  // traced mode u (increasing mode order)
};

template <typename T>
__global__ void __launch_bounds__(256)
k_marginal(const C<T>* __restrict__ st, Marg mg, int D, int mixed, int64_t NR, int64_t chunk, double* partial) {
  __shared__ double sh[8];
  int64_t offA = 0;
  {
    int64_t o = blockIdx.x;
    for (int t = mg.k - 1; t >= 0; --t) {
      offA += (o % D) * mg.kstride[t];
      o /= D;
    }
  }
  double v = 0.0;
  const int64_t r0 = (int64_t)blockIdx.y * chunk;
  const int64_t r1 = (r0 + chunk < NR) ? r0 + chunk : NR;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    int64_t rr = r, off = 0;
    for (int u = mg.nt - 1; u >= 0; --u) {
      off += (rr % D) * mg.tstride[u];
      rr /= D;
    }
    const C<T> x = st[offA + off];
    v += mixed ? (double)x.re : (double)x.re * x.re + (double)x.im * x.im;
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * gridDim.y + blockIdx.y] = s;
}

__global__ void __launch_bounds__(256)
k_marginal_final(const double* __restrict__ partial, double* __restrict__ out, int64_t NO, int chunks) {
  GRID_STRIDE(o, NO) {
    double v = 0.0;
    for (int c = 0; c < chunks; ++c) v += partial[o * chunks + c];
    out[o] = v;
  }
}

// |0><x| on the measured modes, times `scale`: an output element whose
// measured digits (bra and ket for a DM) are all zero copies the input element
// with those digits set to the outcome (offset `shift`); every other output
// element is zero.
struct Proj {
  int32_t k;
  int32_t pad;
  int64_t stride[64];  // strides of the measured axes (bra and ket for a DM)
  int64_t shift;       // sum of outcome * stride over the measured axes
};

template <typename T>
__global__ void __launch_bounds__(256)
k_project_reset(const C<T>* __restrict__ in, C<T>* __restrict__ out, int64_t N, int D, Proj pj, T scale) {
  GRID_STRIDE(e, N) {
    bool zero = true;
    for (int t = 0; t < pj.k; ++t) zero = zero && ((e / pj.stride[t]) % D) == 0;
    C<T> v = mk<T>((T)0, (T)0);
    if (zero) {
      const C<T> x = in[e + pj.shift];
      v = mk<T>(x.re * scale, x.im * scale);
    }
    out[e] = v;
  }
}

// Homodyne density on q_b = -q_mag + b dq, dq = 2 q_mag / (num_bins - 1), of a
// single-mode DM rotated by R(-phi): with x = sqrt(m_omega) q and the
// normalised Hermite functions h_n(x) = H_n(x) e^{-x^2/2} / sqrt(2^n n!)
// (h_{n+1} = sqrt(2/(n+1)) x h_n - sqrt(n/(n+1)) h_{n-1}),
//   pdf(q) = sqrt(m_omega / pi) dq sum_{n,m} Re(rho_nm e^{-i (n-m) phi}) h_n h_m.
// One wave per workgroup, one bin per lane; h of every lane and Re(rho') in LDS.
constexpr int kPdfMaxD = SFAMD_MAX_CUTOFF;

template <typename T>
__global__ void __launch_bounds__(64)
k_homodyne_pdf(const C<T>* __restrict__ rho, int D, double phi, double q_mag, int64_t nbins, double mw,
               double* __restrict__ out) {
  __shared__ double rr[kPdfMaxD * kPdfMaxD];
  __shared__ double hs[kPdfMaxD * 64];
  for (int e = threadIdx.x; e < D * D; e += 64) {
    const int a = e / D, b = e % D;
    const C<T> x = rho[e];
    double sn, cs;
    sincos(-(double)(a - b) * phi, &sn, &cs);
    rr[e] = (double)x.re * cs - (double)x.im * sn;
  }
  __syncthreads();
  const double dq = nbins > 1 ? 2.0 * q_mag / (double)(nbins - 1) : 0.0;
  const double pref = sqrt(mw / M_PI) * dq;
  for (int64_t b0 = (int64_t)blockIdx.x * 64; b0 < nbins; b0 += (int64_t)gridDim.x * 64) {
    const int64_t b = b0 + threadIdx.x;
    const double q = -q_mag + (double)b * dq, x = sqrt(mw) * q;
    double* h = hs + threadIdx.x;  // h_n at h[64 n]
    double hm = exp(-0.5 * x * x), hn = sqrt(2.0) * x * hm;
    h[0] = hm;
    if (D > 1) h[64] = hn;
    for (int m = 1; m + 1 < D; ++m) {
      const double nx = sqrt(2.0 / (m + 1)) * x * hn - sqrt((double)m / (m + 1)) * hm;
      hm = hn;
      hn = nx;
      h[64 * (m + 1)] = nx;
    }
    double acc = 0.0;
    for (int a = 0; a < D; ++a) {
      const double ha = h[64 * a];
      double row = 0.5 * rr[a * D + a] * ha;
      for (int c = 0; c < a; ++c) row += rr[a * D + c] * h[64 * c];
      acc += 2.0 * row * ha;
    }
    if (b < nbins) out[b] = pref * acc;
  }
}

// Embedding of a prepared k-mode state: out = rest (x) sigma with sigma's t-th
// mode on register mode modes[t] and the complement's modes in increasing
// order (the product state of prepare_multimode after its axis permutation).
// DM layouts throughout; for a ket (k == n) out is the permuted ket.  Per
// register mode the host supplies the strides of its bra and ket digits in
// sigma (or in rest, the other factor's strides being zero).
struct Embed {
  int32_t n, nr, mixed, pad;
  int64_t sb[32], sk[32];  // strides in sigma of register mode m's bra / ket digit
  int64_t rb[32], rk[32];  // strides in rest
};

template <typename T>
__global__ void __launch_bounds__(256)
k_embed(const C<T>* __restrict__ rest, const C<T>* __restrict__ sigma, C<T>* __restrict__ out, int64_t N, int D,
        Embed em) {
  GRID_STRIDE(e, N) {
    int64_t r = e, si = 0, ri = 0;
    for (int m = em.n - 1; m >= 0; --m) {
      int j = 0;
      if (em.mixed) {
        j = (int)(r % D);
        r /= D;
      }
      const int i = (int)(r % D);
      r /= D;
      si += i * em.sb[m] + j * em.sk[m];
      ri += i * em.rb[m] + j * em.rk[m];
    }
    const C<T> s = sigma[si];
    out[e] = em.nr > 0 ? cmul(rest[ri], s) : s;
  }
}

}  // namespace

extern "C" int64_t sfamd_marginal_partials(int n, int D, int k) {
  if (n < 0 || D < 1 || k < 0 || k > n) return -1;
  return ipow(D, k) * marginal_chunks(ipow(D, n - k));
}

extern "C" int sfamd_marginal_probs(const void* state, int n, int D, int mixed, const int* modes, int k,
                                    double* partial, double* out_dev, int dtype, void* stream) {
  CHECK_ARG(state && partial && out_dev && (modes || k == 0) && D >= 1);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  ModeList ml;
  RET_IF(make_modelist(n, modes, k, &ml));
  Marg mg{};
  mg.k = k;
  mg.nt = n - k;
  auto dstride = [&](int m) { return mixed ? dm_ket_stride(n, D, m) * (D + 1) : ket_stride(n, D, m); };
  for (int t = 0; t < k; ++t) mg.kstride[t] = dstride(ml.kept[t]);
  for (int u = 0; u < n - k; ++u) mg.tstride[u] = dstride(ml.traced[u]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t NO = ipow(D, k), NR = ipow(D, n - k);
  const int64_t chunks = marginal_chunks(NR), chunk = (NR + chunks - 1) / chunks;
  CHECK_ARG(NO < (int64_t)1 << 31);
  dim3 grid((unsigned)NO, (unsigned)chunks);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_marginal<float>, grid, dim3(kBlock), 0, s, CP(float, state), mg, D, mixed, NR, chunk,
                       partial);
  else
    hipLaunchKernelGGL(k_marginal<double>, grid, dim3(kBlock), 0, s, CP(double, state), mg, D, mixed, NR, chunk,
                       partial);
  HIP_RET(hipGetLastError());
  hipLaunchKernelGGL(k_marginal_final, dim3(grid_for(NO)), dim3(kBlock), 0, s, partial, out_dev, NO, (int)chunks);
  return (int)hipGetLastError();
}

extern "C" int sfamd_project_reset(const void* in, void* out, int n, int D, int mixed, const int* modes,
                                   const int* outcome, int k, double scale, int dtype, void* stream) {
  CHECK_ARG(in && out && in != out && n >= 1 && D >= 1 && k >= 0 && k <= n && k <= 32);
  CHECK_ARG(k == 0 || (modes && outcome));
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  Proj pj{};
  pj.k = mixed ? 2 * k : k;
  pj.shift = 0;
  bool used[32] = {false};
  for (int t = 0; t < k; ++t) {
    const int m = modes[t], x = outcome[t];
    CHECK_ARG(m >= 0 && m < n && !used[m] && x >= 0 && x < D);
    used[m] = true;
    if (mixed) {
      const int64_t sk = dm_ket_stride(n, D, m);
      pj.stride[2 * t] = sk * D;
      pj.stride[2 * t + 1] = sk;
      pj.shift += x * (sk * D + sk);
    } else {
      pj.stride[t] = ket_stride(n, D, m);
      pj.shift += x * pj.stride[t];
    }
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t N = ipow(D, mixed ? 2 * n : n);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_project_reset<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, in), MP(float, out),
                       N, D, pj, (float)scale);
  else
    hipLaunchKernelGGL(k_project_reset<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, in),
                       MP(double, out), N, D, pj, scale);
  return (int)hipGetLastError();
}

extern "C" int sfamd_homodyne_pdf(const void* rho, int D, double phi, double q_mag, int num_bins, double hbar,
                                  double* out_dev, int dtype, void* stream) {
  CHECK_ARG(rho && out_dev && D >= 1 && D <= kPdfMaxD && num_bins >= 1 && hbar > 0.0 && q_mag >= 0.0);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  hipStream_t s = (hipStream_t)stream;
  int64_t g = ((int64_t)num_bins + 63) / 64;
  if (g > 4096) g = 4096;
  const double mw = 1.0 / hbar;
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_homodyne_pdf<float>, dim3((unsigned)g), dim3(64), 0, s, CP(float, rho), D, phi, q_mag,
                       (int64_t)num_bins, mw, out_dev);
  else
    hipLaunchKernelGGL(k_homodyne_pdf<double>, dim3((unsigned)g), dim3(64), 0, s, CP(double, rho), D, phi, q_mag,
                       (int64_t)num_bins, mw, out_dev);
  return (int)hipGetLastError();
}

extern "C" int sfamd_embed_modes(const void* rest, const void* sigma, void* out, int n, int D, const int* modes,
                                 int k, int mixed, int dtype, void* stream) {
  CHECK_ARG(sigma && out && modes && n >= 1 && n <= 32 && D >= 1 && k >= 1 && k <= n);
  CHECK_ARG(dtype == SFAMD_C64 || dtype == SFAMD_C128);
  CHECK_ARG(mixed || k == n);
  CHECK_ARG(rest || k == n);
  Embed em{};
  em.n = n;
  em.nr = n - k;
  em.mixed = mixed ? 1 : 0;
  int slot[32];
  for (int m = 0; m < n; ++m) slot[m] = -1000;
  for (int t = 0; t < k; ++t) {
    const int m = modes[t];
    CHECK_ARG(m >= 0 && m < n && slot[m] == -1000);
    slot[m] = t;
  }
  int u = 0;
  for (int m = 0; m < n; ++m) {
    if (slot[m] >= 0) {
      const int t = slot[m];
      em.sb[m] = mixed ? ipow(D, 2 * (k - 1 - t) + 1) : ipow(D, k - 1 - t);
      em.sk[m] = mixed ? ipow(D, 2 * (k - 1 - t)) : 0;
    } else {
      em.rb[m] = ipow(D, 2 * (em.nr - 1 - u) + 1);
      em.rk[m] = ipow(D, 2 * (em.nr - 1 - u));
      ++u;
    }
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t N = ipow(D, mixed ? 2 * n : n);
  if (dtype == SFAMD_C64)
    hipLaunchKernelGGL(k_embed<float>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(float, rest), CP(float, sigma),
                       MP(float, out), N, D, em);
  else
    hipLaunchKernelGGL(k_embed<double>, dim3(grid_for(N)), dim3(kBlock), 0, s, CP(double, rest), CP(double, sigma),
                       MP(double, out), N, D, em);
  return (int)hipGetLastError();
}
