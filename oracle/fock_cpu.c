/*
 * fock_cpu.c -- native part of the CPU oracle (test and baseline
 * infrastructure only; nothing in strawberryfields_amd/ links or calls it).
 *
 * Restates the numba kernel of the reference fock backend's beamsplitter,
 * Circuit._apply_two_mode_passive (strawberryfields/backends/fockbackend/
 * circuit.py:306-335), which numba compiles to native code:
 *
 *   ret[i, j, ...] = sum_{k = max(1+i+j-D, 0)}^{min(i+j, D-1)}
 *                    mat[i, k, j, i+j-k] * state[k, i+j-k, ...]
 *
 * on a state whose two gate axes lead (the reference transposes them to the
 * front, circuit.py:236-261), and the same loop for the two-mode squeezer's
 * difference-conserving rule (Circuit._apply_S2, circuit.py:337-365):
 *
 *   ret[i, k, ...] = sum_{j} mat[i, j, k, k+j-i] * state[j, k+j-i, ...].
 *
 * Complex128 values are interleaved (re, im) doubles, numpy's layout; `rest`
 * is the number of trailing elements per (i, j) pair.  Single-threaded, as
 * numba's nopython loop is.  Built with gcc -O3 by oracle/build_oracle.py.
 */
#include <stddef.h>
#include <string.h>

static inline void cmac_row(double* r, double mre, double mim, const double* s, long rest) {
  for (long t = 0; t < rest; ++t) {
    const double sre = s[2 * t], sim = s[2 * t + 1];
    r[2 * t] += mre * sre - mim * sim;
    r[2 * t + 1] += mre * sim + mim * sre;
  }
}

/* mat: [D][D][D][D] = mat[out1][in1][out2][in2]; state, ret: [D][D][rest] */
void fock_two_mode_passive(const double* mat, const double* state, double* ret, int D, long rest) {
  const long D2 = (long)D * D, D3 = D2 * D;
  memset(ret, 0, sizeof(double) * 2 * D2 * rest);
  for (int i = 0; i < D; ++i) {
    for (int j = 0; j < D; ++j) {
      double* r = ret + 2 * ((long)i * D + j) * rest;
      const int lo = 1 + i + j - D > 0 ? 1 + i + j - D : 0;
      const int hi = i + j < D - 1 ? i + j : D - 1;
      for (int k = lo; k <= hi; ++k) {
        const int l = i + j - k;
        const double* m = mat + 2 * (i * D3 + k * D2 + (long)j * D + l);
        cmac_row(r, m[0], m[1], state + 2 * ((long)k * D + l) * rest, rest);
      }
    }
  }
}

/* mat: [D][D][D][D] = mat[out1][in1][out2][in2]; state, ret: [D][D][rest] */
void fock_two_mode_diffcons(const double* mat, const double* state, double* ret, int D, long rest) {
  const long D2 = (long)D * D, D3 = D2 * D;
  memset(ret, 0, sizeof(double) * 2 * D2 * rest);
  for (int i = 0; i < D; ++i) {
    for (int j = 0; j < D; ++j) {
      const int lo = i - j > 0 ? i - j : 0;
      const int hi = D + (i - j < 0 ? i - j : 0);
      for (int k = lo; k < hi; ++k) {
        const int l = k + j - i;
        double* r = ret + 2 * ((long)i * D + k) * rest;
        const double* m = mat + 2 * (i * D3 + (long)j * D2 + (long)k * D + l);
        cmac_row(r, m[0], m[1], state + 2 * ((long)j * D + l) * rest, rest);
      }
    }
  }
}
