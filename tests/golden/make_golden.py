"""Writes the golden fixtures of tests/golden (run from the repo root).

* gates_mpmath.npz -- displacement, squeezing and beamsplitter matrices from
  their closed forms evaluated with mpmath at 50 digits (independent of the
  recurrences of the oracle and of the HIP gate builder):
    D: <m|D(a)|n> = sqrt(n!/m!) a^(m-n) e^(-|a|^2/2) L_n^(m-n)(|a|^2) for m >= n
       (strawberryfields/ops.py:1525-1530), D(a)^dagger = D(-a) for m < n;
    S: <m|S(r)|n> of tests/backend/test_squeeze_operation.py:31-50 (real r),
       S(r e^{i phi}) = R(phi/2) S(r) R(-phi/2);
    BS: <N,M|B|n,m> = e^{i phi (n-N)} sum_k (-1)^(N-k)
        sqrt(C(n,k) C(m,N-k) C(N,k) C(M,n-k)) cos^(m-N+2k) sin^(n+N-2k)
       (the combinatorial factors are those of the reference data
       strawberryfields/backends/data/fock_beamsplitter_factors_*.npz);
    S2: <i,k|S2|j,l> = sum_q (-conj t)^q t^p / (q! p!) sqrt(i! j! k! l!)
        / ((j-q)! (l-q)!) sech^(j+l-2q+1), p = i - j + q, t = e^{i phi} tanh r
       (normal ordering of exp(r (e^{i phi} a1^dag a2^dag - h.c.))).
* c1.npz -- the 3-mode D = 8 circuit of strawberryfields_amd.circuits.config1:
  full complex128 ket and all Fock probabilities from the CPU oracle (which
  tests/test_oracle.py pins against the closed forms above and the reference's
  known-answer tests).
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

mp.mp.dps = 50


def disp_mp(r, phi, D):
    a = mp.mpf(r) * mp.expj(phi)
    U = np.zeros((D, D), dtype=complex)

    def elem(m, n, al):
        # m >= n
        x = abs(al) ** 2
        return (
            mp.sqrt(mp.factorial(n) / mp.factorial(m))
            * al ** (m - n)
            * mp.exp(-x / 2)
            * mp.laguerre(n, m - n, x)
        )

    for m in range(D):
        for n in range(D):
            if m >= n:
                v = elem(m, n, a)
            else:
                v = mp.conj(elem(n, m, -a))
            U[m, n] = complex(v)
    return U


def sq_real_mp(n, r, m):
    """<n|S(r)|m> for real r (test_squeeze_operation.py:31-50, exact)."""
    if (n + m) % 2:
        return mp.mpf(0)
    r = mp.mpf(r)
    s = mp.mpf(0)
    for k in range(m % 2, min(m, n) + 1, 2):
        s += (
            (-1) ** ((n - k) // 2)
            * mp.sqrt(mp.factorial(m) * mp.factorial(n))
            / (mp.factorial(k) * mp.factorial((m - k) // 2) * mp.factorial((n - k) // 2))
            * (mp.sinh(r) / 2) ** ((n + m - 2 * k) // 2)
            / mp.cosh(r) ** (mp.mpf(n + m + 1) / 2)
        )
    return s


def sq_mp(r, phi, D):
    U = np.zeros((D, D), dtype=complex)
    for m in range(D):
        for n in range(D):
            U[m, n] = complex(mp.expj(mp.mpf(phi) * (m - n) / 2) * sq_real_mp(m, r, n))
    return U


def bs_mp(theta, phi, D):
    th, ph = mp.mpf(theta), mp.mpf(phi)
    ct, st = mp.cos(th), mp.sin(th)
    mat = np.zeros((D, D, D, D), dtype=complex)  # [N, n, M, m]
    for N in range(D):
        for M in range(D):
            for n in range(D):
                m = N + M - n
                if m < 0 or m >= D:
                    continue
                s = mp.mpf(0)
                for k in range(0, min(n, N) + 1):
                    if N - k > m or n - k > M or N - k < 0 or n - k < 0:
                        continue
                    pref = mp.sqrt(mp.binomial(n, k) * mp.binomial(m, N - k) * mp.binomial(N, k) * mp.binomial(M, n - k))
                    e1, e2 = m - N + 2 * k, n + N - 2 * k
                    s += (-1) ** (N - k) * pref * ct**e1 * st**e2
                mat[N, n, M, m] = complex(mp.expj(ph * (n - N)) * s)
    return mat


def s2_mp(r, phi, D):
    """<i,k|S2|j,l> as mat[i, j, k, l] from the disentangled form
    S2 = exp(t a1^dag a2^dag) sech^(n1+n2+1) exp(-conj(t) a1 a2), t = e^{i phi} tanh r."""
    r = mp.mpf(r)
    t = mp.expj(phi) * mp.tanh(r)
    sech = 1 / mp.cosh(r)
    f = mp.factorial
    mat = np.zeros((D, D, D, D), dtype=complex)
    for i in range(D):
        for j in range(D):
            for k in range(D):
                l = k + j - i
                if l < 0 or l >= D:
                    continue
                s = mp.mpc(0)
                for q in range(max(0, j - i), min(j, l) + 1):
                    p = i - j + q
                    s += (
                        (-mp.conj(t)) ** q
                        * t**p
                        / (f(q) * f(p))
                        * mp.sqrt(f(j) * f(l) * f(i) * f(k))
                        / (f(j - q) * f(l - q))
                        * sech ** (j + l - 2 * q + 1)
                    )
                mat[i, j, k, l] = complex(s)
    return mat


CASES = {
    "D": [(0.3, 0.7, 6), (-0.45, 2.1, 12), (0.9, -1.3, 20), (0.05, 0.0, 15), (0.9, 0.7, 64)],
    "S": [(0.3, 0.7, 6), (-0.4, 2.1, 12), (0.8, -1.3, 20), (0.1, 0.0, 15), (0.6, -0.4, 64)],
    "BS": [(0.6, 0.4, 5), (1.1, -2.3, 8), (-0.35, 1.7, 12), (0.6, 0.4, 20)],
    "S2": [(0.3, 0.7, 6), (-0.4, 2.1, 8), (0.9, -1.3, 12), (0.6, 0.4, 20)],
}


def main():
    out = {}
    for kind, cases in CASES.items():
        for i, (p0, p1, D) in enumerate(cases):
            key = "%s_%d" % (kind, i)
            if kind == "D":
                m = disp_mp(p0, p1, D)
            elif kind == "S":
                m = sq_mp(p0, p1, D)
            elif kind == "S2":
                m = s2_mp(p0, p1, D)
            else:
                m = bs_mp(p0, p1, D)
            out[key] = m
            out[key + "_params"] = np.array([p0, p1, D], dtype=float)
    np.savez_compressed(os.path.join(HERE, "gates_mpmath.npz"), **out)

    from oracle import fock_oracle as fo
    from strawberryfields_amd import circuits

    c1 = circuits.config1()
    st, pure = fo.run_ops(c1["ops"], c1["n"], c1["D"], pure=True)
    probs = fo.all_fock_probs(st, c1["n"], pure)
    params = np.array([list(p) + [0.0] * (2 - len(p)) for _, _, p in c1["ops"]])
    np.savez_compressed(os.path.join(HERE, "c1.npz"), ket=st, probs=probs, params=params)
    print("wrote gates_mpmath.npz and c1.npz")


if __name__ == "__main__":
    main()
