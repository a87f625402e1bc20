"""Fock expansions of the single-mode states that preparations take as input
and that are not a column of one gate matrix: the displaced squeezed state and
the square-lattice GKP state.

They are D-element vectors built once per preparation on the host in FP64,
as the reference builds them (``fockbackend/ops.py:420-446, 518-596``); the
state they replace a mode with is then formed on the GPU by the library's
preparation kernels.  Coherent and squeezed states are the first columns of
the D and S gate matrices the library builds, which hold the exact matrix
elements, so they need nothing here.
"""
import math

import numpy as np


def _coherent(r, phi, D):
    """e^{-|a|^2/2} a^n / sqrt(n!), a = r e^{i phi} (fockbackend/ops.py:392-402)."""
    a = complex(r * np.exp(1j * phi))
    return np.array([np.exp(-abs(a) ** 2 / 2) * a**n / math.exp(0.5 * math.lgamma(n + 1)) for n in range(D)],
                    dtype=np.complex128)


def _squeezed(r, theta, D):
    """sqrt(sech r) sqrt((2k)!) / (2^k k!) (-e^{i theta} tanh r)^k at n = 2k
    (fockbackend/ops.py:405-416)."""
    v = np.zeros(D, dtype=np.complex128)
    z = -np.exp(1j * theta) * np.tanh(r)
    for n in range(0, D, 2):
        k = n // 2
        c = math.exp(0.5 * math.lgamma(2 * k + 1) - k * math.log(2) - math.lgamma(k + 1))
        v[n] = c * z**k
    return np.sqrt(1 / np.cosh(r)) * v


def _hermite(z, D):
    """Physicists' Hermite polynomials H_0..H_{D-1} at complex z (three-term
    recurrence H_{n+1} = 2 z H_n - 2 n H_{n-1})."""
    h = np.zeros(D, dtype=np.complex128)
    h[0] = 1.0
    if D > 1:
        h[1] = 2 * z
    for n in range(1, D - 1):
        h[n + 1] = 2 * z * h[n] - 2 * n * h[n - 1]
    return h


def displaced_squeezed(r_d, phi_d, r_s, phi_s, D):
    """D(a) S(r_s e^{i phi_s}) |0>, a = r_d e^{i phi_d}, truncated to D levels,
    from the closed form of its amplitudes (Yuen, PRA 13, 2226 (1976))

        <n|a, z> = N (e^{i phi_s} tanh r_s / 2)^{n/2} / sqrt(n! cosh r_s)
                   H_n(gamma / sqrt(e^{i phi_s} sinh 2 r_s)),
        gamma = a cosh r_s + conj(a) e^{i phi_s} sinh r_s,
        N = exp(-|a|^2/2 - conj(a)^2 e^{i phi_s} tanh(r_s) / 2),

    with the branches, the zero cases (coherent state when r_s ~ 0, squeezed
    vacuum when r_d ~ 0, np.allclose's default tolerance) and the 1e-10 added
    under the square root exactly as fockbackend/ops.py:420-446 has them, so the
    amplitudes match the reference's; unlike D[:, :D] @ S[:D, 0] they carry no
    truncation error from the finite cutoff."""
    if np.allclose(r_s, 0.0):
        return _coherent(r_d, phi_d, D)
    if np.allclose(r_d, 0.0):
        return _squeezed(r_s, phi_s, D)
    ph = np.exp(1j * phi_s)
    ch, sh, th = np.cosh(r_s), np.sinh(r_s), np.tanh(r_s)
    a = r_d * np.exp(1j * phi_d)
    gamma = a * ch + np.conj(a) * ph * sh
    arg = gamma / np.sqrt(ph * np.sinh(2 * r_s) + 1e-10)
    norm = np.exp(-0.5 * abs(a) ** 2 - 0.5 * np.conj(a) ** 2 * ph * th)
    n = np.arange(D)
    coeff = np.power(0.5 * ph * th, n / 2) / np.sqrt(np.array([math.exp(math.lgamma(k + 1)) for k in n]) * ch)
    return norm * coeff * _hermite(arg, D)


def square_gkp(theta, phi, epsilon, ampl_cutoff, D):
    """Normalised Fock expansion of cos(theta/2)|0>_gkp + e^{-i phi} sin(theta/2)|1>_gkp
    on the square lattice with finite-energy parameter epsilon
    (fockbackend/ops.py:518-596): each logical state mu is the sum over teeth t
    of exp(-pi tanh(eps) (mu + 2t)^2 / 2) D(sqrt(pi/2) (2t + mu) / cosh eps)
    S(r)|0>, r = -log(tanh eps) / 2, over |t| <= ceil(sqrt(-log(ampl_cutoff) /
    (4 pi tanh eps)))."""
    zmax = int(np.ceil(np.sqrt(-0.25 / np.pi * np.log(ampl_cutoff) / np.tanh(epsilon))))
    r = -0.5 * np.log(np.tanh(epsilon))

    def logical(mu):
        ket = np.zeros(D, dtype=np.complex128)
        for t in range(-zmax, zmax + 1):
            c = np.exp(-0.5 * np.pi * np.tanh(epsilon) * (mu + 2 * t) ** 2)
            alpha = np.sqrt(0.5 * np.pi) * (2 * t + mu) / np.cosh(epsilon)
            ket = ket + c * displaced_squeezed(alpha, 0.0, r, 0.0, D)
        return ket

    ket = np.cos(theta / 2) * logical(0) + np.sin(theta / 2) * np.exp(-1j * phi) * logical(1)
    return ket / np.linalg.norm(ket)
