"""CPU oracle: a numpy complex128 restatement of Strawberry Fields' ``fock`` backend.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP library.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import it; the product path (``strawberryfields_amd``) never does.

What it restates (reference files, read as text, never imported):

* gate matrices -- ``strawberryfields/backends/fockbackend/ops.py``:
  ``phase`` 309-314, ``kerr`` 274-281, ``displacement`` 219-235, ``squeezing``
  238-254 and ``beamsplitter`` 318-329 (the latter three call
  ``thewalrus.fock_gradients`` 0.22.0, not vendored; its published algorithm is
  a ladder-operator recurrence, restated below from the Heisenberg action of
  each gate), ``lossChannel`` Kraus set 471-490, ``a`` 208-216, ``proj`` 346-353;
  the loss superoperator of ``tfbackend/ops.py:288-315``; ``two_mode_squeeze``
  257-271 and ``mzgate`` 332-343 (thewalrus, not vendored: S2 by its Gaussian
  Hermite recurrence, MZ as the product of its documented factors),
  ``cross_kerr`` 284-294, ``cubicPhase`` 296-306.
* state updates -- ``strawberryfields/backends/fockbackend/circuit.py``:
  ``apply_gate_BLAS`` 118-217 (the slice loop with ``np.dot``), the
  selection-rule beamsplitter loop ``_apply_two_mode_passive`` 306-335, the
  difference-conserving ``_apply_S2`` 337-365 (both numba in the reference;
  in C here, ``oracle/fock_cpu.c``, for the faithful mode),
  ``_apply_channel`` 65-87, ``prepare_multimode`` 393-473, ``norm`` 367-371,
  ``is_vacuum`` 600-609; ``ops.mix`` 110-120, ``partial_trace`` 144-157.
* readout -- ``strawberryfields/backends/states.py`` ``BaseFockState``:
  ``dm`` 544-554, ``trace`` 556-578, ``all_fock_probs`` 580-608,
  ``reduced_dm`` 613-642, ``fock_prob`` 644-655, ``mean_photon`` 657-663,
  ``diagonal_expectation`` / ``number_expectation`` 947-980.

Deliberate difference: for a beamsplitter with ``mode1 > mode2`` the fock
backend's transposes (``circuit.py:236-261``) place the wrong axes in front;
this oracle (like the TF backend, ``tfbackend/ops.py:800-811``, and the
``BaseBackend.beamsplitter`` contract) applies the matrix's first mode pair to
``mode1``.  For ``mode1 < mode2`` both agree.

Pinning (see tests/test_oracle.py): the D, S, BS and S2 matrices are checked
against 50-digit mpmath closed forms (tests/golden/gates_mpmath.npz), S2 also
against expm of its generator on a 48-level basis, the BS combinatorial
prefactors against the reference's own ``fock_beamsplitter_factors_{5,10}.npz``,
MZ against its documented mode unitary, special cases and gate decomposition,
V against the eigendecomposition of x^3, the Kraus loss against the
superoperator identity, the slice-loop and selection-rule gate applications
against a brute-force Kronecker operator, and the restated reference
known-answer tests (TMSV amplitudes, MZ on |1,1>, cross-Kerr and cubic phase on
uniform states, coherent and squeezed states, Hong-Ou-Mandel, loss).
"""
from itertools import product
import ctypes
import os
import string

import numpy as np

DEF_TYPE = np.complex128
_IDX = string.ascii_lowercase

# The faithful mode's selection-rule loops (numba-compiled native code in the
# reference) run from oracle/libfock_cpu.so (oracle/fock_cpu.c, gcc -O3) when
# it is built, else as Python loops with the same arithmetic.
_NATIVE = {}


def native_lib():
    """The oracle's C loops (ctypes) or None when oracle/libfock_cpu.so is absent."""
    if "lib" not in _NATIVE:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfock_cpu.so")
        lib = None
        if os.path.exists(path):
            lib = ctypes.CDLL(path)
            for fn in (lib.fock_two_mode_passive, lib.fock_two_mode_diffcons):
                fn.restype = None
                fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_long]
        _NATIVE["lib"] = lib
    return _NATIVE["lib"]


def _native_rule(name, mat, state, D):
    lib = native_lib()
    if lib is None:
        return None
    m = np.ascontiguousarray(mat, dtype=DEF_TYPE)
    st = np.ascontiguousarray(state, dtype=DEF_TYPE)
    ret = np.empty_like(st)
    rest = st.size // (D * D)
    getattr(lib, name)(m.ctypes.data, st.ctypes.data, ret.ctypes.data, D, rest)
    return ret


# ----------------------------------------------------------------- matrices
def phase(theta, D):
    """R(theta) = diag(e^{i n theta})  (fockbackend/ops.py:309-314)."""
    return np.diag(np.exp(1j * theta * np.arange(D))).astype(DEF_TYPE)


def kerr(kappa, D):
    """K(kappa) = diag(e^{i kappa n^2})  (fockbackend/ops.py:274-281)."""
    n = np.arange(D)
    return np.diag(np.exp(1j * kappa * n**2)).astype(DEF_TYPE)


def displacement(r, phi, D):
    """<m|D(r e^{i phi})|n>, truncated, not renormalised.

    Recurrence from D a^dag = (a^dag - alpha*) D and a D = D (a + alpha):
    D[m,0] = alpha/sqrt(m) D[m-1,0], D[0,0] = e^{-r^2/2},
    D[m,n] = (sqrt(m) D[m-1,n-1] - alpha* D[m,n-1]) / sqrt(n).
    """
    alpha = r * np.exp(1j * phi)
    U = np.zeros((D, D), dtype=DEF_TYPE)
    U[0, 0] = np.exp(-0.5 * r**2)
    for m in range(1, D):
        U[m, 0] = alpha / np.sqrt(m) * U[m - 1, 0]
    for n in range(1, D):
        for m in range(D):
            v = -np.conj(alpha) * U[m, n - 1]
            if m > 0:
                v += np.sqrt(m) * U[m - 1, n - 1]
            U[m, n] = v / np.sqrt(n)
    return U


def squeezing(r, phi, D):
    """<m|S(r e^{i phi})|n>, S(z) = exp((z* a^2 - z a^dag^2)/2), truncated.

    Recurrence from S a^dag S^dag = a^dag cosh r + a e^{-i phi} sinh r:
    S[m,0] = -e^{i phi} tanh r sqrt((m-1)/m) S[m-2,0], S[0,0] = sqrt(sech r),
    S[m,n] = (sqrt(n-1) e^{-i phi} tanh r S[m,n-2] + sqrt(m) sech r S[m-1,n-1]) / sqrt(n).
    """
    th, sech = np.tanh(r), 1.0 / np.cosh(r)
    e = np.exp(1j * phi)
    U = np.zeros((D, D), dtype=DEF_TYPE)
    U[0, 0] = np.sqrt(sech)
    for m in range(2, D, 2):
        U[m, 0] = -e * th * np.sqrt((m - 1) / m) * U[m - 2, 0]
    for n in range(1, D):
        for m in range(D):
            if (m + n) % 2:
                continue
            v = 0.0j
            if n >= 2:
                v += np.sqrt(n - 1) * np.conj(e) * th * U[m, n - 2]
            if m >= 1:
                v += np.sqrt(m) * sech * U[m - 1, n - 1]
            U[m, n] = v / np.sqrt(n)
    return U


def beamsplitter(theta, phi, D):
    """BS(theta, phi) as mat[out1, in1, out2, in2] (fockbackend/ops.py:318-329 layout).

    BS = exp(theta (e^{i phi} a1 a2^dag - e^{-i phi} a1^dag a2)).  Recurrence from
    BS a1^dag = (a1^dag cos + e^{i phi} sin a2^dag) BS and
    BS a2^dag = (a2^dag cos - e^{-i phi} sin a1^dag) BS, photon-number block by block.
    """
    ct, st = np.cos(theta), np.sin(theta)
    e = np.exp(1j * phi)
    Z = np.zeros((D, D, D, D), dtype=DEF_TYPE)  # Z[N, M, n, m] = <N,M|B|n,m>
    Z[0, 0, 0, 0] = 1.0
    for s in range(1, 2 * D - 1):
        lo, hi = max(0, s - D + 1), min(s, D - 1)
        for N in range(lo, hi + 1):
            M = s - N
            for n in range(lo, hi + 1):
                m = s - n
                v = 0.0j
                if m > 0:
                    if M >= 1:
                        v += ct * np.sqrt(M) * Z[N, M - 1, n, m - 1]
                    if N >= 1:
                        v -= np.conj(e) * st * np.sqrt(N) * Z[N - 1, M, n, m - 1]
                    v /= np.sqrt(m)
                else:
                    if N >= 1:
                        v += ct * np.sqrt(N) * Z[N - 1, M, n - 1, 0]
                    if M >= 1:
                        v += e * st * np.sqrt(M) * Z[N, M - 1, n - 1, 0]
                    v /= np.sqrt(n)
                Z[N, M, n, m] = v
    return Z.transpose(0, 2, 1, 3).copy()


def annihilation(D):
    """a (fockbackend/ops.py:208-216)."""
    a = np.zeros((D, D), dtype=DEF_TYPE)
    for i in range(1, D):
        a[i - 1, i] = np.sqrt(i)
    return a


def loss_kraus(T, D):
    """Kraus operators of N(T)  (fockbackend/ops.py:471-490)."""
    if T == 0:
        out = []
        for i in range(D):
            P = np.zeros((D, D), dtype=DEF_TYPE)
            P[0, i] = 1.0
            out.append(P)
        return out
    a = annihilation(D)
    TT = np.diag([T ** (i / 2) for i in range(D)]).astype(DEF_TYPE)
    ops = []
    for n in range(D):
        an = np.linalg.matrix_power(a, n)
        fac = np.prod(np.arange(1, n + 1, dtype=float))
        ops.append(((1 - T) / T) ** (n / 2) * (an / np.sqrt(fac)) @ TT)
    return ops


def loss_superop(T, D):
    """S[a, b, c, d]: |b><c| -> sum S |a><d|  (tfbackend/ops.py:288-315)."""
    from math import factorial

    S = np.zeros((D, D, D, D))
    for b in range(D):
        for c in range(D):
            for l in range(min(b, c) + 1):
                coef = (
                    (1 - T) ** l
                    * T ** ((b + c) / 2 - l)
                    * np.sqrt(factorial(b) * factorial(c) / (factorial(b - l) * factorial(c - l)))
                    / factorial(l)
                )
                S[b - l, b, c, c - l] += coef
    return S


def two_mode_squeezing(r, phi, D):
    """S2(r, phi) as mat[out1, in1, out2, in2] (fockbackend/ops.py:257-271 layout).

    S2 = exp(r (e^{i phi} a1^dag a2^dag - e^{-i phi} a1 a2)), the convention the
    reference's known-answer test fixes (test_twomode_squeezing_operation.py:30-42:
    S2|0,0> = sech r sum_k (e^{i phi} tanh r)^k |k,k>).  thewalrus 0.22.0
    ``two_mode_squeezing`` is not vendored; its published method is a
    Gaussian-kernel (Hermite) recurrence in which every index step down-shifts
    the others, restated here from the Heisenberg action
    S2^dag a1 S2 = cosh r a1 + e^{i phi} sinh r a2^dag (and 1 <-> 2):
      sqrt(N) G[N, M, n, m] = sech r sqrt(n) G[N-1, M, n-1, m] + e^{i phi} tanh r sqrt(M) G[N-1, M-1, n, m]
      sqrt(M) G[N, M, n, m] = sech r sqrt(m) G[N, M-1, n, m-1] + e^{i phi} tanh r sqrt(N) G[N-1, M-1, n, m]
      sqrt(n) G[N, M, n, m] = sech r sqrt(N) G[N-1, M, n-1, m] - e^{-i phi} tanh r sqrt(m) G[N, M, n-1, m-1]
    with G[N, M, n, m] = <N,M|S2|n,m>, G[0,0,0,0] = sech r, and entries zero
    unless N - M = n - m.  Every coefficient is at most 1 in magnitude and no
    entry outside the D^4 block is read, so the truncation is exact.
    """
    sech, th = 1.0 / np.cosh(r), np.tanh(r)
    e = np.exp(1j * phi)
    sq = np.sqrt(np.arange(D + 1, dtype=float))
    G = np.zeros((D, D, D, D), dtype=DEF_TYPE)
    G[0, 0, 0, 0] = sech
    for N in range(D):
        for M in range(D):
            for n in range(D):
                m = n - N + M
                if m < 0 or m >= D or N + M + n + m == 0:
                    continue
                if N > 0:
                    v = (sech * sq[n] * G[N - 1, M, n - 1, m] if n > 0 else 0.0) + (
                        e * th * sq[M] * G[N - 1, M - 1, n, m] if M > 0 else 0.0
                    )
                    v /= sq[N]
                elif M > 0:  # N = 0
                    v = (sech * sq[m] * G[0, M - 1, n, m - 1] if m > 0 else 0.0) / sq[M]
                else:  # N = M = 0, so n = m > 0
                    v = -np.conj(e) * th * sq[m] * G[0, 0, n - 1, m - 1] / sq[n]
                G[N, M, n, m] = v
    return G.transpose(0, 2, 1, 3).copy()


_BS50 = {}  # BS(pi/4, pi/2) at cutoff L as an (L^2, L^2) operator, by L


def mzgate(phi_in, phi_ex, D):
    """MZ(phi_in, phi_ex) as mat[out1, in1, out2, in2] (fockbackend/ops.py:332-343 layout).

    The operator of strawberryfields/ops.py:1951-1990,
    MZ = BS(pi/4, pi/2) (R(phi_in) x I) BS(pi/4, pi/2) (R(phi_ex) x I), whose mode
    unitary is 1/2 [[(e^{i phi_in} - 1) e^{i phi_ex}, i (1 + e^{i phi_in})],
    [i (1 + e^{i phi_in}) e^{i phi_ex}, 1 - e^{i phi_in}]] (thewalrus 0.22.0
    ``mzgate`` is not vendored).  The four factors are multiplied as Fock
    operators at cutoff 2D - 1: every factor conserves n1 + n2 <= 2D - 2 on the
    kept block, so the product is exact there before it is truncated to D.
    """
    L = max(2 * D - 1, 1)
    if L not in _BS50:
        _BS50[L] = beamsplitter(np.pi / 4, np.pi / 2, L).transpose(0, 2, 1, 3).reshape(L * L, L * L)
    B = _BS50[L]
    ones = np.ones(L)
    Rin = np.kron(np.exp(1j * phi_in * np.arange(L)), ones)
    Rex = np.kron(np.exp(1j * phi_ex * np.arange(L)), ones)
    M = (B * Rin[None, :]) @ (B * Rex[None, :])
    M = M.reshape(L, L, L, L)[:D, :D, :D, :D]  # [out1, out2, in1, in2]
    return M.transpose(0, 2, 1, 3).copy()


def cross_kerr(kappa, D):
    """CK(kappa) = exp(i kappa n1 n2) as mat[out1, in1, out2, in2] (fockbackend/ops.py:284-294)."""
    n = np.arange(D)
    mat = np.zeros((D, D, D, D), dtype=DEF_TYPE)
    mat[n[:, None], n[:, None], n[None, :], n[None, :]] = np.exp(1j * kappa * np.outer(n, n))
    return mat


def cubic_phase(gamma, D, hbar=2.0):
    """V(gamma) = expm(i gamma / (3 hbar) x^3), x = sqrt(hbar/2) (a + a^dag) on the
    truncated basis (fockbackend/ops.py:297-306; scipy.linalg.expm as there)."""
    from scipy.linalg import expm

    a = annihilation(D)
    x = (a + a.conj().T) * np.sqrt(hbar / 2)
    return expm(1j * gamma / (3 * hbar) * (x @ x @ x)).astype(DEF_TYPE)


# ----------------------------------------------------------------- states
def mix(state, n):
    """rho[i0,j0,i1,j1,...] = psi[i...] psi*[j...]  (fockbackend/ops.py:110-120)."""
    left = [_IDX[i] for i in range(0, 2 * n, 2)]
    right = [_IDX[i] for i in range(1, 2 * n, 2)]
    eq = "".join(left) + "," + "".join(right) + "->" + _IDX[: 2 * n]
    return np.einsum(eq, state, state.conj())


def partial_trace(state, n, modes):
    """Trace out `modes` of a DM  (fockbackend/ops.py:144-157)."""
    left = [
        _IDX[2 * i] + _IDX[2 * i] if i in modes else _IDX[2 * i : 2 * i + 2] for i in range(n)
    ]
    out = ["" if i in modes else _IDX[2 * i : 2 * i + 2] for i in range(n)]
    return np.einsum("".join(left) + "->" + "".join(out), state)


def vacuum(n, D, pure=True):
    st = np.zeros([D] * (n if pure else 2 * n), dtype=DEF_TYPE)
    st.ravel()[0] = 1.0
    return st


class Circuit:
    """Restatement of fockbackend.circuit.Circuit (the parts on the gate path).

    ``faithful=True`` keeps the reference's loop structure (the slice loop with
    ``np.dot`` of apply_gate_BLAS and the selection-rule beamsplitter loop); it
    is the CPU baseline.  ``faithful=False`` contracts with ``tensordot`` /
    vectorised selection rules and is used to keep the tests fast.  Both are
    checked against each other and against a brute-force Kronecker operator.
    """

    def __init__(self, num, trunc, pure=True, faithful=False):
        if num < 0:
            raise ValueError("Number of modes must be non-negative -- got {}".format(num))
        if trunc <= 0:
            raise ValueError("Truncation must be positive -- got {}".format(trunc))
        self._num_modes = num
        self._trunc = trunc
        self._pure = pure
        self._hbar = 2.0
        self.faithful = faithful
        self.reset(pure)

    # -- state management
    def reset(self, pure=None, cutoff_dim=None, num_subsystems=None):
        if pure is not None:
            self._pure = pure
        if cutoff_dim is not None:
            self._trunc = cutoff_dim
        if num_subsystems is not None:
            self._num_modes = num_subsystems
        self._state = vacuum(self._num_modes, self._trunc, self._pure)

    def get_state(self):
        return self._state, self._pure

    # -- generic single-mode gate (circuit.py:118-217)
    def apply_gate_BLAS(self, mat, modes, state=None, pure=None, n=None):
        state = self._state if state is None else state
        pure = self._pure if pure is None else pure
        n = self._num_modes if n is None else n
        D = self._trunc
        size = len(modes)
        dim = D**size
        stshape = [D] * size
        tl = [2 * i for i in range(size)] + [2 * i + 1 for i in range(size)]
        matview = np.transpose(mat, tl).reshape((dim, dim))
        diag = np.all(matview == np.diag(np.diagonal(matview)))
        if pure:
            if n == 1:
                return mat.diagonal() * state if diag else mat @ state
            tlist = [i for i in range(n) if i not in modes] + list(modes)
            view = np.transpose(state, tlist)
            if self.faithful:
                ret = np.zeros([D] * n, dtype=DEF_TYPE)
                for i in product(*([range(D)] * (n - size))):
                    if diag:
                        ret[i] = np.multiply(matview.diagonal(), view[i].ravel()).reshape(stshape)
                    else:
                        ret[i] = np.dot(matview, view[i].ravel()).reshape(stshape)
            else:
                flat = view.reshape(-1, dim)
                ret = (flat @ matview.T).reshape([D] * n)
            return np.transpose(ret, np.argsort(tlist))
        if n == 1:
            if diag:
                d = mat.diagonal().reshape(-1, 1)
                return d * (state * d.conj().T)
            return mat @ state @ mat.conj().T
        tlist = [i for i in range(2 * n) if i // 2 not in modes]
        tlist = tlist + [2 * i for i in modes] + [2 * i + 1 for i in modes]
        view = np.transpose(state, tlist)
        if self.faithful:
            ret = np.zeros([D] * (2 * n), dtype=DEF_TYPE)
            for i in product(*([range(D)] * ((n - size) * 2))):
                if diag:
                    d = matview.diagonal().reshape(-1, 1)
                    ret[i] = (d * (view[i].reshape((dim, dim)) * d.conj().T)).reshape(stshape + stshape)
                else:
                    ret[i] = (matview @ view[i].reshape((dim, dim)) @ matview.conj().T).reshape(
                        stshape + stshape
                    )
        else:
            flat = view.reshape(-1, dim, dim)
            ret = np.matmul(np.matmul(matview, flat), matview.conj().T).reshape([D] * (2 * n))
        return np.transpose(ret, np.argsort(tlist))

    # -- beamsplitter (circuit.py:219-335), first matrix pair on modes[0]
    @staticmethod
    def _two_mode_passive(mat, state, D, faithful):
        """ret[i, j, ...] = sum_k mat[i, k, j, i+j-k] state[k, i+j-k, ...]  (circuit.py:306-335)."""
        if faithful:
            ret = _native_rule("fock_two_mode_passive", mat, state, D)
            if ret is not None:
                return ret
        ret = np.zeros_like(state, dtype=DEF_TYPE)
        if faithful:
            for i in range(D):
                for j in range(D):
                    for k in range(max(1 + i + j - D, 0), min(i + j, D - 1) + 1):
                        ret[i, j] += mat[i, k, j, i + j - k] * state[k, i + j - k]
            return ret
        for s in range(2 * D - 1):
            lo, hi = max(0, s - D + 1), min(s, D - 1)
            ks = np.arange(lo, hi + 1)
            blk = mat[ks[:, None], ks[None, :], (s - ks)[:, None], (s - ks)[None, :]]
            sub = state[ks, s - ks]
            ret[ks, s - ks] = np.tensordot(blk, sub, axes=(1, 0))
        return ret

    # -- two-mode squeezing (circuit.py:337-365), first matrix pair on modes[0]
    @staticmethod
    def _two_mode_diffcons(mat, state, D, faithful):
        """ret[i, k, ...] = sum_j mat[i, j, k, k+j-i] state[j, k+j-i, ...]  (circuit.py:337-365)."""
        if faithful:
            ret = _native_rule("fock_two_mode_diffcons", mat, state, D)
            if ret is not None:
                return ret
        ret = np.zeros_like(state, dtype=DEF_TYPE)
        if faithful:
            for i in range(D):
                for j in range(D):
                    for k in range(max(i - j, 0), D + min(i - j, 0)):
                        ret[i, k] += mat[i, j, k, k + j - i] * state[j, k + j - i]
            return ret
        for d in range(-(D - 1), D):
            # out1 - out2 = in1 - in2 = d
            ks = np.arange(max(0, d), min(D, D + d))
            blk = mat[ks[:, None], ks[None, :], (ks - d)[:, None], (ks - d)[None, :]]
            ret[ks, ks - d] = np.tensordot(blk, state[ks, ks - d], axes=(1, 0))
        return ret

    def apply_twomode_gate(self, mat, modes, gate="BSgate"):
        """circuit.py:219-304: BSgate / MZgate by the passive selection rule,
        S2gate by the difference-conserving one; bra side with mat, ket side
        with conj(mat)."""
        D = self._trunc
        st = self._state
        rule = {"BSgate": self._two_mode_passive, "MZgate": self._two_mode_passive}.get(gate)
        if gate == "S2gate":
            rule = self._two_mode_diffcons
        if rule is None:
            raise NotImplementedError(gate)
        if self._pure:
            axes = [modes[0], modes[1]]
            v = np.moveaxis(st, axes, [0, 1])
            v = rule(mat, v, D, self.faithful)
            return np.moveaxis(v, [0, 1], axes)
        bra = [2 * modes[0], 2 * modes[1]]
        ket = [2 * modes[0] + 1, 2 * modes[1] + 1]
        v = np.moveaxis(st, bra, [0, 1])
        v = rule(mat, v, D, self.faithful)
        v = np.moveaxis(v, [0, 1], bra)
        v = np.moveaxis(v, ket, [0, 1])
        v = rule(mat.conj(), v, D, self.faithful)
        return np.moveaxis(v, [0, 1], ket)

    # -- channels (circuit.py:65-87)
    def _apply_channel(self, kraus_ops, modes):
        if self._pure:
            self._state = mix(self._state, self._num_modes)
            self._pure = False
        states = [self.apply_gate_BLAS(k, modes, pure=False) for k in kraus_ops]
        self._state = sum(states)

    # -- gates (circuit.py:537-621)
    def phase_shift(self, theta, mode):
        self._state = self.apply_gate_BLAS(phase(theta, self._trunc), [mode])

    def kerr_interaction(self, kappa, mode):
        self._state = self.apply_gate_BLAS(kerr(kappa, self._trunc), [mode])

    def displacement(self, r, phi, mode):
        self._state = self.apply_gate_BLAS(displacement(r, phi, self._trunc), [mode])

    def squeeze(self, r, phi, mode):
        self._state = self.apply_gate_BLAS(squeezing(r, phi, self._trunc), [mode])

    def beamsplitter(self, theta, phi, mode1, mode2):
        self._state = self.apply_twomode_gate(beamsplitter(theta, phi, self._trunc), [mode1, mode2])

    def loss(self, T, mode):
        self._apply_channel(loss_kraus(T, self._trunc), [mode])

    def mzgate(self, phi_in, phi_ex, mode1, mode2):
        self._state = self.apply_twomode_gate(mzgate(phi_in, phi_ex, self._trunc), [mode1, mode2], "MZgate")

    def two_mode_squeeze(self, r, phi, mode1, mode2):
        self._state = self.apply_twomode_gate(two_mode_squeezing(r, phi, self._trunc), [mode1, mode2], "S2gate")

    def cross_kerr_interaction(self, kappa, mode1, mode2):
        self._state = self.apply_gate_BLAS(cross_kerr(kappa, self._trunc), [mode1, mode2])

    def cubic_phase_shift(self, gamma, mode):
        self._state = self.apply_gate_BLAS(cubic_phase(gamma, self._trunc), [mode])

    # -- preparations (circuit.py:393-498), single mode
    def prepare(self, st, mode):
        """Replace `mode` by the single-mode ket or DM `st`."""
        D, n = self._trunc, self._num_modes
        if n == 1:
            self._state = np.array(st, dtype=DEF_TYPE)
            self._pure = st.ndim == 1
            return
        if self._pure:
            self._state = mix(self._state, n)
            self._pure = False
        if st.ndim == 1:
            st = np.outer(st, st.conj())
        red = partial_trace(self._state, n, [mode])
        full = np.tensordot(red, st, axes=0)
        perm = [x for x in range(n) if x != mode] + [mode]
        index_perm = [2 * x + i for x in perm for i in (0, 1)]
        self._state = np.transpose(full, np.argsort(index_perm))

    def prepare_multimode(self, state, modes):
        """circuit.py:393-473: the listed modes (any order) take `state` (ket or
        DM, tensor or vector/matrix shape); the others keep their reduced state."""
        if isinstance(modes, int):
            modes = [modes]
        D, n, k = self._trunc, self._num_modes, len(modes)
        state = np.asarray(state)
        pure_shape, mixed_shape = (D,) * k, (D,) * (2 * k)
        if state.shape == (D**k,):
            state = state.reshape(pure_shape)
        elif state.shape == (D**k, D**k):
            state = state.reshape(mixed_shape)
        if state.shape not in (pure_shape, mixed_shape):
            raise ValueError("Incorrect shape for state preparation")
        if len(modes) != len(set(modes)):
            raise ValueError("The specified modes cannot appear multiple times.")
        if n == k:
            self._state = state.astype(DEF_TYPE)
            self._pure = state.shape == pure_shape
        else:
            if self._pure:
                self._state = mix(self._state, n)
                self._pure = False
            if state.shape == pure_shape:
                state = mix(state, k)
            self._state = np.tensordot(partial_trace(self._state, n, modes), state, axes=0)
        if list(modes) != list(range(n - k, n)):
            perm = [x for x in range(n) if x not in modes] + list(modes)
            if not self._pure:
                perm = [2 * x + i for x in perm for i in (0, 1)]
            self._state = np.transpose(self._state, np.argsort(perm))

    # -- measurements (circuit.py:623-801)
    def project_reset(self, modes, x, state=None):
        """|0..0><x| on `modes` (fockbackend/ops.py:179-198)."""
        state = self._state if state is None else state
        n = self._num_modes
        src = tuple(x[modes.index(i)] if i in modes else slice(None) for i in range(n))
        dst = tuple(0 if i in modes else slice(None) for i in range(n))
        ret = np.zeros_like(state, dtype=DEF_TYPE)
        if self._pure:
            ret[dst] = state[src]
        else:
            ret[tuple(dst[i // 2] for i in range(2 * n))] = state[tuple(src[i // 2] for i in range(2 * n))]
        return ret

    def measure_fock(self, modes, select=None):
        """circuit.py:623-715; numpy's global generator draws the outcome.  The
        sampled modes' distribution comes from the state before any
        post-selection, as there."""
        if select is not None and np.any(np.array(select) == None):  # noqa: E711
            raise NotImplementedError("Post-selection lists must only contain numerical values.")
        n, D = self._num_modes, self._trunc
        state = mix(self._state, n) if self._pure else self._state
        if select is not None:
            if len(select) != len(modes):
                raise ValueError("select and modes differ in length")
            measure = [i for i, v in zip(modes, select) if v is None]
            selected = [i for i, v in zip(modes, select) if v is not None]
            values = [v for v in select if v is not None]
            self._state = self.project_reset(selected, values)
            if self.norm() == 0:
                raise ZeroDivisionError("Measurement has zero probability.")
            self._state = self._state / self.norm()
        else:
            measure = list(modes)
        outcome = []
        if measure:
            unmeasured = [i for i in range(n) if i not in measure]
            reduced = partial_trace(state, n, unmeasured)
            k = len(measure)
            dist = np.ravel(np.einsum("".join(_IDX[i] * 2 for i in range(k)) + "->" + _IDX[:k], reduced).real)
            dist = dist * ~np.isclose(dist, 0.0)
            if sum(dist) != 1:
                i = np.random.choice(list(range(len(dist))), p=dist / sum(dist))
            else:
                i = np.random.choice(list(range(len(dist))), p=dist)
            permuted = [i // D ** (k - 1 - m) % D for m in range(k)]
            perm = np.argsort(measure)
            outcome = [0] * k
            for j in range(k):
                outcome[perm[j]] = permuted[j]
            self._state = self.project_reset(measure, outcome)
            if self.norm() == 0:
                raise ZeroDivisionError("Measurement has zero probability.")
            self._state = self._state / self.norm()
        if select is not None:
            outcome = list(select)
        return np.array([outcome])

    def homodyne_pdf(self, rho, phi, q_mag=10, num_bins=100000):
        """The density of circuit.py:729-760 with ops.hermiteVals (ops.py:500-515):
        raw physicists' Hermite polynomials, as there."""
        from math import factorial

        D, mw = self._trunc, 1 / self._hbar
        rho = phase(-phi, D) @ rho @ phase(-phi, D).conj().T
        q = np.linspace(-q_mag, q_mag, num_bins)
        x = np.sqrt(mw) * q
        H = [np.ones_like(x), 2 * x]
        for i in range(2, D):
            H.append(2 * x * H[i - 1] - 2 * (i - 1) * H[i - 2])
        acc = np.zeros(num_bins, dtype=DEF_TYPE)
        for a in range(D):
            for b in range(D):
                acc += rho[a, b] * H[a] * H[b] / np.sqrt(float(2**a * factorial(a)) * float(2**b * factorial(b)))
        return q, (acc * (mw / np.pi) ** 0.5 * np.exp(-mw * q**2) * (q[1] - q[0])).real

    def homodyne_eigenstate(self, x, phi):
        """R(phi) D(alpha) |truncated infinitely squeezed vacuum>, alpha = x sqrt(1/(2 hbar))
        (circuit.py:772-785)."""
        from math import factorial

        D = self._trunc
        v = np.array(
            [(-0.5) ** (k // 2) * np.sqrt(factorial(k)) / factorial(k // 2) if k % 2 == 0 else 0.0 for k in range(D)],
            dtype=DEF_TYPE,
        )
        alpha = x * np.sqrt(1 / self._hbar / 2)
        return phase(phi, D) @ displacement(np.abs(alpha), np.angle(alpha), D) @ v

    def measure_homodyne(self, phi, mode, select=None, **kwargs):
        """circuit.py:717-801 (hbar = 2)."""
        n = self._num_modes
        if select is not None:
            sample = float(select)
        else:
            state = mix(self._state, n) if self._pure else self._state
            reduced = partial_trace(state, n, [i for i in range(n) if i != mode])
            q, pdf = self.homodyne_pdf(reduced, phi, kwargs.get("max", 10), kwargs.get("num_bins", 100000))
            probs = pdf / np.sum(pdf)
            probs[np.abs(probs) < 1e-10] = 0
            hist = np.random.multinomial(1, probs)
            sample = q[list(hist).index(1)]
        e = self.homodyne_eigenstate(sample, phi)
        P = np.zeros((self._trunc, self._trunc), dtype=DEF_TYPE)
        P[0] = e.conj()
        self._state = self.apply_gate_BLAS(P, [mode])
        self._state = self._state / self.norm()
        return np.array([[sample]])

    def prepare_mode_fock(self, k, mode):
        st = np.zeros(self._trunc, dtype=DEF_TYPE)
        st[k] = 1.0
        self.prepare(st if self._pure else np.outer(st, st.conj()), mode)

    def norm(self):
        if self._pure:
            return np.sqrt(np.vdot(self._state, self._state).real)
        return np.einsum("".join(_IDX[i] * 2 for i in range(self._num_modes)), self._state)

    def is_vacuum(self, tol):
        fid = np.abs(self._state.flat[0]) ** 2 if self._pure else self._state.flat[0]
        return np.abs(fid - 1) <= tol


# ----------------------------------------------------------------- readout
def dm_of(state, n, pure):
    return mix(state, n) if pure else state


def trace(state, n, pure):
    if pure:
        return np.vdot(state, state).real
    return np.einsum("".join(_IDX[i] * 2 for i in range(n)), state).real


def all_fock_probs(state, n, pure):
    D = state.shape[0]
    if pure:
        return (np.abs(state) ** 2).reshape([D] * n)
    ev = list(range(0, 2 * n, 2))
    od = list(range(1, 2 * n, 2))
    flat = D**n
    return np.diag(np.transpose(state, ev + od).reshape(flat, flat)).real.reshape([D] * n)


def reduced_dm(state, n, pure, modes):
    """Reduced DM of sorted `modes`, interleaved axes (states.py:613-642).

    The reference forms the full DM and traces out the other modes.  For a ket
    the same sum, rho_red = sum over traced indices of psi psi*, is contracted
    directly so the D^(2n) array is never built; mixed states take the
    reference's route."""
    traced = [m for m in range(n) if m not in modes]
    if pure:
        a = list(_IDX[:n])
        b = [a[m] if m in traced else _IDX[n + m] for m in range(n)]
        out = "".join(a[m] + b[m] for m in sorted(modes))
        return np.einsum("".join(a) + "," + "".join(b) + "->" + out, state, state.conj())
    rho = dm_of(state, n, pure)
    return partial_trace(rho, n, traced)


def fock_prob(state, n, pure, idx):
    if pure:
        return np.abs(state[tuple(idx)]) ** 2
    return state[tuple(idx[i // 2] for i in range(2 * n))].real


def diagonal_expectation(state, n, pure, modes, values):
    """states.py:947-980 (probability-weighted product of per-mode values)."""
    p = all_fock_probs(state, n, pure)
    traced = tuple(m for m in range(n) if m not in modes)
    ps = p.sum(axis=traced) if traced else p
    # remaining axes are the kept modes in increasing order; weight each
    kept_sorted = sorted(modes)
    w = np.ones_like(ps)
    for ax, _m in enumerate(kept_sorted):
        shape = [1] * ps.ndim
        shape[ax] = -1
        w = w * np.asarray(values).reshape(shape)
    return float((ps * w).sum())


def mean_photon(state, n, pure, mode):
    D = state.shape[0]
    r = reduced_dm(state, n, pure, [mode])
    p = np.diagonal(r).real
    k = np.arange(D)
    mean = float((k * p).sum())
    return mean, float((k**2 * p).sum() - mean**2)


def passive_fock_action(U, psi, n, D):
    """Exact action of the passive interferometer U (a_j^dag -> sum_i U_ij a_i^dag)
    on a ket whose support has fewer than D photons in total, by expanding
    prod_j (b_j^dag)^{k_j} / sqrt(k_j!) |0> photon by photon.  Independent of
    any gate decomposition."""
    from math import factorial

    U = np.asarray(U)
    out = np.zeros([D] * n, dtype=DEF_TYPE)
    for occ_in in product(*([range(D)] * n)):
        amp = psi[occ_in]
        if amp == 0:
            continue
        if sum(occ_in) >= D:
            raise ValueError("input support must hold fewer than D photons")
        terms = {(0,) * n: amp / np.sqrt(float(np.prod([factorial(k) for k in occ_in])))}
        for j, k in enumerate(occ_in):
            for _ in range(k):
                nxt = {}
                for occ, a in terms.items():
                    for i in range(n):
                        o = list(occ)
                        o[i] += 1
                        o = tuple(o)
                        nxt[o] = nxt.get(o, 0) + a * U[i, j] * np.sqrt(o[i])
                terms = nxt
        for occ, a in terms.items():
            out[occ] += a
    return out


# ----------------------------------------------------------------- circuits
def run_ops(ops, n, D, pure=True, faithful=False, state=None):
    """Apply a list of (name, modes, params) ops to vacuum (or `state`).

    name in {"R", "K", "D", "S", "BS", "LOSS", "S2", "MZ", "CK", "V"} (V takes the
    hbar = 2 parameter of the backend call); returns (state, pure).
    """
    c = Circuit(n, D, pure=pure, faithful=faithful)
    if state is not None:
        c._state = np.array(state, dtype=DEF_TYPE)
        c._pure = state.ndim == n
    for name, modes, p in ops:
        if name == "R":
            c.phase_shift(p[0], modes[0])
        elif name == "K":
            c.kerr_interaction(p[0], modes[0])
        elif name == "D":
            c.displacement(p[0], p[1], modes[0])
        elif name == "S":
            c.squeeze(p[0], p[1], modes[0])
        elif name == "BS":
            c.beamsplitter(p[0], p[1], modes[0], modes[1])
        elif name == "LOSS":
            c.loss(p[0], modes[0])
        elif name == "S2":
            c.two_mode_squeeze(p[0], p[1], modes[0], modes[1])
        elif name == "MZ":
            c.mzgate(p[0], p[1], modes[0], modes[1])
        elif name == "CK":
            c.cross_kerr_interaction(p[0], modes[0], modes[1])
        elif name == "V":
            c.cubic_phase_shift(p[0], modes[0])
        else:
            raise ValueError(name)
    return c.get_state()
