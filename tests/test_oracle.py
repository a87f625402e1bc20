"""Pins the CPU oracle (oracle/fock_oracle.py) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign,
a transposed operand, a swapped mode) fails it:
  * gate matrices vs 50-digit mpmath closed forms (tests/golden/gates_mpmath.npz);
  * BS combinatorial factors vs the reference's own data files;
  * Kraus loss vs the independent superoperator identity;
  * slice-loop gate application vs a brute-force Kronecker operator;
  * the reference's known-answer tests, restated at tight tolerances.
"""
import itertools
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.special import comb, factorial

from oracle import fock_oracle as fo

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ matrices
def test_gate_matrices_match_mpmath_closed_forms():
    g = np.load(os.path.join(GOLDEN, "gates_mpmath.npz"))
    fns = {"D": fo.displacement, "S": fo.squeezing, "BS": fo.beamsplitter, "S2": fo.two_mode_squeezing}
    n = 0
    for k in g.files:
        if k.endswith("_params"):
            continue
        p0, p1, D = g[k + "_params"]
        m = fns[k.split("_")[0]](p0, p1, int(D))
        # the FP64 recurrences lose ~1e-12 (S2: ~3e-12) at cutoff 20 and ~1e-9
        # (squeezing) at cutoff 64: conditioning, not a slip
        tol = 1e-13 if D < 15 else (5e-12 if D <= 20 else 5e-9)
        assert np.abs(m - g[k]).max() < tol, k
        n += 1
    assert n == 18


@pytest.mark.parametrize("D", [5, 10])
def test_beamsplitter_factors_match_reference_data(D):
    """fock_beamsplitter_factors_D.npz (reference data) holds prefac[N,n,M,m,k]."""
    ref = sp.load_npz(os.path.join(GOLDEN, "fock_beamsplitter_factors_%d.npz" % D)).toarray()
    ref = ref.reshape(D, D, D, D, D).real
    P = np.zeros((D,) * 5)
    for N, n, M, m in itertools.product(range(D), repeat=4):
        if N + M != n + m:
            continue
        for k in range(D):
            v = comb(n, k) * comb(m, N - k) * comb(N, k) * comb(M, n - k)
            if v > 0:
                P[N, n, M, m, k] = (-1) ** (N - k) * np.sqrt(v)
    assert np.abs(ref - P).max() == 0.0
    # the oracle's recurrence reproduces the closed form built on the data
    for theta, phi in [(0.3, 1.1), (1.2, -0.4), (-0.7, 2.5)]:
        B = fo.beamsplitter(theta, phi, D)
        ct, st = np.cos(theta), np.sin(theta)
        C = np.zeros((D,) * 4, dtype=complex)
        for N, n, M, m in itertools.product(range(D), repeat=4):
            if N + M != n + m:
                continue
            s = sum(
                ref[N, n, M, m, k] * ct ** (m - N + 2 * k) * st ** (n + N - 2 * k)
                for k in range(D)
                if ref[N, n, M, m, k] != 0
            )
            C[N, n, M, m] = np.exp(1j * phi * (n - N)) * s
        assert np.abs(B - C).max() < 1e-12


def test_beamsplitter_blocks_unitary_on_complete_blocks():
    D = 12
    B = fo.beamsplitter(0.83, 0.41, D)
    for s in range(D):
        ks = np.arange(0, s + 1)
        blk = B[ks[:, None], ks[None, :], (s - ks)[:, None], (s - ks)[None, :]]
        assert np.abs(blk @ blk.conj().T - np.eye(s + 1)).max() < 1e-13


def test_rotation_kerr_exact():
    D = 9
    assert np.allclose(np.diag(fo.phase(0.37, D)), np.exp(1j * 0.37 * np.arange(D)), atol=0, rtol=1e-15)
    assert np.allclose(np.diag(fo.kerr(0.2, D)), np.exp(1j * 0.2 * np.arange(D) ** 2), atol=0, rtol=1e-15)


@pytest.mark.parametrize("T", [0.0, 0.3, 0.77, 1.0])
def test_loss_kraus_equals_superoperator(T):
    D = 7
    rng = np.random.default_rng(3)
    rho = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    kr = sum(E @ rho @ E.conj().T for E in fo.loss_kraus(T, D))
    S = fo.loss_superop(T, D)
    sup = np.einsum("abcd,bc->ad", S, rho)
    assert np.abs(kr - sup).max() < 1e-13


# ------------------------------------------------------------------ application
def _kron_op(mats_by_mode, n, D):
    op = np.array([[1.0 + 0j]])
    for m in range(n):
        op = np.kron(op, mats_by_mode.get(m, np.eye(D)))
    return op


def _bs_full(mat, m1, m2, n, D):
    """Dense D^n operator of a beamsplitter with mat's first pair on m1."""
    dim = D**n
    op = np.zeros((dim, dim), dtype=complex)
    for I in range(dim):
        di = np.unravel_index(I, [D] * n)
        for J in range(dim):
            dj = np.unravel_index(J, [D] * n)
            if any(di[k] != dj[k] for k in range(n) if k not in (m1, m2)):
                continue
            op[I, J] = mat[di[m1], dj[m1], di[m2], dj[m2]]
    return op


@pytest.mark.parametrize("faithful", [True, False])
def test_single_mode_gates_vs_kronecker(faithful):
    rng = np.random.default_rng(11)
    n, D = 3, 4
    psi = rng.normal(size=[D] * n) + 1j * rng.normal(size=[D] * n)
    rho = fo.mix(psi, n)
    mats = {
        "R": fo.phase(0.7, D),
        "K": fo.kerr(0.3, D),
        "D": fo.displacement(0.4, 1.2, D),
        "S": fo.squeezing(0.3, -0.8, D),
    }
    for name, U in mats.items():
        for mode in range(n):
            c = fo.Circuit(n, D, pure=True, faithful=faithful)
            c._state = psi.copy()
            out = c.apply_gate_BLAS(U, [mode])
            ref = (_kron_op({mode: U}, n, D) @ psi.reshape(-1)).reshape([D] * n)
            assert np.abs(out - ref).max() < 1e-12, (name, mode)
            c2 = fo.Circuit(n, D, pure=False, faithful=faithful)
            c2._state = rho.copy()
            outm = c2.apply_gate_BLAS(U, [mode])
            K = _kron_op({mode: U}, n, D)
            refm = fo.mix((K @ psi.reshape(-1)).reshape([D] * n), n)
            assert np.abs(outm - refm).max() < 1e-10, (name, mode, "mixed")


@pytest.mark.parametrize("faithful", [True, False])
def test_beamsplitter_vs_kronecker_both_orders(faithful):
    rng = np.random.default_rng(12)
    n, D = 3, 4
    psi = rng.normal(size=[D] * n) + 1j * rng.normal(size=[D] * n)
    mat = fo.beamsplitter(0.7, 0.9, D)
    for m1, m2 in itertools.permutations(range(n), 2):
        c = fo.Circuit(n, D, pure=True, faithful=faithful)
        c._state = psi.copy()
        c.beamsplitter(0.7, 0.9, m1, m2)
        ref = (_bs_full(mat, m1, m2, n, D) @ psi.reshape(-1)).reshape([D] * n)
        assert np.abs(c._state - ref).max() < 1e-12, (m1, m2)
        cm = fo.Circuit(n, D, pure=False, faithful=faithful)
        cm._state = fo.mix(psi, n)
        cm.beamsplitter(0.7, 0.9, m1, m2)
        assert np.abs(cm._state - fo.mix(ref, n)).max() < 1e-10, (m1, m2, "mixed")


@pytest.mark.parametrize("rule", ["_two_mode_passive", "_two_mode_diffcons"])
def test_native_selection_rule_loops_match_the_python_loops(rule):
    """oracle/fock_cpu.c (the faithful mode's stand-in for the reference's numba
    loops) against the same loops in Python, on random matrices and states with
    trailing axes; both are pinned against the Kronecker operator above."""
    if fo.native_lib() is None:
        pytest.skip("oracle/libfock_cpu.so not built")
    rng = np.random.default_rng(13)
    for D, rest in [(1, 3), (2, 1), (5, 7), (8, 16)]:
        mat = rng.normal(size=[D] * 4) + 1j * rng.normal(size=[D] * 4)
        st = rng.normal(size=(D, D, rest)) + 1j * rng.normal(size=(D, D, rest))
        got = getattr(fo.Circuit, rule)(mat, st, D, True)
        saved = fo._NATIVE["lib"]
        try:
            fo._NATIVE["lib"] = None
            want = getattr(fo.Circuit, rule)(mat, st, D, True)
        finally:
            fo._NATIVE["lib"] = saved
        assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(want).max()), (rule, D, rest)


def test_faithful_and_fast_agree_on_config1():
    from strawberryfields_amd import circuits

    c1 = circuits.config1()
    a, _ = fo.run_ops(c1["ops"], 3, 8, faithful=True)
    b, _ = fo.run_ops(c1["ops"], 3, 8, faithful=False)
    assert np.abs(a - b).max() < 1e-14
    g = np.load(os.path.join(GOLDEN, "c1.npz"))
    assert np.abs(a - g["ket"]).max() < 1e-14


def test_loss_on_pure_then_mixed_matches_kraus_by_hand():
    n, D = 2, 5
    ops = [("D", (0,), (0.3, 0.2)), ("LOSS", (0,), (0.6,)), ("S", (1,), (0.2, 0.5)), ("LOSS", (1,), (0.8,))]
    st, pure = fo.run_ops(ops, n, D)
    assert not pure
    psi = fo.vacuum(n, D)
    K0 = _kron_op({0: fo.displacement(0.3, 0.2, D)}, n, D)
    v = K0 @ psi.reshape(-1)
    rho = np.outer(v, v.conj())
    rho = sum(_kron_op({0: E}, n, D) @ rho @ _kron_op({0: E}, n, D).conj().T for E in fo.loss_kraus(0.6, D))
    S1 = _kron_op({1: fo.squeezing(0.2, 0.5, D)}, n, D)
    rho = S1 @ rho @ S1.conj().T
    rho = sum(_kron_op({1: E}, n, D) @ rho @ _kron_op({1: E}, n, D).conj().T for E in fo.loss_kraus(0.8, D))
    got = st.transpose(0, 2, 1, 3).reshape(D * D, D * D)
    assert np.abs(got - rho).max() < 1e-13


# ------------------------------------------------------------------ known answers
# restated from the reference's tests/backend at tight tolerances
@pytest.mark.parametrize("r,p", [(0.0, 0.0), (0.05, 1.0), (0.1, 2.5)])
def test_coherent_state_fock_elements(r, p):
    """test_displacement_operation.py:72-94."""
    D = 12
    st, _ = fo.run_ops([("D", (0,), (r, p))], 1, D)
    a = r * np.exp(1j * p)
    n = np.arange(D)
    ref = np.exp(-0.5 * abs(a) ** 2) * a**n / np.sqrt(factorial(n))
    assert np.abs(st - ref).max() < 1e-14


@pytest.mark.parametrize("r,p", [(0.05, 0.0), (0.15, 2.0)])
def test_squeezed_vacuum_loudon(r, p):
    """test_squeeze_operation.py:104-133 (Loudon 5.5.6)."""
    D = 10
    st, _ = fo.run_ops([("S", (0,), (r, p))], 1, D)
    k = np.arange(0, D, 2)
    ref = np.sqrt(1 / np.cosh(r)) * np.sqrt(factorial(k)) / factorial(k / 2) * (-0.5 * np.exp(1j * p) * np.tanh(r)) ** (k / 2)
    assert np.abs(st[::2] - ref).max() < 1e-14
    assert np.all(st[1::2] == 0)


def test_squeezed_fock_states():
    """test_squeeze_operation.py:135-164 (arXiv quant-ph/0108024 Eq. 20)."""
    from scipy.special import gammaln as lg

    def matrix_elem(n, r, m):
        if n % 2 != m % 2:
            return 0.0
        if r == 0.0:
            return complex(n == m)
        k = np.arange(m % 2, min([m, n]) + 1, 2)
        return np.sum(
            (-1) ** ((n - k) / 2)
            * np.exp((lg(m + 1) + lg(n + 1)) / 2 - lg(k + 1) - lg((m - k) / 2 + 1) - lg((n - k) / 2 + 1))
            * (np.sinh(r) / 2) ** ((n + m - 2 * k) / 2)
            / (np.cosh(r) ** ((n + m + 1) / 2))
        )

    D, r = 8, 0.12
    S = fo.squeezing(r, 0.0, D)
    for m in range(D):
        ref = np.array([matrix_elem(n, r, m) for n in range(D)])
        assert np.abs(S[:, m] - ref).max() < 1e-13


@pytest.mark.parametrize("modes", list(itertools.combinations(range(4), 2)))
def test_beamsplitter_coherent_vacuum(modes):
    """test_beamsplitter_operation.py:139-179: |a>|0> -> |t a>|r a> on a mode subset."""
    D, t, r_phi, mag = 12, 0.6, 2.0, 0.1
    alpha = mag * np.exp(1j * np.pi / 5)
    r = np.exp(1j * r_phi) * np.sqrt(1 - t**2)
    ops = [("D", (modes[0],), (mag, np.pi / 5)), ("BS", modes, (np.arccos(t), r_phi))]
    st, pure = fo.run_ops(ops, 4, D)
    red = fo.reduced_dm(st, 4, pure, list(modes))
    n = np.arange(D)
    A = np.exp(-0.5 * abs(t * alpha) ** 2) * (t * alpha) ** n / np.sqrt(factorial(n))
    B = np.exp(-0.5 * abs(r * alpha) ** 2) * (r * alpha) ** n / np.sqrt(factorial(n))
    ref = np.einsum("i,j,k,l->ijkl", A, A.conj(), B, B.conj())
    assert np.abs(red - ref).max() < 1e-13


def test_hong_ou_mandel():
    """test_tf_integration.py:522-530: |1,1> through a 50:50 BS has p(1,1) = cos^2(2 theta)."""
    D = 5
    for theta in [np.pi / 4, 0.3]:
        psi = np.zeros((D, D), dtype=complex)
        psi[1, 1] = 1
        st, _ = fo.run_ops([("BS", (0, 1), (theta, 0.4))], 2, D, state=psi)
        assert abs(abs(st[1, 1]) ** 2 - np.cos(2 * theta) ** 2) < 1e-14
        assert abs(abs(st[0, 2]) ** 2 - np.sin(2 * theta) ** 2 / 2) < 1e-14


def test_kerr_on_uniform_state():
    """test_nongaussian_gates.py:63-77."""
    D, kappa = 7, 0.21
    st, _ = fo.run_ops([("K", (0,), (kappa,))], 1, D, state=np.ones(D) / D)
    assert np.abs(st - np.exp(1j * kappa * np.arange(D) ** 2) / D).max() < 1e-15


@pytest.mark.parametrize("T", [0.0, 0.25, 0.5, 1.0])
def test_loss_on_coherent_state(T):
    """test_loss_channel.py:186-207 (rho -> |sqrt(T) a><sqrt(T) a|) and T=0 -> vacuum."""
    D, mag, ph = 16, 0.2, 0.7
    a0 = mag * np.exp(1j * ph)
    n = np.arange(D)
    coh = np.exp(-0.5 * abs(a0) ** 2) * a0**n / np.sqrt(factorial(n))
    st, _ = fo.run_ops([("LOSS", (0,), (T,))], 1, D, state=coh)
    a = np.sqrt(T) * a0
    ref = np.exp(-0.5 * abs(a) ** 2) * a**n / np.sqrt(factorial(n))
    assert np.abs(st - np.outer(ref, ref.conj())).max() < 1e-13


def test_gate_then_inverse_returns_vacuum():
    """test_ops_integration.py:53-73 (G then G.H restores vacuum) for the exactly invertible gates."""
    D = 10
    for name, p in [("R", (0.5,)), ("K", (0.3,)), ("BS", (0.4, 0.2))]:
        n = 2 if name == "BS" else 1
        modes = (0, 1) if name == "BS" else (0,)
        start = [("D", (0,), (0.1, 0.3))]
        p_inv = (-p[0],) + tuple(p[1:])
        st_a, _ = fo.run_ops(start + [(name, modes, p), (name, modes, p_inv)], n, D)
        st_b, _ = fo.run_ops(start, n, D)
        assert np.abs(st_a - st_b).max() < 1e-14


# ------------------------------------------------------------------ readout
def test_readouts_pure_vs_mixed():
    from strawberryfields_amd import circuits

    c1 = circuits.config1()
    psi, _ = fo.run_ops(c1["ops"], 3, 8)
    rho = fo.mix(psi, 3)
    assert abs(fo.trace(psi, 3, True) - fo.trace(rho, 3, False)) < 1e-14
    assert np.abs(fo.all_fock_probs(psi, 3, True) - fo.all_fock_probs(rho, 3, False)).max() < 1e-15
    for modes in ([0], [1], [2], [0, 2], [1, 2]):
        a = fo.reduced_dm(psi, 3, True, modes)
        b = fo.reduced_dm(rho, 3, False, modes)
        assert np.abs(a - b).max() < 1e-15
    m, v = fo.mean_photon(psi, 3, True, 1)
    p = fo.all_fock_probs(psi, 3, True)
    assert abs(m - (p.sum(axis=(0, 2)) * np.arange(8)).sum()) < 1e-14
    e = fo.diagonal_expectation(psi, 3, True, [0, 2], np.arange(8))
    w = np.arange(8)[:, None, None] * np.arange(8)[None, None, :]
    assert abs(e - (p * w).sum()) < 1e-13
    assert abs(fo.fock_prob(psi, 3, True, [1, 0, 2]) - p[1, 0, 2]) < 1e-16


def test_coherent_mean_photon():
    D, r = 20, 0.8
    st, pure = fo.run_ops([("D", (0,), (r, 0.1))], 1, D)
    m, v = fo.mean_photon(st, 1, pure, 0)
    assert abs(m - r**2) < 1e-10 and abs(v - r**2) < 1e-9


def test_reduced_dm_of_ket_equals_partial_trace_of_full_dm():
    """The ket route of reduced_dm contracts psi psi* over the traced modes;
    it must equal the reference route (full DM, then partial trace)."""
    rng = np.random.default_rng(7)
    for n, D in [(2, 5), (3, 4), (4, 3)]:
        psi = rng.normal(size=[D] * n) + 1j * rng.normal(size=[D] * n)
        for k in range(1, n + 1):
            for modes in itertools.combinations(range(n), k):
                traced = [m for m in range(n) if m not in modes]
                full = fo.partial_trace(fo.dm_of(psi, n, True), n, traced)
                assert np.abs(fo.reduced_dm(psi, n, True, list(modes)) - full).max() < 1e-12


# ------------------------------------------------------------ S2, MZ, CK, V
def _two_mode_generator_expm(gen_of, D, L):
    """expm of a two-mode generator on an L x L basis, cut to mat[out1, in1, out2, in2]."""
    from scipy.linalg import expm

    a = fo.annihilation(L)
    I = np.eye(L)
    a1, a2 = np.kron(a, I), np.kron(I, a)
    U = expm(gen_of(a1, a2)).reshape(L, L, L, L)[:D, :D, :D, :D]
    return U.transpose(0, 2, 1, 3)


@pytest.mark.parametrize("r,phi", [(0.3, 0.7), (-0.45, 2.0), (0.2, -2.9)])
def test_two_mode_squeezing_matches_generator_exponential(r, phi):
    """The column recurrence against expm of r (e^{i phi} a1^dag a2^dag - h.c.)
    on a 48-level basis (leakage past it below 1e-13 at these r)."""
    D = 8
    ref = _two_mode_generator_expm(
        lambda a1, a2: r * (np.exp(1j * phi) * a1.conj().T @ a2.conj().T - np.exp(-1j * phi) * a1 @ a2), D, D + 40
    )
    assert np.abs(fo.two_mode_squeezing(r, phi, D) - ref).max() < 1e-12


@pytest.mark.parametrize("r", [0.0, 0.1, 0.35])
@pytest.mark.parametrize("p", [0.0, 2.0943951023931953, 4.1887902047863905])
def test_two_mode_squeezed_vacuum_known_answer(r, p):
    """test_twomode_squeezing_operation.py:30-79: amplitude delta_kl
    (e^{ip} tanh r)^k / cosh r, pure and mixed, on every mode pair of 4 modes."""
    D = 5
    amp = np.zeros((D, D), dtype=complex)
    k = np.arange(D)
    amp[k, k] = (np.exp(1j * p) * np.tanh(r)) ** k / np.cosh(r)
    st, pure = fo.run_ops([("S2", (0, 1), (r, p))], 2, D)
    assert np.abs(st - amp).max() < 1e-14
    dm, _ = fo.run_ops([("S2", (0, 1), (r, p))], 2, D, pure=False)
    assert np.abs(dm - np.einsum("ac,bd->abcd", amp, amp.conj())).max() < 1e-14
    for modes in itertools.combinations(range(4), 2):
        st, pure = fo.run_ops([("S2", modes, (r, p))], 4, D)
        red = fo.reduced_dm(st, 4, pure, list(modes))
        assert np.abs(red - np.einsum("ac,bd->abcd", amp, amp.conj())).max() < 1e-14, modes


def test_two_mode_squeezing_selection_rule_and_order():
    """Only out1 - out2 = in1 - in2 entries are nonzero; the difference-conserving
    application equals the dense operator for both mode orders, pure and mixed,
    faithful loop and vectorised."""
    D, n = 4, 3
    mat = fo.two_mode_squeezing(0.4, 1.1, D)
    i, j, k, l = np.indices((D,) * 4)
    assert np.abs(mat[(i - k) != (j - l)]).max() == 0.0
    rng = np.random.default_rng(21)
    psi = rng.normal(size=[D] * n) + 1j * rng.normal(size=[D] * n)
    for faithful in (True, False):
        for m1, m2 in itertools.permutations(range(n), 2):
            c = fo.Circuit(n, D, pure=True, faithful=faithful)
            c._state = psi.copy()
            c.two_mode_squeeze(0.4, 1.1, m1, m2)
            ref = (_bs_full(mat, m1, m2, n, D) @ psi.reshape(-1)).reshape([D] * n)
            assert np.abs(c._state - ref).max() < 1e-12, (m1, m2)
            cm = fo.Circuit(n, D, pure=False, faithful=faithful)
            cm._state = fo.mix(psi, n)
            cm.two_mode_squeeze(0.4, 1.1, m1, m2)
            assert np.abs(cm._state - fo.mix(ref, n)).max() < 1e-10, (m1, m2, "mixed")


@pytest.mark.parametrize("phi_in", [0.0, np.pi / 2, np.pi, 3 * np.pi / 2])
@pytest.mark.parametrize("phi_ex", [0.0, np.pi / 2, np.pi, 3 * np.pi / 2])
def test_mzgate_known_answers(phi_in, phi_ex):
    """test_mzgate_operation.py:21-89 on |1,1>, cutoff 6: normalised output,
    p11 = cos^2 phi_in, p02 = sin^2 phi_in / 2, and the gate equals its
    decomposition R(phi_ex) BS(pi/4, pi/2) R(phi_in) BS(pi/4, pi/2) amplitude by amplitude."""
    D = 6
    one = np.zeros((D, D), dtype=complex)
    one[1, 1] = 1.0
    st, _ = fo.run_ops([("MZ", (0, 1), (phi_in, phi_ex))], 2, D, state=one)
    p = np.abs(st) ** 2
    assert abs(p.sum() - 1) < 1e-13
    assert abs(p[1, 1] - np.cos(phi_in) ** 2) < 1e-13
    assert abs(p[0, 2] - 0.5 * np.sin(phi_in) ** 2) < 1e-13
    dec = [
        ("R", (0,), (phi_ex,)),
        ("BS", (0, 1), (np.pi / 4, np.pi / 2)),
        ("R", (0,), (phi_in,)),
        ("BS", (0, 1), (np.pi / 4, np.pi / 2)),
    ]
    st2, _ = fo.run_ops(dec, 2, D, state=one)
    assert np.abs(st - st2).max() < 1e-13


def test_mzgate_mode_unitary_and_special_cases():
    """Single-photon block equals the documented mode unitary
    (strawberryfields/ops.py:1960-1985), U(pi, pi) = I and U(0, 0) = i X; the
    Fock matrix conserves n1 + n2 and is unitary on every complete block."""
    D = 5
    for pin, pex in [(0.3, 1.1), (np.pi / 2, 0.4), (2.0, -0.5)]:
        M = fo.mzgate(pin, pex, D)
        e, f = np.exp(1j * pin), np.exp(1j * pex)
        U = 0.5 * np.array([[(e - 1) * f, 1j * (1 + e)], [1j * (1 + e) * f, 1 - e]])
        got = np.array([[M[1, 1, 0, 0], M[1, 0, 0, 1]], [M[0, 1, 1, 0], M[0, 0, 1, 1]]])
        assert np.abs(got - U).max() < 1e-14
        i, j, k, l = np.indices((D,) * 4)
        assert np.abs(M[(i + k) != (j + l)]).max() < 1e-15
        for s in range(D):
            ks = np.arange(s + 1)
            blk = M[ks[:, None], ks[None, :], (s - ks)[:, None], (s - ks)[None, :]]
            assert np.abs(blk @ blk.conj().T - np.eye(s + 1)).max() < 1e-13, s
    M = fo.mzgate(np.pi, np.pi, D)
    i, j, k, l = np.indices((D,) * 4)
    assert np.abs(M - ((i == j) & (k == l))).max() < 1e-14
    M = fo.mzgate(0.0, 0.0, D)
    assert abs(M[0, 1, 1, 0] - 1j) < 1e-15 and abs(M[1, 0, 0, 1] - 1j) < 1e-15


@pytest.mark.parametrize("kappa", np.linspace(0, 2 * np.pi, 7))
def test_cross_kerr_known_answer(kappa):
    """test_nongaussian_gates.py:79-102: uniform product state -> e^{i kappa n1 n2} / D,
    through the dense two-mode apply_gate_BLAS route; the passive selection-rule
    route gives the same state (CK conserves n1 and n2)."""
    D = 6
    ket = np.ones((D, D), dtype=complex) / D
    st, _ = fo.run_ops([("CK", (0, 1), (kappa,))], 2, D, state=ket)
    n = np.arange(D)
    assert np.abs(st - np.exp(1j * kappa * np.outer(n, n)) / D).max() < 1e-14
    rng = np.random.default_rng(5)
    psi = rng.normal(size=[D] * 3) + 1j * rng.normal(size=[D] * 3)
    for m1, m2 in itertools.permutations(range(3), 2):
        a, _ = fo.run_ops([("CK", (m1, m2), (kappa,))], 3, D, state=psi)
        c = fo.Circuit(3, D)
        c._state = psi.copy()
        b = c.apply_twomode_gate(fo.cross_kerr(kappa, D), [m1, m2], "BSgate")
        assert np.abs(a - b).max() < 1e-13
        dm, _ = fo.run_ops([("CK", (m1, m2), (kappa,))], 3, D, state=fo.mix(psi, 3))
        assert np.abs(dm - fo.mix(a, 3)).max() < 1e-11


@pytest.mark.parametrize("gamma", np.linspace(0, 6, 7))
def test_cubic_phase_known_answer_and_spectral_form(gamma):
    """test_nongaussian_gates.py:30-61 (uniform ket through expm(i gamma x^3 / 6),
    x = a + a^dag) and the same unitary from the eigendecomposition of the real
    symmetric x^3; hbar scaling: V_hbar(gamma) = V_2(gamma sqrt(hbar / 2))."""
    D = 10
    a = fo.annihilation(D).real
    x3 = np.linalg.matrix_power(a + a.T, 3)
    w, Q = np.linalg.eigh(x3)
    spec = (Q * np.exp(1j * gamma * w / 6)) @ Q.T
    V = fo.cubic_phase(gamma, D)
    assert np.abs(V - spec).max() < 1e-11
    ket = np.ones(D) / np.sqrt(D)
    st, _ = fo.run_ops([("V", (0,), (gamma,))], 1, D, state=ket.astype(complex))
    assert np.abs(st - spec @ ket).max() < 1e-11
    assert np.abs(fo.cubic_phase(gamma, D, hbar=1.3) - fo.cubic_phase(gamma * np.sqrt(1.3 / 2), D)).max() < 1e-12


# ------------------------------------------------------------ measurements, preparations
def _fock_product(ns, D, pure):
    c = fo.Circuit(len(ns), D, pure=pure)
    for m, k in enumerate(ns):
        c.prepare_mode_fock(k, m)
    return c


@pytest.mark.parametrize("pure", [True, False])
def test_fock_measurement_known_answers(pure):
    """test_fock_measurement.py:86-150: Fock product states give their photon
    numbers on every mode subset, and conditional states stay normalised;
    test_fock_measurement.py:169-177: the vacuum measures zero."""
    D = 5
    np.random.seed(3)
    choices = [(0,), (1,), (2,), (0, 1), (0, 2), (1, 2), (0, 1, 2), (2, 0)]
    for idx in range(16):
        ns = [idx % D, (idx + 1) % D, (idx + 2) % D]
        modes = list(choices[idx % len(choices)])
        c = _fock_product(ns, D, pure)
        res = c.measure_fock(modes)
        assert res.tolist() == [[ns[m] for m in modes]]
        assert abs(fo.trace(c._state, 3, c._pure) - 1) < 1e-13
        # the measured modes are reset to vacuum
        probs = fo.all_fock_probs(c._state, 3, c._pure)
        expect = np.zeros([D] * 3)
        expect[tuple(0 if m in modes else ns[m] for m in range(3))] = 1
        assert np.abs(probs - expect).max() < 1e-13
    c = fo.Circuit(3, D, pure=pure)
    assert c.measure_fock([0, 1, 2]).tolist() == [[0, 0, 0]]


@pytest.mark.parametrize("pure", [True, False])
def test_fock_measurement_tmsv_and_postselection(pure):
    """test_fock_measurement.py:151-167 (two-mode squeezed vacuum: equal counts)
    and test_postselection.py:52-68 (post-selecting one beamsplitter output
    conserves the total photon number in the other)."""
    D = 6
    np.random.seed(5)
    for _ in range(10):
        c = fo.Circuit(2, D, pure=pure)
        c.squeeze(-0.25, 0, 0)
        c.squeeze(0.25, 0, 1)
        c.beamsplitter(np.pi / 4, np.pi, 0, 1)
        r = c.measure_fock([0, 1])
        assert r[0][0] == r[0][1]
    for nphot in range(D - 1):
        c = fo.Circuit(2, D, pure=pure)
        c.prepare_mode_fock(nphot, 0)
        c.prepare_mode_fock(D - 1 - nphot, 1)
        c.beamsplitter(np.pi / 4, 0, 0, 1)
        try:
            r1 = c.measure_fock([0], select=[D // 2])
        except ZeroDivisionError:
            continue
        r2 = c.measure_fock([1])
        assert r1[0][0] + r2[0][0] == D - 1


def test_fock_measurement_distribution_matches_marginal():
    """Many draws from one state follow its marginal distribution (chi-square
    style bound on the largest deviation)."""
    D = 4
    st, _ = fo.run_ops([("D", (0,), (0.6, 0.2)), ("BS", (0, 1), (0.7, 0.3))], 2, D)
    p = fo.all_fock_probs(st, 2, True).sum(axis=1)
    np.random.seed(11)
    counts = np.zeros(D)
    N = 4000
    for _ in range(N):
        c = fo.Circuit(2, D)
        c._state = st.copy()
        counts[c.measure_fock([0])[0][0]] += 1
    assert np.abs(counts / N - p / p.sum()).max() < 5 * np.sqrt(0.25 / N)


def test_homodyne_density_closed_forms():
    """pdf of the vacuum is N(0, hbar/2) and of |1> is q^2 e^{-q^2/hbar} (2/hbar)
    / sqrt(pi hbar) (hbar = 2), times dq; a rotation by phi of a coherent state
    moves the mean to 2 |alpha| cos(arg alpha - phi)."""
    D = 12
    c = fo.Circuit(1, D)
    vac = np.zeros((D, D), dtype=complex)
    vac[0, 0] = 1
    q, pdf = c.homodyne_pdf(vac, 0.0, 10, 20001)
    dq = q[1] - q[0]
    assert np.abs(pdf - np.exp(-(q**2) / 2) / np.sqrt(2 * np.pi) * dq).max() < 1e-15
    one = np.zeros((D, D), dtype=complex)
    one[1, 1] = 1
    q, pdf = c.homodyne_pdf(one, 0.7, 10, 20001)
    assert np.abs(pdf - q**2 * np.exp(-(q**2) / 2) / np.sqrt(2 * np.pi) * dq).max() < 1e-15
    st, _ = fo.run_ops([("D", (0,), (0.8, 0.4))], 1, 20)
    c20 = fo.Circuit(1, 20)
    for phi in [0.0, 0.9]:
        q, pdf = c20.homodyne_pdf(np.outer(st, st.conj()), phi, 10, 20001)
        assert abs((q * pdf).sum() - 2 * 0.8 * np.cos(0.4 - phi)) < 1e-9


def test_homodyne_known_answers():
    """test_homodyne.py:27-38 (the measured mode is reset to vacuum),
    test_postselection.py:20-33 (post-selection returns the selected value) and
    test_homodyne.py:40-55 (vacuum: mean 0, std 1 over many draws)."""
    D = 10
    c = fo.Circuit(1, D)
    c.squeeze(np.arcsinh(1.0), 0.0, 0)
    c.displacement(0.8, 0.0, 0)
    np.random.seed(2)
    c.measure_homodyne(0, 0, num_bins=4001)
    assert c.is_vacuum(1e-12)
    c2 = fo.Circuit(2, D)
    c2.squeeze(1.0, 0, 0)
    c2.squeeze(-1.0, 0, 1)
    c2.beamsplitter(np.pi / 4, 0, 0, 1)
    assert c2.measure_homodyne(0, 0, select=0.2)[0][0] == 0.2
    assert abs(fo.trace(c2._state, 2, c2._pure) - 1) < 1e-12
    red = fo.reduced_dm(c2._state, 2, c2._pure, [0])
    assert abs(red[0, 0] - 1) < 1e-12
    xs = []
    np.random.seed(7)
    for _ in range(300):
        v = fo.Circuit(1, D)
        xs.append(v.measure_homodyne(0, 0, num_bins=4001)[0][0])
    xs = np.array(xs)
    std10 = 10 / np.sqrt(len(xs))
    assert abs(xs.mean()) < std10 and abs(xs.std() - 1) < std10


def test_multimode_preparation_known_answers():
    """test_multimode_state_preparations.py:34-222: product kets and DMs on mode
    pairs in any order equal successive single-mode preparations; the vector /
    matrix shapes equal the tensor shapes; other modes keep their state."""
    rng = np.random.default_rng(8)
    D = 4
    k0 = rng.normal(size=D) + 1j * rng.normal(size=D)
    k1 = rng.normal(size=D) + 1j * rng.normal(size=D)
    k0, k1 = k0 / np.linalg.norm(k0), k1 / np.linalg.norm(k1)
    ket = np.outer(k0, k1)
    rho = np.einsum("ij,kl->ikjl", ket, ket.conj())
    for pure in (True, False):
        c = fo.Circuit(4, D, pure=pure)
        c.prepare_multimode(ket, [3, 1])
        red = fo.reduced_dm(c._state, 4, c._pure, [1, 3])  # modes in increasing order
        assert np.abs(red - rho.transpose(2, 3, 0, 1)).max() < 1e-14
        single = fo.Circuit(4, D, pure=pure)
        single.prepare(k0, 0)
        single.prepare(k1, 1)
        multi = fo.Circuit(4, D, pure=pure)
        multi.prepare_multimode(ket.reshape(-1), [0, 1])
        assert np.abs(single._state - multi._state).max() < 1e-14
        dmat = fo.Circuit(4, D, pure=pure)
        dmat.prepare_multimode(rho.reshape(D * D, D * D), [0, 1])
        assert np.abs(dmat._state - multi._state).max() < 1e-14
    # all modes at once, permuted: a ket stays a ket
    full = rng.normal(size=[D] * 3) + 1j * rng.normal(size=[D] * 3)
    c = fo.Circuit(3, D)
    c.prepare_multimode(full, [2, 0, 1])
    assert c._pure and np.abs(c._state - full.transpose(1, 2, 0)).max() == 0.0


def test_passive_fock_action_known_answers():
    """Single photons follow the columns of U; Hong-Ou-Mandel for a 50:50
    splitter; the two-mode action equals the (mpmath-pinned) beamsplitter
    matrix; a 3-mode unitary equals its Clements gate list run through the
    oracle's R and BS gates."""
    from scipy.stats import unitary_group
    from strawberryfields_amd.interferometer import interferometer_gates

    D = 4
    U = unitary_group.rvs(3, random_state=2)
    for j in range(3):
        psi = np.zeros([D] * 3, dtype=complex)
        psi[tuple(1 if m == j else 0 for m in range(3))] = 1
        out = fo.passive_fock_action(U, psi, 3, D)
        for i in range(3):
            assert abs(out[tuple(1 if m == i else 0 for m in range(3))] - U[i, j]) < 1e-15
    bs = np.array([[1, -1], [1, 1]]) / np.sqrt(2)
    psi = np.zeros((D, D), dtype=complex)
    psi[1, 1] = 1
    out = fo.passive_fock_action(bs, psi, 2, D)
    assert abs(out[1, 1]) < 1e-15 and abs(abs(out[2, 0]) ** 2 - 0.5) < 1e-15
    th, ph = 0.7, 0.3
    U2 = np.array([[np.cos(th), -np.exp(-1j * ph) * np.sin(th)], [np.exp(1j * ph) * np.sin(th), np.cos(th)]])
    rng = np.random.default_rng(1)
    psi = np.zeros((D, D), dtype=complex)
    for a in range(D):
        for b in range(D - a):
            psi[a, b] = rng.normal() + 1j * rng.normal()
    st, _ = fo.run_ops([("BS", (0, 1), (th, ph))], 2, D, state=psi)
    assert np.abs(fo.passive_fock_action(U2, psi, 2, D) - st).max() < 1e-14
    psi3 = np.zeros([D] * 3, dtype=complex)
    for occ in itertools.product(range(D), repeat=3):
        if sum(occ) < D:
            psi3[occ] = rng.normal() + 1j * rng.normal()
    for mesh in ("rectangular", "rectangular_phase_end", "rectangular_symmetric"):
        st, _ = fo.run_ops(interferometer_gates(U, mesh=mesh), 3, D, state=psi3)
        assert np.abs(fo.passive_fock_action(U, psi3, 3, D) - st).max() < 1e-13, mesh
