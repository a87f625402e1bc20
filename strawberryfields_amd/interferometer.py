"""Decompositions of an interferometer into the gates of this backend.

``Interferometer(U, mesh)`` of the front end (``strawberryfields/ops.py:2556-2718``)
is compiled, as in the reference, into rotations and beamsplitters (or
Mach-Zehnder gates for the symmetric mesh) that the HIP kernels then run.  The
rectangular meshes follow Clements et al., Optica 3, 1460 (2016), the
triangular one Reck et al., PRL 73, 58 (1994), as
``strawberryfields/decompositions.py:233-633`` states them: the two-mode block

    T(m, n, theta, phi) = [[e^{i phi} cos theta, -sin theta],
                           [e^{i phi} sin theta,  cos theta]]   (rows/cols m, n)

equals BS(theta, 0) R_m(phi) in the beamsplitter convention of this backend
(mode unitary [[cos, -e^{-i phi} sin], [e^{i phi} sin, cos]]).  Diagonals of
U are cleared alternately from the right (U <- U T^{-1}) and from the left
(U <- T U), leaving U = (prod of left T)^{-1} diag (prod of right T^{-1})^{-1}.
The elimination order and the angle formulas are those of the paper, so the
gate list matches the reference's one gate for gate.
"""
import numpy as np


def _t(m, n, theta, phi, N):
    M = np.identity(N, dtype=np.complex128)
    M[m, m] = np.exp(1j * phi) * np.cos(theta)
    M[m, n] = -np.sin(theta)
    M[n, m] = np.exp(1j * phi) * np.sin(theta)
    M[n, n] = np.cos(theta)
    return M


def _t_inv(m, n, theta, phi, N):
    # T^{-1} = T(theta, -phi)^T
    return _t(m, n, theta, -phi, N).T


def _angles(num, den):
    """(theta, phi) with tan(theta) e^{i phi} = num / den; the zero and
    divide-by-zero cases as in the reference (no rotation, full swap)."""
    if num == 0:
        return 0.0, 0.0
    if den == 0:
        return np.pi / 2, 0.0
    r = num / den
    return float(np.arctan(np.abs(r))), float(np.angle(r))


def rectangular(U, tol=1e-11):
    """(right_list, diag, left_list): entries [m, n, theta, phi, N] of the T
    blocks cleared from the right and from the left, and the remaining phases
    (decompositions.py:294-339)."""
    V = np.array(U, dtype=np.complex128)
    N = V.shape[0]
    if V.shape != (N, N) or not np.allclose(V @ V.conj().T, np.identity(N), atol=tol, rtol=0):
        raise ValueError("The input matrix is not unitary")
    right, left = [], []
    for k, i in enumerate(range(N - 2, -1, -1)):
        if k % 2 == 0:
            # clear V[i+j+1, j] from the right with T^{-1} on columns (j, j+1)
            for j in reversed(range(N - 1 - i)):
                row = i + j + 1
                theta, phi = _angles(V[row, j], V[row, j + 1])
                right.append([j, j + 1, theta, phi, N])
                V = V @ _t_inv(j, j + 1, theta, phi, N)
        else:
            # clear V[i+j+1, j] from the left with T on rows (i+j, i+j+1)
            for j in range(N - 1 - i):
                row = i + j + 1
                theta, phi = _angles(-V[row, j], V[row - 1, j])
                left.append([row - 1, row, theta, phi, N])
                V = _t(row - 1, row, theta, phi, N) @ V
    return right, np.diag(V), left


def rectangular_phase_end(U, tol=1e-11):
    """(tlist, diag, None): the rectangular mesh with every local phase moved
    to the output (Clements et al., Eq. 5; decompositions.py:342-381).

    Each left block T^{-1} of ``rectangular`` is pushed through the diagonal,
    last block first: T(theta, phi)^{-1} diag(.., e^{i a}, e^{i b}, ..) =
    diag(.., e^{i(b - phi + pi)}, e^{i b}, ..) T(theta, (a - b + pi) mod 2 pi),
    so the whole mesh becomes T blocks followed by one diagonal."""
    right, diag, left = rectangular(U, tol)
    tlist, d = list(right), np.array(diag, dtype=np.complex128)
    for m, n, theta, phi, N in reversed(left):
        a, b = np.angle(d[m]), np.angle(d[n])
        tlist.append([m, n, theta, (a - b + np.pi) % (2 * np.pi), N])
        d[m], d[n] = np.exp(1j * (b - phi + np.pi)), np.exp(1j * b)
    return tlist, d, None


def mach_zehnder(m, n, phi_in, phi_ex, N):
    """Mode unitary of MZgate(phi_in, phi_ex) on (m, n): B R_m(phi_in) B
    R_m(phi_ex) with B the symmetric 50:50 beamsplitter [[1, i], [i, 1]] / sqrt 2,
    rounded to 14 decimals as decompositions.py:384-419 rounds it (the rounding
    is part of the elimination, so the angles follow it)."""
    B = np.identity(N, dtype=np.complex128)
    h = 1 / np.sqrt(2)
    B[m, m], B[m, n], B[n, m], B[n, n] = h, 1j * h, 1j * h, h
    Ri = np.identity(N, dtype=np.complex128)
    Ri[m, m] = np.exp(1j * phi_in)
    Re = np.identity(N, dtype=np.complex128)
    Re[m, m] = np.exp(1j * phi_ex)
    return np.round(B @ Ri @ B @ Re, 14)


def _mz_angles(num, den):
    """(phi_in, phi_ex) of the Mach-Zehnder block that clears the entry `den`,
    with r = num / den (num from its partner): phi_in = 2 arctan|r|,
    phi_ex = -arg r; an entry that is already zero takes (pi, 0) and a zero
    partner the full swap (0, 0) (decompositions.py:429-492), both reduced mod
    2 pi twice as the reference reduces them (an identity would otherwise leave
    2 pi)."""
    if den == 0:
        pi, pe = np.pi, 0.0
    elif num == 0:
        pi, pe = 0.0, 0.0
    else:
        r = num / den
        pi, pe = 2 * np.arctan(np.abs(r)), -np.angle(r)
    for _ in range(2):
        pi, pe = pi % (2 * np.pi), pe % (2 * np.pi)
    return float(pi), float(pe)


def rectangular_mz(U, tol=1e-11):
    """(right, diag, left) of the rectangular mesh built from Mach-Zehnder
    blocks instead of T blocks (decompositions.py:495-545): the same
    elimination order as ``rectangular``, entries [m, n, phi_in, phi_ex, N]."""
    V = np.array(U, dtype=np.complex128)
    N = V.shape[0]
    if V.shape != (N, N) or not np.allclose(V @ V.conj().T, np.identity(N), atol=tol, rtol=0):
        raise ValueError("The input matrix is not unitary")
    right, left = [], []
    for k, i in enumerate(range(N - 2, -1, -1)):
        if k % 2 == 0:
            for j in reversed(range(N - 1 - i)):
                row = i + j + 1
                pi, pe = _mz_angles(-V[row, j + 1], V[row, j])
                right.append([j, j + 1, pi, pe, N])
                V = V @ mach_zehnder(j, j + 1, pi, pe, N).conj().T
        else:
            for j in range(N - 1 - i):
                row = i + j + 1
                pi, pe = _mz_angles(V[row - 1, j], V[row, j])
                left.append([row - 1, row, pi, pe, N])
                V = mach_zehnder(row - 1, row, pi, pe, N) @ V
    return right, np.diag(V), left


def rectangular_symmetric(U, tol=1e-11):
    """(tlist, diag, None): Mach-Zehnder mesh with the local phases at the
    output (decompositions.py:548-601).  Each left block M^{-1} is pushed
    through the diagonal, last first: M(phi_in, phi_ex)^{-1} diag(.., e^{i a},
    e^{i b}, ..) = diag(.., e^{i(b - phi_ex - phi_in + pi)}, e^{i(b - phi_in +
    pi)}, ..) M(phi_in, a - b), every phase reduced mod 2 pi."""
    right, diag, left = rectangular_mz(U, tol)
    tlist, d = list(right), np.array(diag, dtype=np.complex128)
    tp = 2 * np.pi
    for m, n, phi_in, phi_ex, N in reversed(left):
        a, b = np.angle(d[m]), np.angle(d[n])
        new_ex = ((a - b) % tp) % tp
        new_in = (phi_in % tp) % tp
        tlist.append([m, n, new_in, new_ex, N])
        d[m] = np.exp(1j * ((b - phi_ex - phi_in + np.pi) % tp))
        d[n] = np.exp(1j * ((b - phi_in + np.pi) % tp))
    return tlist, d, None


def triangular(U, tol=1e-11):
    """(tlist, diag, None) of the Reck et al. triangular mesh
    (decompositions.py:604-633): entries below the diagonal are cleared column
    by column from the left with T blocks on rows (r - 1, r), bottom row first;
    the list is returned last block first.  The decomposition satisfies
    U = prod_{t in tlist} T_t^{-1} diag (product written right to left), as the
    reference's own test of it states."""
    V = np.array(U, dtype=np.complex128)
    N = V.shape[0]
    if V.shape != (N, N) or not np.allclose(V @ V.conj().T, np.identity(N), atol=tol, rtol=0):
        raise ValueError("The input matrix is not unitary")
    tlist = []
    for i in range(N - 2, -1, -1):
        col = N - i - 2
        for j in range(i + 1):
            row = N - j - 1
            theta, phi = _angles(-V[row, col], V[row - 1, col])
            tlist.append([row - 1, row, theta, phi, N])
            V = _t(row - 1, row, theta, phi, N) @ V
    return list(reversed(tlist)), np.diag(V), None


# ---- compact meshes of symmetric Mach-Zehnder blocks (Bell & Walmsley,
# APL Photonics 6, 070804 (2021), arXiv:2104.07561), as decompositions.py:636-905
# applies them.  A block on modes (n, n + 1) is
#   M(sigma, delta) = e^{i sigma} [[sin delta, cos delta], [cos delta, -sin delta]],
# the symmetric MZ with internal phases sigma +- delta; the elimination works on
# V = conj(U) and clears entries column by column with blocks applied from the
# right (triangular) or alternately from the right and the left (rectangular).


def _smz(n, sigma, delta, N):
    M = np.identity(N, dtype=np.complex128)
    e = np.exp(1j * sigma)
    M[n, n], M[n, n + 1], M[n + 1, n], M[n + 1, n + 1] = e * np.sin(delta), e * np.cos(delta), e * np.cos(delta), -e * np.sin(delta)
    return M


def _phase(j, phi, N):
    M = np.identity(N, dtype=np.complex128)
    M[j, j] = np.exp(1j * phi)
    return M


def _check_unitary(U, rtol, atol):
    if U.shape[0] != U.shape[1]:
        raise ValueError("Matrix is not square")
    if not np.allclose(U @ U.conj().T, np.eye(U.shape[0]), rtol=rtol, atol=atol):
        raise ValueError("The input matrix is not unitary")


def _delta(big, small):
    """delta with tan(delta) = -|small| / |big| (arctan2 of the two moduli); a
    zero entry `big` gives pi / 2."""
    return 0.5 * np.pi if big == 0 else float(np.arctan2(-abs(small), abs(big)))


def _final_phases(V, N, rtol, atol):
    """Residual diagonal: zeta_0 = -arg V[0, 0], zeta_j = arg V[0, 0] - arg V[j, j]
    after the earlier ones are applied; V must end as the identity."""
    zetas = {}
    zetas[0] = float(-np.angle(V[0, 0]))
    V = V @ _phase(0, zetas[0], N)
    for j in range(1, N):
        zetas[j] = float(np.angle(V[0, 0]) - np.angle(V[j, j]))
        V = V @ _phase(j, zetas[j], N)
    if not np.allclose(V, np.eye(N), rtol=rtol, atol=atol):
        raise RuntimeError("decomposition failed")
    return zetas


def triangular_compact(U, rtol=1e-12, atol=1e-12):
    """Phases of the compact triangular (Reck) mesh (decompositions.py:672-737):
    dict with m, phi_ins (mode -> phi), deltas and sigmas ((mode, layer) ->
    value) and zetas (mode -> output phase)."""
    U = np.asarray(U)
    _check_unitary(U, rtol, atol)
    V = U.conj()
    N = U.shape[0]
    out = {"m": N, "phi_ins": {}, "deltas": {}, "sigmas": {}, "zetas": {}}
    for j in range(N - 1):
        x, y = N - 1, j
        phi = float(np.angle(V[x, y]) - np.angle(V[x, y + 1]))
        out["phi_ins"][j] = phi
        V = V @ _phase(j + 1, phi, N)
        for k in range(j + 1):
            n = j - k
            delta = _delta(V[x, y], V[x, y + 1])
            W = V @ _smz(n, 0, delta, N)
            # y - 1 is -1 for the last block of a column: numpy's wrap-around to
            # the last column, as in the reference (the output phases absorb it)
            sigma = float(np.angle(W[x - 1, y - 1]) - np.angle(W[x - 1, y]))
            out["deltas"][n, k], out["sigmas"][n, k] = delta, sigma
            V = V @ _smz(n, sigma, delta, N)
            x, y = x - 1, y - 1
    out["zetas"] = _final_phases(V, N, rtol, atol)
    return out


def _rectangular_compact_init(U, rtol, atol):
    """Phases of the compact rectangular (Clements) mesh before the residual
    phases are moved to the edge (decompositions.py:740-826)."""
    V = U.conj()
    N = U.shape[0]
    out = {"m": N, "phi_ins": {}, "deltas": {}, "sigmas": {}, "zetas": {}, "phi_outs": {}}
    for j in range(N - 1):
        if j % 2 == 0:
            x, y = N - 1, j
            phi = float(np.angle(V[x, y + 1]) - np.angle(V[x, y]))
            V = V @ _phase(j, phi, N)
            out["phi_ins"][j] = phi
            for k in range(j + 1):
                delta = _delta(V[x, y], V[x, y + 1])
                n = j - k
                W = V @ _smz(n, 0, delta, N)
                sigma = float(np.angle(W[x - 1, y - 1]) - np.angle(W[x - 1, y]))
                V = V @ _smz(n, sigma, delta, N)
                out["deltas"][n, k], out["sigmas"][n, k] = delta, sigma
                x, y = x - 1, y - 1
        else:
            x, y = N - j - 1, 0
            phi = float(np.angle(V[x - 1, y]) - np.angle(V[x, y]))
            V = _phase(x, phi, N) @ V
            out["phi_outs"][x] = phi
            for k in range(j + 1):
                delta = 0.5 * np.pi if V[x, y] == 0.0 else float(np.arctan2(abs(V[x - 1, y]), abs(V[x, y])))
                W = _smz(x - 1, 0, delta, N) @ V
                n = N + k - j - 2
                sigma = float(np.angle(W[x + 1, y + 1]) - np.angle(W[x, y + 1])) if j != k else 0.0
                out["deltas"][n, N - k - 1], out["sigmas"][n, N - k - 1] = delta, sigma
                V = _smz(n, sigma, delta, N) @ V
                x, y = x + 1, y + 1
    out["zetas"] = _final_phases(V, N, rtol, atol)
    return out


def _absorb_zetas(ph):
    """Move the residual phases zeta_j into the sMZ sigmas of the layers they
    cross and into phase shifters on the last mode's edge
    (decompositions.py:829-876)."""
    import collections

    N = ph["m"]
    out = dict(ph)
    zetas = out.pop("zetas")
    out["phi_edges"] = collections.defaultdict(float)
    if N % 2 == 0:
        out["phi_outs"][0] = zetas[0]
        js, shift, parity = range(1, N), 0, 1
    else:
        js, shift, parity = range(N), 1, 0
    for j in js:
        zeta = zetas[j]
        layer = N - j - shift
        for mode in range(j, N - 1, 2):
            out["sigmas"][mode, layer] += zeta
        for mode in range(j + 1, N - 1, 2):
            out["sigmas"][mode, layer - 1] -= zeta
        if layer % 2 == parity:
            out["phi_edges"][N - 1, layer] += zeta
        else:
            out["phi_edges"][N - 1, layer - 1] -= zeta
    return out


def rectangular_compact(U, rtol=1e-12, atol=1e-12):
    """Phases of the compact rectangular mesh with the residual phases on the
    edge (decompositions.py:879-903)."""
    U = np.asarray(U)
    _check_unitary(U, rtol, atol)
    return _absorb_zetas(_rectangular_compact_init(U, rtol, atol))


def smz_gates(modes, phi1, phi2):
    """sMZgate(phi1, phi2) on (m, n) as the fock compiler runs it
    (ops.py:2015-2031): BS(pi/4, pi/2), R(phi2 - pi/2) on n, R(phi1 - pi/2) on m,
    BS(pi/4, pi/2)."""
    m, n = modes
    return [("BS", (m, n), (np.pi / 4, np.pi / 2)), ("R", (n,), (phi2 - np.pi / 2,)),
            ("R", (m,), (phi1 - np.pi / 2,)), ("BS", (m, n), (np.pi / 4, np.pi / 2))]


def _compact_gates(mesh, U, tol):
    """Gate list of a compact mesh (ops.py:2461-2499): sMZ blocks with internal
    phases sigma +- delta, input, edge and output phase shifters."""
    gates = []
    if mesh == "rectangular_compact":
        ph = rectangular_compact(U, rtol=tol, atol=tol)
        N = ph["m"]
        for j in range(0, N - 1, 2):
            gates.append(("R", (j,), (ph["phi_ins"][j],)))
        for layer in range(N):
            if (layer + N + 1) % 2 == 0:
                gates.append(("R", (N - 1,), (ph["phi_edges"][N - 1, layer],)))
            for mode in range(layer % 2, N - 1, 2):
                s, d = ph["sigmas"][mode, layer], ph["deltas"][mode, layer]
                gates += smz_gates((mode, mode + 1), s + d, s - d)
        for j, phi in ph["phi_outs"].items():
            gates.append(("R", (j,), (phi,)))
    else:
        ph = triangular_compact(U, rtol=tol, atol=tol)
        N = ph["m"]
        for j in range(N - 1):
            gates.append(("R", (j + 1,), (ph["phi_ins"][j],)))
            for k in range(j + 1):
                n = j - k
                s, d = ph["sigmas"][n, k], ph["deltas"][n, k]
                gates += smz_gates((n, n + 1), s + d, s - d)
        for j in range(N):
            gates.append(("R", (j,), (ph["zetas"][j],)))
    return gates


# ---- SU(n) factorisation into SU(2) blocks on neighbouring modes
# (decompositions.py:1073-1381): U = e^{i phi / n} S, S in SU(n) reduced by a
# staircase of SU(2) rotations to SU(n - 1) down to SU(3), which factors as
# SU_23 SU_12 SU_23 (Rowe et al., J. Math. Phys. 40, 3604 (1999)).  A block
# with Euler angles (a, b, g) is diag(e^{ia/2}, e^{-ia/2}) [[cos b/2, -sin b/2],
# [sin b/2, cos b/2]] diag(e^{ig/2}, e^{-ig/2}).


def _su2_params(U, tol=1e-10):
    """Euler angles [a, b, g] of a 2x2 special unitary (decompositions.py:1293-1335)."""
    if U.shape != (2, 2):
        raise ValueError("an SU(2) block needs a 2x2 matrix, got shape %s" % (U.shape,))
    if not np.isclose(np.linalg.det(U), 1, atol=tol, rtol=0):
        raise ValueError("an SU(2) block needs determinant 1, got %s" % np.linalg.det(U))
    if np.isclose(np.abs(U[0, 1]), 1, atol=tol, rtol=0):
        b = 2 * np.arcsin(1)
    else:
        b = 2 * np.arcsin(np.abs(U[0, 1]))
    half_sum = np.angle(U[0, 0])  # (a + g) / 2
    half_diff = -np.angle(U[1, 0])  # (a - g) / 2
    return [half_sum + half_diff, b, half_sum - half_diff]


def _su3_params(U):
    """Three SU(2) blocks (on modes 2-3, 1-2, 2-3) of a 3x3 special unitary,
    with the identity-column and phase-only special cases
    (decompositions.py:1338-1381)."""
    if U.shape != (3, 3):
        raise ValueError("the SU(3) base needs a 3x3 matrix, got shape %s" % (U.shape,))
    if not np.isclose(np.linalg.det(U), 1):
        raise ValueError("the SU(3) base needs determinant 1, got %s" % np.linalg.det(U))
    x, y, z = U[0, 0], U[1, 0], U[2, 0]
    if np.isclose(x, 1):
        return [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], _su2_params(U[1:, 1:])]
    if np.isclose(np.abs(x), 1):
        ph = np.array([[np.conj(x), 0], [0, x]])
        full = np.identity(3, dtype=complex)
        full[0:2, 0:2] = ph
        rest = (full @ U)[1:, 1:]
        return [[0.0, 0.0, 0.0], _su2_params(ph.conj().T), _su2_params(rest)]
    cf = np.sqrt(1 - np.abs(x) ** 2)
    Y, Z = y / cf, z / cf
    left = np.array([[1, 0, 0], [0, Y, -np.conj(Z)], [0, Z, np.conj(Y)]])
    middle = np.array([[x, -cf, 0], [cf, np.conj(x), 0], [0, 0, 1]])
    right = middle.conj().T @ left.conj().T @ U
    return [_su2_params(left[1:, 1:]), _su2_params(middle[0:2, 0:2]), _su2_params(right[1:, 1:])]


def _staircase(U, rtol, atol):
    """SU(2) blocks on modes (n-1, n), ..., (1, 2) that reduce the first column
    of U to e_1, and the reduced matrix (decompositions.py:1187-1290), with the
    reference's branches for a column that already has a single nonzero entry."""
    n = U.shape[0]
    out = []
    cur = U
    moduli = np.abs(U[:, 0])
    if np.allclose(sorted(moduli), [0.0] * (n - 1) + [1], rtol, atol):
        if np.isclose(cur[0, 0], 1, rtol, atol):
            out = [[0.0, 0.0, 0.0]] * (n - 1)
        elif np.isclose(np.abs(cur[0, 0]), 1, rtol, atol):
            ph = np.array([[cur[0, 0].conjugate(), 0], [0, cur[0, 0]]])
            out = [[0.0, 0.0, 0.0]] * (n - 2) + [_su2_params(ph.conj().T)]
            full = np.identity(n, dtype=complex)
            full[0:2, 0:2] = ph
            cur = full @ cur
        else:
            for r in range(n - 1, 0, -1):
                if not np.isclose(cur[r, 0], 0, rtol, atol):
                    perm = np.array([[0, -1], [1, 0]])
                    full = np.identity(n, dtype=complex)
                    full[r - 1 : r + 1, r - 1 : r + 1] = perm
                    tmp = full @ cur
                    if r == 1:  # at the top the phase of the entry joins the block
                        perm = np.array([[tmp[0, 0].conjugate(), 0], [0, tmp[0, 0]]]) @ perm
                    out.append(_su2_params(perm.conj().T))
                    full = np.identity(n, dtype=complex)
                    full[r - 1 : r + 1, r - 1 : r + 1] = perm
                    cur = full @ cur
                else:
                    out.append([0, 0, 0])
    else:
        for t in range(n - 1):
            i, j = n - t - 2, n - t - 1
            full = np.identity(n, dtype=complex)
            if t != n - 2:
                cf = np.sqrt(1 - sum(np.abs(cur[k, 0]) ** 2 for k in range(i)))
                Y, Z = cur[i, 0] / cf, cur[j, 0] / cf
                inv = np.array([[np.conj(Y), np.conj(Z)], [-Z, Y]])
            else:
                x = U[0, 0]
                cf = np.sqrt(1 - np.abs(x) ** 2)
                inv = np.array([[np.conj(x), cf], [-cf, x]])
            out.append(_su2_params(inv.conj().T))
            full[i : j + 1, i : j + 1] = inv
            cur = full @ cur
    return out, cur


def _sun_params(U, rtol, atol):
    if U.shape == (3, 3):
        return _su3_params(U)
    stairs, rest = _staircase(U, rtol, atol)
    return stairs + _sun_params(rest[1:, 1:], rtol, atol)


def sun_compact(U, rtol=1e-12, atol=1e-12):
    """([((i, i + 1), [a, b, g]), ...], global phase or None) of the SU(n)
    factorisation (decompositions.py:1073-1155); n >= 3.  U is taken as complex,
    so a real orthogonal U with det -1 gets the complex root det^(-1/n) (the same
    values the reference computes for a complex-typed U) rather than a NaN."""
    U = np.asarray(U, dtype=np.complex128)
    n = U.shape[0]
    if n < 3:
        raise ValueError("sun_compact needs a matrix of size 3 or more, got %d" % n)
    if not np.allclose(U @ U.conj().T, np.identity(n), rtol=rtol, atol=atol):
        raise ValueError("The input matrix is not unitary.")
    det = np.linalg.det(U)
    SU = np.array(U, dtype=complex)
    phase = None
    if not np.isclose(det, 1, rtol=rtol, atol=atol):
        SU = SU * det ** (-1 / n)
        phase = float(np.angle(det))
    flat = _sun_params(SU, rtol, atol)
    out, k = [], 0
    for m2 in range(2, n + 1):
        for m1 in range(n - 1, m2 - 2, -1):
            out.append([(m1 - 1, m1), flat[k]])
            k += 1
    return out, phase


def _sun_gates(U, tol):
    """Gate list of the sun_compact mesh (ops.py:2502-2526): per block, in matrix
    order, R(a/2) and R(-a/2), BS(b/2, 0), R(g/2) and R(-g/2), after the global
    phase R(phi / n) on every mode; the list is then reversed into application
    order."""
    params, phase = sun_compact(U, rtol=tol, atol=tol)
    n = np.asarray(U).shape[0]
    cmds = []
    if phase is not None:
        cmds += [("R", (m,), (phase / n,)) for m in range(n)]
    for (m1, m2), (a, b, g) in params:
        cmds += [("R", (m1,), (a / 2,)), ("R", (m2,), (-a / 2,)), ("BS", (m1, m2), (b / 2, 0.0)),
                 ("R", (m1,), (g / 2,)), ("R", (m2,), (-g / 2,))]
    return cmds[::-1]


# every mesh the reference's Interferometer accepts (ops.py:2638-2646)
MESHES = ("rectangular", "rectangular_phase_end", "rectangular_symmetric", "triangular",
          "rectangular_compact", "triangular_compact", "sun_compact")


# reference tolerances (strawberryfields/ops.py:48-52)
DECOMPOSITION_MERGE_TOL = 1e-13
DECOMPOSITION_TOL = 1e-13


def interferometer_gates(U, drop_identity=True, tol=1e-6, mesh="rectangular"):
    """Gate list [(name, modes, params)] of the mesh of U, in application order
    (ops.py:2655-2718): per first-list block R(phi) and BS(theta, 0), or
    MZgate(phi_in, phi_ex) for the symmetric mesh; the diagonal phases; then, for
    the rectangular mesh, BS(-theta, 0) and R(-phi) per left block in reverse.
    The compact meshes compile to R and BS through their sMZ / SU(2) blocks.

    The triangular mesh is compiled with the same per-block rule as the
    reference compiles it (R(phi), BS(theta, 0)), which realises T where the
    decomposition's identity needs T^{-1}: the reference's own decomposition
    test leaves it out (tests/frontend/test_ops_decompositions.py:32-41), and
    the gate list here is the reference's, not a corrected one."""
    if mesh not in MESHES:
        raise ValueError(f"Unknown mesh '{mesh}'. Mesh must be one of {set(MESHES)}")
    U = np.asarray(U)
    if mesh in ("rectangular_compact", "triangular_compact"):
        # no identity shortcut for these meshes in the reference (ops.py:2663-2673)
        return _compact_gates(mesh, U, tol)
    if mesh == "sun_compact":
        return _sun_gates(U, tol)
    if np.allclose(U, np.identity(len(U)), atol=DECOMPOSITION_MERGE_TOL, rtol=0) and drop_identity:
        return []
    decomp = {"rectangular": rectangular, "rectangular_phase_end": rectangular_phase_end,
              "rectangular_symmetric": rectangular_symmetric, "triangular": triangular}[mesh]
    right, diag, left = decomp(U, tol=tol)
    gates = []

    def clip(x):
        return x if np.abs(x) >= DECOMPOSITION_TOL else 0

    for m, n, theta, phi, _ in right:
        theta, phi = clip(theta), clip(phi)
        if mesh == "rectangular_symmetric":
            gates.append(("MZ", (m, n), (float(np.mod(theta, 2 * np.pi)), float(np.mod(phi, 2 * np.pi)))))
            continue
        if not (drop_identity and phi == 0):
            gates.append(("R", (m,), (phi,)))
        if not (drop_identity and theta == 0):
            gates.append(("BS", (m, n), (theta, 0.0)))
    for m, e in enumerate(diag):
        q = np.log(e).imag if np.abs(e - 1) >= DECOMPOSITION_TOL else 0
        if not (drop_identity and q == 0):
            gates.append(("R", (m,), (float(np.mod(q, 2 * np.pi)),)))
    for m, n, theta, phi, _ in reversed(left or []):
        theta, phi = clip(theta), clip(phi)
        if not (drop_identity and theta == 0):
            gates.append(("BS", (m, n), (-theta, 0.0)))
        if not (drop_identity and phi == 0):
            gates.append(("R", (m,), (-phi,)))
    return gates
