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


MESHES = ("rectangular", "rectangular_phase_end", "rectangular_symmetric", "triangular")
# meshes the reference also accepts, whose phase parametrisations are not built here
COMPACT_MESHES = ("rectangular_compact", "triangular_compact", "sun_compact")


# reference tolerances (strawberryfields/ops.py:48-52)
DECOMPOSITION_MERGE_TOL = 1e-13
DECOMPOSITION_TOL = 1e-13


def interferometer_gates(U, drop_identity=True, tol=1e-6, mesh="rectangular"):
    """Gate list [(name, modes, params)] of the mesh of U, in application order
    (ops.py:2655-2718): per first-list block R(phi) and BS(theta, 0), or
    MZgate(phi_in, phi_ex) for the symmetric mesh; the diagonal phases; then, for
    the rectangular mesh, BS(-theta, 0) and R(-phi) per left block in reverse.

    The triangular mesh is compiled with the same per-block rule as the
    reference compiles it (R(phi), BS(theta, 0)), which realises T where the
    decomposition's identity needs T^{-1}: the reference's own decomposition
    test leaves it out (tests/frontend/test_ops_decompositions.py:32-41), and
    the gate list here is the reference's, not a corrected one."""
    if mesh in COMPACT_MESHES:
        raise NotImplementedError(f"mesh {mesh!r} is not implemented; use one of {MESHES}")
    if mesh not in MESHES:
        raise ValueError(f"Unknown mesh '{mesh}'. Mesh must be one of {set(MESHES + COMPACT_MESHES)}")
    U = np.asarray(U)
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
