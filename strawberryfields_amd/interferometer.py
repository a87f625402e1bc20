"""Rectangular (Clements) decomposition of an interferometer into the gates of
this backend.

``Interferometer(U)`` of the front end (``strawberryfields/ops.py:2556-2718``)
is compiled, as in the reference, into rotations and beamsplitters that the
HIP kernels then run.  The mesh follows Clements et al., Optica 3, 1460
(2016), as ``strawberryfields/decompositions.py:233-339`` states it: the
two-mode block

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


# reference tolerances (strawberryfields/ops.py:48-52)
DECOMPOSITION_MERGE_TOL = 1e-13
DECOMPOSITION_TOL = 1e-13


def interferometer_gates(U, drop_identity=True, tol=1e-6):
    """Gate list [(name, modes, params)] of the rectangular mesh of U, in
    application order (ops.py:2655-2718 with mesh "rectangular"):
    R(phi) and BS(theta, 0) per right block, the diagonal phases, then
    BS(-theta, 0) and R(-phi) per left block in reverse."""
    U = np.asarray(U)
    if np.allclose(U, np.identity(len(U)), atol=DECOMPOSITION_MERGE_TOL, rtol=0) and drop_identity:
        return []
    right, diag, left = rectangular(U, tol=tol)
    gates = []

    def clip(x):
        return x if np.abs(x) >= DECOMPOSITION_TOL else 0

    for m, n, theta, phi, _ in right:
        theta, phi = clip(theta), clip(phi)
        if not (drop_identity and phi == 0):
            gates.append(("R", (m,), (phi,)))
        if not (drop_identity and theta == 0):
            gates.append(("BS", (m, n), (theta, 0.0)))
    for m, e in enumerate(diag):
        q = np.log(e).imag if np.abs(e - 1) >= DECOMPOSITION_TOL else 0
        if not (drop_identity and q == 0):
            gates.append(("R", (m,), (float(np.mod(q, 2 * np.pi)),)))
    for m, n, theta, phi, _ in reversed(left):
        theta, phi = clip(theta), clip(phi)
        if not (drop_identity and theta == 0):
            gates.append(("BS", (m, n), (-theta, 0.0)))
        if not (drop_identity and phi == 0):
            gates.append(("R", (m,), (-phi,)))
    return gates
