"""Seeded circuit generators for the benchmark and parity workloads.

Each generator returns a list of operations ``(name, modes, params)`` with
``name`` in {"R", "K", "D", "S", "BS", "LOSS"}; parameters are drawn from one
``numpy.random.default_rng(seed)`` in circuit order (within a gate, in the
order of the backend method's signature), so the HIP path, the CPU oracle and
the comparators see identical values.

The layered circuit is the continuous-variable neural-network layer of the
reference example ``examples/quantum_neural_network.py:14-85``: a rectangular
beamsplitter mesh (pairs (k, k+1) with (l + k) even, for l in range(N)) and
rotations, squeezers, a second interferometer, displacements and Kerr gates.
Here the interferometer rotates every mode (N rotations, e.g. 6 BSgate +
4 Rgate at N = 4) and squeezers and displacements take both parameters.
"""
import numpy as np

SEEDS = {"c1": 20261201, "c2": 20261202, "c3": 20261203, "c4": 20261204, "c5": 20261205}


def mesh_pairs(N):
    """Beamsplitter pairs of the rectangular mesh, in application order."""
    pairs = []
    for l in range(N):
        for k in range(N - 1):
            if (l + k) % 2 != 1:
                pairs.append((k, k + 1))
    return pairs


def _interferometer(rng, N, ops, sd):
    for (a, b) in mesh_pairs(N):
        ops.append(("BS", (a, b), (rng.uniform(0, np.pi / 2), rng.normal(0, sd))))
    for i in range(N):
        ops.append(("R", (i,), (rng.normal(0, sd),)))


def layered(N, L, rng, sd=0.1, loss=False):
    """L layers of interferometer -> S -> interferometer -> [loss] -> D -> K."""
    ops = []
    for _ in range(L):
        _interferometer(rng, N, ops, sd)
        for i in range(N):
            ops.append(("S", (i,), (rng.normal(0, sd), rng.normal(0, sd))))
        _interferometer(rng, N, ops, sd)
        if loss:
            for i in range(N):
                ops.append(("LOSS", (i,), (rng.uniform(0.7, 1.0),)))
        for i in range(N):
            ops.append(("D", (i,), (rng.normal(0, sd), rng.normal(0, sd))))
        for i in range(N):
            ops.append(("K", (i,), (rng.normal(0, sd),)))
    return ops


def config1():
    """3-mode pure ket, D = 8: S, D, R per mode, BS on (0,1), (1,2), (0,1), K per mode."""
    rng = np.random.default_rng(SEEDS["c1"])
    ops = []
    for i in range(3):
        ops.append(("S", (i,), (rng.uniform(0, 0.5), rng.uniform(0, 2 * np.pi))))
        ops.append(("D", (i,), (rng.uniform(0, 0.5), rng.uniform(0, 2 * np.pi))))
        ops.append(("R", (i,), (rng.uniform(0, 2 * np.pi),)))
    for (a, b) in [(0, 1), (1, 2), (0, 1)]:
        ops.append(("BS", (a, b), (rng.uniform(0, np.pi / 2), rng.uniform(0, 2 * np.pi))))
    for i in range(3):
        ops.append(("K", (i,), (rng.uniform(0, 0.2),)))
    return dict(name="c1", n=3, D=8, pure=True, ops=ops)


def config2():
    """4-mode pure ket, D = 15, 8 layers (256 operations)."""
    rng = np.random.default_rng(SEEDS["c2"])
    return dict(name="c2", n=4, D=15, pure=True, ops=layered(4, 8, rng))


def target_state(n, D, nmax, rng):
    """Seeded complex normal vector, zero where the total photon number exceeds nmax, normalised."""
    v = rng.normal(size=D**n) + 1j * rng.normal(size=D**n)
    idx = np.indices([D] * n).reshape(n, -1).sum(axis=0)
    v[idx > nmax] = 0.0
    v /= np.linalg.norm(v)
    return v.reshape([D] * n)


def config3():
    """6-mode pure ket, D = 8, 4 layers; loss 1 - |<target|psi>|^2."""
    rng = np.random.default_rng(SEEDS["c3"])
    ops = layered(6, 4, rng)
    target = target_state(6, 8, 6, rng)
    return dict(name="c3", n=6, D=8, pure=True, ops=ops, target=target)


def config4(count=256):
    """256 independent pure circuits of uneven size: n in {2..5}, D in {6..20}, D^n <= 2^20."""
    rng = np.random.default_rng(SEEDS["c4"])
    circuits = []
    for _ in range(count):
        while True:
            n = int(rng.integers(2, 6))
            D = int(rng.integers(6, 21))
            if D**n <= 2**20:
                break
        L = int(rng.integers(2, 7))
        circuits.append(dict(n=n, D=D, pure=True, ops=layered(n, L, rng)))
    return dict(name="c4", circuits=circuits)


def config5():
    """3-mode DM, D = 20, 6 layers with a loss channel per mode after the second interferometer."""
    rng = np.random.default_rng(SEEDS["c5"])
    return dict(name="c5", n=3, D=20, pure=False, ops=layered(3, 6, rng, loss=True), layer_len=24)


def n_params(ops):
    """Number of real gate parameters of an op list."""
    return sum(len(p) for _, _, p in ops)
