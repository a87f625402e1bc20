"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Tolerances: complex128 max|d psi| <= 1e-12 (1e-11 for long circuits),
complex64 <= 1e-5; gradients against central finite differences of the
complex128 oracle.
"""
import itertools
import os

import numpy as np
import pytest
import torch

from oracle import fock_oracle as fo
from strawberryfields_amd import HIPFockBackend, circuits, frontend, runtime

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
C128, C64 = torch.complex128, torch.complex64
TOL = {C128: 1e-12, C64: 1e-5}
DEV = "cuda:0"


def apply(be, name, modes, p):
    if name == "R":
        be.rotation(p[0], modes[0])
    elif name == "K":
        be.kerr_interaction(p[0], modes[0])
    elif name == "D":
        be.displacement(p[0], p[1], modes[0])
    elif name == "S":
        be.squeeze(p[0], p[1], modes[0])
    elif name == "BS":
        be.beamsplitter(p[0], p[1], modes[0], modes[1])
    elif name == "LOSS":
        be.loss(p[0], modes[0])
    elif name == "S2":
        be.two_mode_squeeze(p[0], p[1], modes[0], modes[1])
    elif name == "MZ":
        be.mzgate(p[0], p[1], modes[0], modes[1])
    elif name == "CK":
        be.cross_kerr_interaction(p[0], modes[0], modes[1])
    elif name == "V":
        be.cubic_phase(p[0], modes[0])
    else:
        raise ValueError(name)


def run_backend(ops, n, D, dtype, pure=True, init=None, **kw):
    be = HIPFockBackend()
    be.begin_circuit(n, cutoff_dim=D, pure=pure, dtype=dtype, device=DEV, **kw)
    if init is not None:
        if init.ndim == n:
            be.prepare_ket_state(init, list(range(n)))
        else:
            be.prepare_dm_state(init, list(range(n)))
    for name, modes, p in ops:
        apply(be, name, modes, p)
    return be


def to_np(t):
    return t.detach().cpu().numpy()


def rand_state(n, D, pure, seed):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=[D] * n) + 1j * rng.normal(size=[D] * n)
    psi /= np.linalg.norm(psi)
    return psi if pure else fo.mix(psi, n) * 0.7 + 0.3 * fo.mix(np.roll(psi, 1), n)


# ------------------------------------------------------------------ gate tables
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("D", [1, 2, 5, 12, 20])
def test_single_mode_gate_tables(dtype, D):
    tol = 1e-13 if dtype == C128 else 1e-6
    for kind, fn, params in [
        ("D", fo.displacement, [(0.3, 0.7), (-0.45, 2.1), (0.9, -1.3)]),
        ("S", fo.squeezing, [(0.3, 0.7), (-0.4, 2.1), (0.8, -1.3)]),
    ]:
        for p0, p1 in params:
            t = to_np(runtime.build_gate(kind, p0, p1, D, dtype, DEV))
            DD = D * D
            U = t[:DD].reshape(D, D)
            assert np.abs(U - fn(p0, p1, D)).max() < tol, (kind, p0, p1)
            h = 1e-6
            d0 = (fn(p0 + h, p1, D) - fn(p0 - h, p1, D)) / (2 * h)
            d1 = (fn(p0, p1 + h, D) - fn(p0, p1 - h, D)) / (2 * h)
            dtol = 1e-8 if dtype == C128 else 1e-5
            assert np.abs(t[DD : 2 * DD].reshape(D, D) - d0).max() < dtol, (kind, "d0")
            assert np.abs(t[2 * DD :].reshape(D, D) - d1).max() < dtol, (kind, "d1")
    for kind, fn in [("R", fo.phase), ("K", fo.kerr)]:
        t = to_np(runtime.build_gate(kind, 0.37, 0.0, D, dtype, DEV))
        assert np.abs(t[:D] - np.diag(fn(0.37, D))).max() < tol


def _bs_dense_from_blocks(blk, D):
    mat = np.zeros((D,) * 4, dtype=complex)
    off = 0
    for s in range(2 * D - 1):
        lo, hi = max(0, s - D + 1), min(s, D - 1)
        c = hi - lo + 1
        b = blk[off : off + c * c].reshape(c, c)
        for a in range(c):
            for k in range(c):
                N, n = lo + a, lo + k
                mat[N, n, s - N, s - n] = b[a, k]
        off += c * c
    return mat, off


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("D", [1, 2, 6, 15, 20])
def test_beamsplitter_and_loss_tables(dtype, D):
    # FP64 block recurrences carry ~1e-12 rounding at cutoff 20 (vs 50-digit mpmath)
    tol = (1e-13 if D < 15 else 3e-12) if dtype == C128 else 1e-6
    for th, ph in [(0.6, 0.4), (-0.35, 1.7), (1.4, -2.2)]:
        t = to_np(runtime.build_gate("BS", th, ph, D, dtype, DEV))
        B, nb = _bs_dense_from_blocks(t, D)
        assert np.abs(B - fo.beamsplitter(th, ph, D)).max() < tol
        # fourth-order central difference; h large enough that the oracle's own
        # ~1e-12 rounding at cutoff 20 stays below 1e-7 in the quotient
        h = 1e-3
        f = lambda x: fo.beamsplitter(x, ph, D)  # noqa: E731
        dB = (-f(th + 2 * h) + 8 * f(th + h) - 8 * f(th - h) + f(th - 2 * h)) / (12 * h)
        dBg, _ = _bs_dense_from_blocks(t[nb:], D)
        # relative to the largest derivative entry: the difference quotient's
        # truncation error grows with the photon number (powers up to 2D - 2)
        scale = max(1.0, np.abs(dB).max())
        assert np.abs(dBg - dB).max() < (1e-8 if dtype == C128 and D < 15 else 1e-6) * scale
    g = np.load(os.path.join(GOLDEN, "gates_mpmath.npz"))
    for key in [k for k in g.files if k.startswith("BS") and not k.endswith("_params")]:
        th, ph, Dg = g[key + "_params"]
        if int(Dg) == D:
            t = to_np(runtime.build_gate("BS", th, ph, D, dtype, DEV))
            assert np.abs(_bs_dense_from_blocks(t, D)[0] - g[key]).max() < tol, key
    for T in [0.0, 0.42, 1.0]:
        t = runtime.build_gate("LOSS", T, 0.0, D, dtype, DEV)
        c = to_np(torch.view_as_real(t).reshape(-1)[: D**3]).reshape(D, D, D)
        S = fo.loss_superop(T, D)  # S[p, p+l, q+l, q]
        ref = np.zeros((D, D, D))
        for p, q, l in itertools.product(range(D), repeat=3):
            if p + l < D and q + l < D:
                ref[p, q, l] = S[p, p + l, q + l, q]
        assert np.abs(c - ref).max() < tol * 10


# ------------------------------------------------------------------ single operations
SINGLE = [
    ("R", (0.7,)),
    ("K", (-0.3,)),
    ("D", (0.4, 1.2)),
    ("D", (-0.25, -0.6)),
    ("S", (0.3, -0.8)),
    ("S", (-0.2, 0.5)),
    ("LOSS", (0.63,)),
]


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure", [True, False])
@pytest.mark.parametrize("n,D", [(1, 1), (1, 6), (2, 2), (3, 5)])
def test_single_mode_ops(dtype, pure, n, D):
    init = rand_state(n, D, pure, seed=n * 100 + D)
    for (name, p), mode in itertools.product(SINGLE, range(n)):
        ops = [(name, (mode,), p)]
        be = run_backend(ops, n, D, dtype, pure=pure, init=init)
        st = be.state()
        ref, rpure = fo.run_ops(ops, n, D, state=init)
        assert st.is_pure == rpure
        got = to_np(st.ket() if rpure else st.dm())
        assert np.abs(got - ref).max() < TOL[dtype] * max(1, np.abs(ref).max()), (name, p, mode)


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure", [True, False])
@pytest.mark.parametrize("n,D", [(2, 1), (2, 4), (3, 5), (4, 3)])
def test_beamsplitter_all_mode_pairs(dtype, pure, n, D):
    init = rand_state(n, D, pure, seed=7 + n + D)
    for m1, m2 in itertools.permutations(range(n), 2):
        for p in [(0.7, 0.9), (-1.1, 2.4)]:
            ops = [("BS", (m1, m2), p)]
            be = run_backend(ops, n, D, dtype, pure=pure, init=init)
            ref, rpure = fo.run_ops(ops, n, D, state=init)
            st = be.state()
            got = to_np(st.ket() if rpure else st.dm())
            assert np.abs(got - ref).max() < TOL[dtype], (m1, m2, p)


@pytest.mark.parametrize("D", [4, 9, 12, 17])
def test_dm_two_mode_gates_in_one_pass_match_oracle(D):
    """A two-mode gate on a DM whose pair is followed by another mode runs in
    one pass (rho' = G rho G^H on matrix cores): every gate of the beamsplitter
    family, both mode orders, blocks of one, two and three 16-row blocks (the
    last at D = 17, a 24M-entry DM), against the complex128 oracle's two
    passes."""
    n = 3
    init = rand_state(n, D, False, seed=300 + D)
    for kind, p in [("BS", (0.7, 0.9)), ("BS", (-1.1, 2.4)), ("MZ", (1.1, 0.4)), ("S2", (0.3, 1.2)), ("CK", (0.6,))]:
        for modes in [(0, 1), (1, 0)]:
            ops = [(kind, modes, p)]
            be = run_backend(ops, n, D, C64, pure=False, init=init)
            ref, _ = fo.run_ops(ops, n, D, state=init)
            got = to_np(be.state().dm())
            assert np.abs(got - ref).max() < TOL[C64], (kind, modes, p)


def test_loss_extremes_and_vacuum():
    D = 8
    for T in [0.0, 1.0]:
        be = run_backend([("D", (0,), (0.6, 0.3)), ("LOSS", (0,), (T,))], 1, D, C128)
        ref, _ = fo.run_ops([("D", (0,), (0.6, 0.3)), ("LOSS", (0,), (T,))], 1, D)
        assert np.abs(to_np(be.state().dm()) - ref).max() < 1e-13
    # full loss maps rho to Tr(rho)|0><0|; the truncated coherent state has Tr < 1
    be = run_backend([("D", (0,), (0.1, 0.3)), ("LOSS", (0,), (0.0,))], 1, D, C128)
    assert be.is_vacuum(1e-12)
    be = run_backend([("LOSS", (0,), (0.5,))], 1, D, C128)
    assert be.is_vacuum(1e-14)


# ------------------------------------------------------------------ circuits
def _fidelity(a, b):
    a, b = a.reshape(-1), b.reshape(-1)
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


@pytest.mark.parametrize("dtype", [C128, C64])
def test_config1_through_engine_matches_golden(dtype):
    c1 = circuits.config1()
    g = np.load(os.path.join(GOLDEN, "c1.npz"))
    prog = frontend.Program(3)
    gates = {"S": frontend.Sgate, "D": frontend.Dgate, "R": frontend.Rgate, "K": frontend.Kgate, "BS": frontend.BSgate}
    with prog.context as q:
        for name, modes, p in c1["ops"]:
            gates[name](*p) | tuple(q[m] for m in modes)
    eng = frontend.Engine(HIPFockBackend(), backend_options={"cutoff_dim": 8, "dtype": dtype, "device": DEV})
    state = eng.run(prog).state
    ket = to_np(state.ket())
    assert np.abs(ket - g["ket"]).max() < TOL[dtype]
    assert 1 - _fidelity(ket, g["ket"]) < (1e-13 if dtype == C128 else 1e-6)
    probs = to_np(state.all_fock_probs())
    assert np.abs(probs - g["probs"]).max() < TOL[dtype]


@pytest.mark.parametrize("dtype", [C128, C64])
def test_config2_forward(dtype):
    c2 = circuits.config2()
    ref, _ = fo.run_ops(c2["ops"], 4, 15)
    be = run_backend(c2["ops"], 4, 15, dtype)
    st = be.state()
    ket = to_np(st.ket())
    tol = 1e-12 if dtype == C128 else 1e-5
    assert np.abs(ket - ref).max() < tol
    assert abs(st.trace() - fo.trace(ref, 4, True)) < tol
    for idx in [(0, 0, 0, 0), (1, 0, 2, 1), (3, 3, 0, 1)]:
        assert abs(st.fock_prob(list(idx)) - fo.fock_prob(ref, 4, True, idx)) < tol


def test_config3_forward_complex64():
    c3 = circuits.config3()
    ref, _ = fo.run_ops(c3["ops"], 6, 8)
    st = run_backend(c3["ops"], 6, 8, C64).state()
    ket = to_np(st.ket())
    assert np.abs(ket - ref).max() < 1e-5
    assert 1 - _fidelity(ket, ref) < 1e-6


def test_mixed_layer_with_loss_matches_oracle():
    """One layer of the checkpointed mixed-state workload at reduced cutoff."""
    rng = np.random.default_rng(5)
    ops = circuits.layered(3, 1, rng, loss=True)
    for dtype in (C128, C64):
        be = run_backend(ops, 3, 6, dtype, pure=False)
        ref, _ = fo.run_ops(ops, 3, 6, pure=False)
        assert np.abs(to_np(be.state().dm()) - ref).max() < TOL[dtype]


# ------------------------------------------------------------------ readout
@pytest.mark.parametrize("pure", [True, False])
def test_readouts(pure):
    n, D = 4, 4
    rng = np.random.default_rng(9)
    ops = circuits.layered(n, 1, rng)
    init = None if pure else fo.vacuum(n, D, pure=False)
    be = run_backend(ops, n, D, C128, pure=pure)
    ref, rpure = fo.run_ops(ops, n, D, pure=pure)
    st = be.state()
    assert abs(st.trace() - fo.trace(ref, n, rpure)) < 1e-13
    assert np.abs(to_np(st.all_fock_probs()) - fo.all_fock_probs(ref, n, rpure)).max() < 1e-14
    for modes in ([0], [3], [1, 2], [0, 2, 3]):
        assert np.abs(to_np(st.reduced_dm(modes)) - fo.reduced_dm(ref, n, rpure, modes)).max() < 1e-13
    for m in range(n):
        mean, var = st.mean_photon(m)
        rm, rv = fo.mean_photon(ref, n, rpure, m)
        assert abs(mean - rm) < 1e-12 and abs(var - rv) < 1e-12
    vals = np.linspace(-1, 2, D)
    for modes in ([1], [0, 3], [2, 1, 0]):
        assert abs(st.diagonal_expectation(modes, vals) - fo.diagonal_expectation(ref, n, rpure, modes, vals)) < 1e-12
    # reduced state through backend.state(modes) in requested order
    sub = be.state(modes=[2, 0])
    r = fo.reduced_dm(ref, n, rpure, [0, 2]).transpose(2, 3, 0, 1)
    assert np.abs(to_np(sub.dm()) - r).max() < 1e-13
    del init


def test_prepare_vacuum_on_pure_multimode_state_becomes_mixed():
    n, D = 3, 5
    ops = [("D", (0,), (0.5, 0.2)), ("BS", (0, 1), (0.7, 0.1)), ("S", (2,), (0.3, 0.0))]
    be = run_backend(ops, n, D, C128)
    be.prepare_vacuum_state(1)
    st = be.state()
    assert not st.is_pure
    c = fo.Circuit(n, D)
    for name, modes, p in ops:
        getattr(c, {"D": "displacement", "BS": "beamsplitter", "S": "squeeze"}[name])(*p, *modes)
    c.prepare_mode_fock(0, 1)
    assert np.abs(to_np(st.dm()) - c._state).max() < 1e-13


# ------------------------------------------------------------------ gradients
def _loss_pure(state, target):
    ov = torch.vdot(target.reshape(-1), state.reshape(-1))
    return 1 - ov.abs() ** 2


def _run_params(ops, n, D, dtype, pure, params, segment=None):
    be = HIPFockBackend()
    be.begin_circuit(n, cutoff_dim=D, pure=pure, dtype=dtype, device=DEV, checkpoint_segment=segment)
    k = 0
    for name, modes, p in ops:
        vals = [params[k + j] for j in range(len(p))]
        k += len(p)
        apply(be, name, modes, vals)
    return be.state()


def _oracle_loss(ops, n, D, pure, flat, target, weights):
    k = 0
    new = []
    for name, modes, p in ops:
        new.append((name, modes, tuple(flat[k : k + len(p)])))
        k += len(p)
    st, rp = fo.run_ops(new, n, D, pure=pure)
    if target is not None:
        return 1 - abs(np.vdot(target.reshape(-1), st.reshape(-1))) ** 2
    return float((fo.all_fock_probs(st, n, rp) * weights).sum())


def _fd_grad(ops, n, D, pure, target, weights, h=1e-5):
    flat = np.array([v for _, _, p in ops for v in p], dtype=float)
    g = np.zeros_like(flat)
    for i in range(len(flat)):
        a, b = flat.copy(), flat.copy()
        a[i] += h
        b[i] -= h
        g[i] = (_oracle_loss(ops, n, D, pure, a, target, weights) - _oracle_loss(ops, n, D, pure, b, target, weights)) / (2 * h)
    return flat, g


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure", [True, False])
def test_reduced_dm_gradient_vs_finite_differences(dtype, pure):
    """reduced_dm is differentiable, as the TF backend's is
    (tfbackend/states.py:307): the gradient of a real cost Re sum conj(W) rho_02
    of the reduced DM of modes (0, 2), W a seeded complex array, through the
    circuit's adjoint and the reduced DM's backward, against central
    differences of the complex128 oracle; the forward value against the
    oracle's reduced DM."""
    n, D = 3, 4
    rng = np.random.default_rng(77 + int(pure))
    ops = circuits.layered(n, 1, rng, sd=0.3, loss=not pure)
    W = rng.normal(size=[D] * 4) + 1j * rng.normal(size=[D] * 4)

    def cost(flat):
        k, new = 0, []
        for name, modes, p in ops:
            new.append((name, modes, tuple(flat[k : k + len(p)])))
            k += len(p)
        st, rp = fo.run_ops(new, n, D, pure=pure)
        return float(np.real(np.sum(np.conj(W) * fo.reduced_dm(st, n, rp, [0, 2]))))

    flat = np.array([v for _, _, p in ops for v in p], dtype=float)
    h = 1e-5
    g_fd = np.array([(cost(flat + h * e) - cost(flat - h * e)) / (2 * h) for e in np.eye(len(flat))])
    params = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    st = _run_params(ops, n, D, dtype, pure, params)
    red = st.reduced_dm([0, 2])
    st_ref, rp = fo.run_ops(ops, n, D, pure=pure)
    assert np.abs(to_np(red) - fo.reduced_dm(st_ref, n, rp, [0, 2])).max() < TOL[dtype]
    Wt = torch.as_tensor(W, dtype=dtype, device=DEV)
    (Wt.conj() * red).sum().real.to(torch.float64).backward()
    g = params.grad.numpy()
    rel = np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd)
    assert rel < (1e-8 if dtype == C128 else 1e-5), rel


@pytest.mark.parametrize("dtype", [C128, C64])
def test_gradient_pure_layer_vs_finite_differences(dtype):
    n, D = 3, 6
    rng = np.random.default_rng(21)
    ops = circuits.layered(n, 1, rng, sd=0.3)
    target = circuits.target_state(n, D, 4, rng)
    flat, g_fd = _fd_grad(ops, n, D, True, target, None)
    params = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    st = _run_params(ops, n, D, dtype, True, params)
    loss = _loss_pure(st.ket(), torch.as_tensor(target, dtype=dtype, device=DEV))
    loss.backward()
    g = params.grad.numpy()
    rel = np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd)
    assert rel < (1e-8 if dtype == C128 else 1e-5), rel


@pytest.mark.parametrize("dtype", [C128, C64])
def test_gradient_mixed_with_loss_vs_finite_differences(dtype):
    n, D = 2, 6
    rng = np.random.default_rng(22)
    ops = circuits.layered(n, 1, rng, sd=0.3, loss=True)
    idx = np.indices([D] * n).sum(axis=0)
    weights = (idx - n) ** 2.0
    for pure in (True, False):  # pure start: the first loss converts the ket
        flat, g_fd = _fd_grad(ops, n, D, pure, None, weights)
        params = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
        st = _run_params(ops, n, D, dtype, pure, params)
        loss = (st.all_fock_probs().to(torch.float64) * torch.as_tensor(weights, device=DEV)).sum()
        loss.backward()
        g = params.grad.numpy()
        rel = np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd)
        assert rel < (1e-8 if dtype == C128 else 1e-5), (pure, rel)


def test_checkpoint_segments_give_identical_gradients():
    n, D = 3, 5
    rng = np.random.default_rng(23)
    ops = circuits.layered(n, 2, rng, loss=True)
    flat = np.array([v for _, _, p in ops for v in p], dtype=float)
    grads = []
    for seg in (None, 1, 7, 24):
        params = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
        st = _run_params(ops, n, D, C64, False, params, segment=seg)
        st.all_fock_probs().sum().backward()
        grads.append(params.grad.numpy().copy())
    for g in grads[1:]:
        assert np.array_equal(g, grads[0])


def test_mean_photon_gradient_closed_form():
    """test_tf_integration.py: d<n>/dr of a coherent state is 2r."""
    r = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    be = HIPFockBackend()
    be.begin_circuit(1, cutoff_dim=25, dtype=C128, device=DEV)
    be.displacement(r, 0.4, 0)
    mean, _ = be.state().mean_photon(0)
    mean.backward()
    assert abs(float(mean.detach()) - 0.09) < 1e-12
    assert abs(float(r.grad) - 0.6) < 1e-10


# ------------------------------------------------------------------ fused stages
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,D,fuse", [(3, 5, 2), (3, 5, 3), (4, 4, 4), (6, 3, 3), (5, 6, 3), (2, 12, 2)])
def test_fused_stages_match_unfused_and_oracle(dtype, n, D, fuse):
    rng = np.random.default_rng(100 + n * D)
    ops = circuits.layered(n, 2, rng, sd=0.3)
    flat = np.array([v for _, _, p in ops for v in p], dtype=float)
    target = torch.as_tensor(circuits.target_state(n, D, 3, rng), dtype=dtype, device=DEV)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    ref, _ = fo.run_ops(ops, n, D)
    out = {}
    for f in (0, fuse):
        cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, False, dtype, DEV, fuse=f)
        if f:
            assert cc.nsteps < len(ops)
        params = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
        vac = torch.zeros([D] * n, dtype=dtype, device=DEV)
        vac.view(-1)[0] = 1
        psi = runtime.CircuitFunction.apply(params, vac, cc, None)
        ov = torch.vdot(target.reshape(-1), psi.reshape(-1))
        (1 - ov.abs() ** 2).backward()
        out[f] = (to_np(psi), params.grad.cpu().numpy())
        assert np.abs(out[f][0] - ref).max() < TOL[dtype]
    g0, gf = out[0][1], out[fuse][1]
    rel = np.linalg.norm(g0 - gf) / np.linalg.norm(g0)
    assert rel < (1e-10 if dtype == C128 else 1e-5), rel
    del flat


@pytest.mark.parametrize("D", [9, 11, 12, 13, 15, 16])
def test_fused_single_gate_adjoint_matches_unfused_at_large_cutoffs(D):
    """One gate per stage on 3 local modes of 4, every kind and both beamsplitter
    mode orders, random state and cotangent: the fused stage's lam_in and
    gradient equal the one-kernel-per-gate adjoint.  At these cutoffs a thread
    owns 1-4 tile amplitudes (the tile has D^3 > 1024 amplitudes from D = 11),
    and at D = 12-15 a compiler rewrite of the beamsplitter's fiber index into
    byte dot products once broke the first amplitude's cotangent reads."""
    n = 4
    g = torch.Generator(device="cpu").manual_seed(D)
    for op in [("BS", (1, 2)), ("BS", (3, 1)), ("S", (2,)), ("D", (1,)), ("R", (3,)), ("K", (2,))]:
        res = {}
        for fuse in (0, 3):
            cc = runtime.CompiledCircuit([op], n, D, False, C64, DEV, fuse=fuse)
            assert cc.nsteps == 1
            cc.build_table(torch.tensor([[0.4, 0.3]], dtype=torch.float64, device=DEV))
            st = torch.randn([D] * n, dtype=C64, generator=g).to(DEV) if fuse == 0 else res[0][3]
            lam = torch.randn([D] * n, dtype=C64, generator=g).to(DEV) if fuse == 0 else res[0][4]
            out = cc.new_state(1)
            cc.forward(0, 1, st, out=out, saved=[st])
            gin = cc.new_state(0)
            grads = torch.zeros(2, dtype=torch.float64, device=DEV)
            cc.adjoint(0, 1, [st], lam, gin, grads)
            torch.cuda.synchronize()
            res[fuse] = (to_np(out), to_np(gin), grads.cpu().numpy(), st, lam)
        assert np.abs(res[0][0] - res[3][0]).max() < 1e-5, op
        dl = np.abs(res[0][1] - res[3][1]).max()
        assert np.isfinite(dl) and dl < 1e-5, (op, dl)
        dg = np.abs(res[0][2] - res[3][2]).max() / max(1.0, np.abs(res[0][2]).max())
        assert dg < 1e-5, (op, res[0][2], res[3][2])


def test_fused_gradient_vs_finite_differences_c128():
    n, D = 4, 5
    rng = np.random.default_rng(31)
    ops = circuits.layered(n, 1, rng, sd=0.3)
    target = circuits.target_state(n, D, 4, rng)
    flat, g_fd = _fd_grad(ops, n, D, True, target, None)
    params = torch.tensor(flat, dtype=torch.float64, requires_grad=True)
    be = HIPFockBackend()
    be.begin_circuit(n, cutoff_dim=D, dtype=C128, device=DEV)
    k = 0
    for name, modes, p in ops:
        apply(be, name, modes, [params[k + j] for j in range(len(p))])
        k += len(p)
    st = be.state()
    loss = _loss_pure(st.ket(), torch.as_tensor(target, dtype=C128, device=DEV))
    loss.backward()
    g = params.grad.numpy()
    assert np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd) < 1e-8


# ------------------------------------------------------------------ captured evaluation
def test_graph_evaluator_matches_eager_over_replays():
    """Replays of the captured gradient evaluation equal the eager autograd result
    for several parameter sets, with eager work and synchronisations between replays."""
    n, D = 4, 6
    rng = np.random.default_rng(41)
    ops = circuits.layered(n, 2, rng, sd=0.3)
    target = torch.as_tensor(circuits.target_state(n, D, 4, rng), dtype=C64, device=DEV)
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, False, C64, DEV)
    vac = torch.zeros([D] * n, dtype=C64, device=DEV)
    vac.view(-1)[0] = 1
    ev = runtime.GradEvaluator(cc, vac, target, use_graph=True)
    for trial in range(4):
        rows = np.zeros((len(ops), 2))
        for i, (_, _, p) in enumerate(ops):
            rows[i, : len(p)] = np.array(p) * (1 + 0.1 * trial)
        params = torch.tensor(rows, dtype=torch.float64, device=DEV)
        loss_g, grads_g = ev(params)
        loss_g, grads_g = float(loss_g), grads_g.clone()
        p_leaf = params.clone().requires_grad_(True)
        psi = runtime.CircuitFunction.apply(p_leaf, vac, cc, None)
        loss = runtime.fidelity_loss(psi, target)
        loss.backward()
        torch.cuda.synchronize()
        assert abs(loss_g - float(loss.detach())) < 1e-12
        assert torch.equal(grads_g, p_leaf.grad)
        ov = runtime.overlap(target, psi.detach())
        ref = torch.vdot(target.reshape(-1).to(torch.complex128), psi.detach().reshape(-1).to(torch.complex128))
        assert abs(complex(ov) - complex(ref)) < 1e-6


def test_evaluator_replay_between_forward_and_backward_keeps_gradients():
    """A captured evaluation of the same plan with other parameters, run between
    an autograd forward and its backward, rewrites the shared gate table; the
    backward must still use the forward's own parameters."""
    n, D = 4, 6
    rng = np.random.default_rng(43)
    ops = circuits.layered(n, 1, rng, sd=0.3)
    target = torch.as_tensor(circuits.target_state(n, D, 4, rng), dtype=C64, device=DEV)
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, False, C64, DEV)
    vac = torch.zeros([D] * n, dtype=C64, device=DEV)
    vac.view(-1)[0] = 1
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    p = torch.tensor(rows, dtype=torch.float64, device=DEV)
    q = p * 1.7 + 0.05
    ref_leaf = p.clone().requires_grad_(True)
    runtime.fidelity_loss(runtime.CircuitFunction.apply(ref_leaf, vac, cc, None), target).backward()
    ev = runtime.GradEvaluator(cc, vac, target, use_graph=True)
    leaf = p.clone().requires_grad_(True)
    loss = runtime.fidelity_loss(runtime.CircuitFunction.apply(leaf, vac, cc, None), target)
    _, gq = ev(q)  # replays table build + sweeps with q on the shared plan
    gq = gq.clone()
    loss.backward()
    torch.cuda.synchronize()
    assert torch.equal(leaf.grad, ref_leaf.grad)
    assert not torch.equal(gq, ref_leaf.grad)


def test_fidelity_loss_gradient_matches_torch():
    rng = np.random.default_rng(3)
    psi = torch.tensor(rng.normal(size=300) + 1j * rng.normal(size=300), dtype=C128, device=DEV, requires_grad=True)
    t = torch.tensor(rng.normal(size=300) + 1j * rng.normal(size=300), dtype=C128, device=DEV)
    l1 = runtime.fidelity_loss(psi, t)
    (g1,) = torch.autograd.grad(l1, psi)
    ov = torch.vdot(t, psi)
    l2 = 1 - ov.real**2 - ov.imag**2
    (g2,) = torch.autograd.grad(l2, psi)
    assert abs(float(l1) - float(l2)) < 1e-9 * abs(float(l2))
    assert (g1 - g2).abs().max() < 1e-9 * g2.abs().max()


def test_tf_path_comparator_gradient_matches_adjoint():
    """The TF-path restatement (dense einsums + autograd) and the hand-written
    adjoint give the same gradient (complex128)."""
    from strawberryfields_amd import tfpath

    n, D = 3, 6
    rng = np.random.default_rng(51)
    ops = circuits.layered(n, 1, rng, sd=0.3)
    target = torch.as_tensor(circuits.target_state(n, D, 4, rng), dtype=C128, device=DEV)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    p1 = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    psi = tfpath.TFPathCircuit([(o[0], o[1]) for o in ops], n, D, C128, DEV).forward(p1)
    ov = torch.vdot(target.reshape(-1), psi.reshape(-1))
    (1 - ov.abs() ** 2).backward()
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, False, C128, DEV)
    p2 = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    vac = torch.zeros([D] * n, dtype=C128, device=DEV)
    vac.view(-1)[0] = 1
    psi2 = runtime.CircuitFunction.apply(p2, vac, cc, None)
    runtime.fidelity_loss(psi2, target).backward()
    assert (psi.detach() - psi2.detach()).abs().max() < 1e-12
    g1, g2 = p1.grad, p2.grad
    assert float((g1 - g2).norm() / g2.norm()) < 1e-10


def test_tf_path_comparator_mixed_with_loss_matches_adjoint():
    """The TF-path restatement on a density matrix (mix, two-sided gates, loss
    superoperator differentiated by autograd) against the adjoint kernels, and
    its forward against the oracle (complex128)."""
    from strawberryfields_amd import tfpath

    n, D = 2, 5
    rng = np.random.default_rng(52)
    ops = circuits.layered(n, 1, rng, sd=0.3, loss=True)
    ref, _ = fo.run_ops(ops, n, D, pure=False)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    idx = torch.as_tensor(np.indices([D] * n).sum(axis=0), dtype=torch.float64, device=DEV)
    w = (idx - 1.0) ** 2
    p1 = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    tfc = tfpath.TFPathCircuit([(o[0], o[1]) for o in ops], n, D, C128, DEV, mixed=True)
    rho = tfc.forward(p1)
    assert np.abs(to_np(rho) - ref).max() < 1e-12
    (tfc.all_fock_probs(rho) * w).sum().backward()
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, True, C128, DEV)
    p2 = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    rho0 = torch.as_tensor(fo.vacuum(n, D, pure=False), dtype=C128, device=DEV)
    rho2 = runtime.CircuitFunction.apply(p2, rho0, cc, None)
    (runtime.all_fock_probs(rho2, n, D, True) * w).sum().backward()
    g1, g2 = p1.grad, p2.grad
    assert float((g1 - g2).norm() / g2.norm()) < 1e-10


# ------------------------------------------------------------------ per-fiber kernels
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure,n,D", [(True, 4, 6), (True, 5, 4), (False, 3, 5), (False, 3, 4)])
def test_per_fiber_kernels_match_oracle_and_fd(dtype, pure, n, D):
    """States large enough that the outer axes take the per-fiber kernels
    (innermost non-gate run >= 16): forward vs the oracle, gradient vs FD."""
    rng = np.random.default_rng(200 + n * D)
    ops = circuits.layered(n, 1, rng, sd=0.3, loss=not pure)
    ref, rpure = fo.run_ops(ops, n, D, pure=pure)
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, not pure, dtype, DEV, fuse=0)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    params = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    st0 = torch.as_tensor(fo.vacuum(n, D, pure=pure), dtype=dtype, device=DEV)
    out = runtime.CircuitFunction.apply(params, st0, cc, None)
    assert np.abs(to_np(out) - ref).max() < TOL[dtype]
    idx = np.indices([D] * n).sum(axis=0)
    weights = (idx - 1.0) ** 2
    (runtime.all_fock_probs(out, n, D, not rpure).to(torch.float64) * torch.as_tensor(weights, device=DEV)).sum().backward()
    g = params.grad.cpu().numpy()
    flat, g_fd = _fd_grad(ops, n, D, pure, None, weights)
    gflat = np.array([g[i, j] for i, (_, _, p) in enumerate(ops) for j in range(len(p))])
    rel = np.linalg.norm(gflat - g_fd) / np.linalg.norm(g_fd)
    assert rel < (1e-8 if dtype == C128 else 1e-5), rel


@pytest.mark.parametrize("D", [16, 20])
def test_dense_gates_on_the_innermost_axis_match_oracle_and_fd(D):
    """Complex64 dense gates on the last mode of a ket whose fiber count is a
    multiple of 16 take the staged innermost-axis variant of the matrix-core
    kernels (each 16-fiber chain copied through LDS): forward against the
    complex128 oracle, gradient of a photon-number cost against central
    differences of the oracle."""
    n = 3
    rng = np.random.default_rng(500 + D)
    ops = [("S", (2,), (0.3, 0.7)), ("D", (2,), (0.4, -1.1)), ("BS", (1, 2), (0.6, 0.2)),
           ("D", (2,), (-0.2, 0.5)), ("S", (2,), (-0.25, 2.0))]
    ref, _ = fo.run_ops(ops, n, D, pure=True)
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, False, C64, DEV, fuse=0)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    params = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    st0 = torch.as_tensor(fo.vacuum(n, D, pure=True), dtype=C64, device=DEV)
    out = runtime.CircuitFunction.apply(params, st0, cc, None)
    assert np.abs(to_np(out) - ref).max() < TOL[C64]
    idx = np.indices([D] * n).sum(axis=0)
    weights = (idx - 1.0) ** 2
    (runtime.all_fock_probs(out, n, D, False).to(torch.float64) * torch.as_tensor(weights, device=DEV)).sum().backward()
    g = params.grad.cpu().numpy()
    gflat = np.array([g[i, j] for i, (_, _, p) in enumerate(ops) for j in range(len(p))])
    g_fd = _fd_subset(ops, n, D, True, weights, list(range(len(gflat))))
    rel = np.linalg.norm(gflat - g_fd) / np.linalg.norm(g_fd)
    assert rel < 1e-5, rel


# ------------------------------------------------------------------ batched circuits
def test_batch_evaluator_matches_per_circuit_autograd():
    rng = np.random.default_rng(61)
    circs = []
    for n, D, L in [(2, 9, 2), (3, 6, 1), (4, 5, 2), (2, 20, 1), (5, 4, 1)]:
        circs.append(dict(n=n, D=D, ops=circuits.layered(n, L, rng, sd=0.3)))
    tuner = runtime.Autotuner(path="")
    be = runtime.BatchEvaluator(circs, C64, DEV, nstreams=3, tuner=tuner)
    plist = []
    for c in circs:
        rows = np.zeros((len(c["ops"]), 2))
        for i, (_, _, p) in enumerate(c["ops"]):
            rows[i, : len(p)] = p
        plist.append(torch.tensor(rows, dtype=torch.float64, device=DEV))
    for rep in range(3):
        scaled = [p * (1 + 0.05 * rep) for p in plist]
        losses, grads = be(scaled)
        losses = losses.clone()
        grads = [g.clone() for g in grads]
        for k, c in enumerate(circs):
            n, D = c["n"], c["D"]
            cc = runtime.CompiledCircuit([(o[0], o[1]) for o in c["ops"]], n, D, False, C64, DEV, fuse=0)
            p = scaled[k].clone().requires_grad_(True)
            vac = torch.zeros([D] * n, dtype=C64, device=DEV)
            vac.view(-1)[0] = 1
            psi = runtime.CircuitFunction.apply(p, vac, cc, None)
            probs = runtime.all_fock_probs(psi, n, D, False).to(torch.float64)
            w = torch.as_tensor(np.indices([D] * n).sum(axis=0), dtype=torch.float64, device=DEV)
            loss = (probs * w).sum()
            loss.backward()
            assert abs(float(loss) - float(losses[k])) < 1e-5 * max(1.0, abs(float(loss)))
            rel = float((p.grad - grads[k]).norm() / max(p.grad.norm(), 1e-30))
            assert rel < 1e-5, (k, rel)


# ------------------------------------------------------------------ LDS slab kernels
def _fd_subset(ops, n, D, pure, weights, idx, h=1e-5):
    flat = np.array([v for _, _, p in ops for v in p], dtype=float)
    g = np.zeros(len(idx))
    for j, i in enumerate(idx):
        a, b = flat.copy(), flat.copy()
        a[i] += h
        b[i] -= h
        g[j] = (_oracle_loss(ops, n, D, pure, a, None, weights) - _oracle_loss(ops, n, D, pure, b, None, weights)) / (2 * h)
    return g


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize(
    "pure,n,D,pairs",
    [
        (True, 4, 16, [(2, 3), (3, 2), (1, 2)]),  # innermost pair: slab tiles of D^2
        (False, 2, 16, [(0, 1), (1, 0)]),  # ket pair (256, 1): one slab per tile
        (False, 3, 7, [(1, 2), (2, 1), (0, 1)]),  # bra pair (343, 7): Sin = 7, mid axes
        (False, 2, 20, [(0, 1)]),  # ket pair slab 8000: split mid range in the adjoint
    ],
)
def test_slab_kernels_match_oracle_and_fd(dtype, pure, n, D, pairs):
    """States of >= 2^16 amplitudes whose gates touch the innermost axes take
    the LDS slab kernels (beamsplitter, loss) and the matrix-core kernel
    (complex64 dense gates): forward vs the oracle, gradient vs FD on a seeded
    subset of the parameters."""
    rng = np.random.default_rng(300 + n * D)
    ops = []
    for m in range(n):
        ops.append(("S", (m,), [float(rng.uniform(0.05, 0.3)), float(rng.uniform(0, 2 * np.pi))]))
        ops.append(("D", (m,), [float(rng.uniform(0.05, 0.3)), float(rng.uniform(0, 2 * np.pi))]))
    for p, q in pairs:
        ops.append(("BS", (p, q), [float(rng.uniform(0.2, 1.2)), float(rng.uniform(0, 2 * np.pi))]))
    if not pure:
        for m in range(n):
            ops.append(("LOSS", (m,), [float(rng.uniform(0.6, 0.95))]))
    ops.append(("BS", (pairs[0][0], pairs[0][1]), [float(rng.uniform(0.2, 1.2)), float(rng.uniform(0, 2 * np.pi))]))
    ref, rpure = fo.run_ops(ops, n, D, pure=pure)
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, not pure, dtype, DEV, fuse=0)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    params = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    st0 = torch.as_tensor(fo.vacuum(n, D, pure=pure), dtype=dtype, device=DEV)
    out = runtime.CircuitFunction.apply(params, st0, cc, None)
    assert np.abs(to_np(out) - ref).max() < TOL[dtype]
    idx = np.indices([D] * n).sum(axis=0)
    weights = (idx - 1.0) ** 2
    (runtime.all_fock_probs(out, n, D, not rpure).to(torch.float64) * torch.as_tensor(weights, device=DEV)).sum().backward()
    g = params.grad.cpu().numpy()
    gflat = np.array([g[i, j] for i, (_, _, p) in enumerate(ops) for j in range(len(p))])
    sel = np.sort(rng.choice(len(gflat), size=min(8, len(gflat)), replace=False))
    g_fd = _fd_subset(ops, n, D, pure, weights, sel)
    rel = np.linalg.norm(gflat[sel] - g_fd) / np.linalg.norm(g_fd)
    assert rel < (1e-8 if dtype == C128 else 1e-5), rel


# ------------------------------------------------------------------ c5 structure at a smaller cutoff
_C5S = {}


def _c5_structure_reference():
    """Oracle forward and Richardson-extrapolated central differences (4 seeded
    parameters) of the c5-shaped circuit at cutoff 10, computed once for both
    dtypes."""
    if _C5S:
        return _C5S
    n, D, L = 3, 10, 2
    rng = np.random.default_rng(20261205)
    ops = circuits.layered(n, L, rng, sd=0.1, loss=True)
    vac = fo.vacuum(n, D, pure=False)
    ref, _ = fo.run_ops(ops, n, D, pure=False, state=vac)
    idx = np.arange(D, dtype=float)
    w = (idx[:, None, None] + idx[None, :, None] + idx[None, None, :] - 3.0) ** 2
    flat = np.array([v for _, _, p in ops for v in p], dtype=float)
    sel = np.sort(rng.choice(len(flat), size=4, replace=False))

    def loss_at(vals):
        k, new = 0, []
        for name, modes, p in ops:
            new.append((name, modes, tuple(vals[k : k + len(p)])))
            k += len(p)
        st, _ = fo.run_ops(new, n, D, pure=False, state=vac)
        return float((fo.all_fock_probs(st, n, False) * w).sum())

    # central differences at h and h/2 combined as (4 g(h/2) - g(h)) / 3, which
    # removes the O(h^2) term: a step small enough for plain differences
    # leaves ~1e-8 of rounding in the loss (~10) divided by h
    def central(i, h):
        a, b = flat.copy(), flat.copy()
        a[i] += h
        b[i] -= h
        return (loss_at(a) - loss_at(b)) / (2 * h)

    g_fd = [(4 * central(i, 5e-4) - central(i, 1e-3)) / 3 for i in sel]
    _C5S.update(n=n, D=D, L=L, ops=ops, vac=vac, ref=ref, w=w, sel=sel, g_fd=np.array(g_fd))
    return _C5S


@pytest.mark.parametrize("dtype", [C128, C64])
def test_c5_structure_dm_layers_with_loss_checkpoints(dtype):
    """The c5 circuit shape (3-mode DM from vacuum, layered circuit with a loss
    channel per mode, loss sum_n p(n) (n1+n2+n3-3)^2, layer-boundary checkpoints)
    at cutoff 10 (10^6 DM entries), where every c5 kernel route is taken
    (one-pass two-sided dense gates on inner strides 1 and >= 16, beamsplitters
    on strided and slab tiles, slab loss) and the complex128 oracle runs in
    seconds: forward DM vs the oracle, gradient vs central differences."""
    r = _c5_structure_reference()
    n, D, ops = r["n"], r["D"], r["ops"]
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, True, dtype, DEV)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    params = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    rho0 = torch.as_tensor(r["vac"], dtype=dtype, device=DEV)
    rho = runtime.CircuitFunction.apply(params, rho0, cc, len(ops) // r["L"])
    assert np.abs(to_np(rho) - r["ref"]).max() < TOL[dtype]
    probs = runtime.all_fock_probs(rho, n, D, True).to(torch.float64)
    (probs * torch.as_tensor(r["w"], device=DEV)).sum().backward()
    g = params.grad.cpu().numpy()
    gflat = np.array([g[i, j] for i, (_, _, p) in enumerate(ops) for j in range(len(p))])
    rel = np.linalg.norm(gflat[r["sel"]] - r["g_fd"]) / np.linalg.norm(r["g_fd"])
    assert rel < (1e-8 if dtype == C128 else 1e-5), rel


# ------------------------------------------------------------------ S2, MZ, CK, V
def _s2_dense_from_blocks(blk, D):
    """S2 table: block s = D - 1 + d holds G[N][n] for out1 - out2 = in1 - in2 = d."""
    mat = np.zeros((D,) * 4, dtype=complex)
    off = 0
    for s in range(2 * D - 1):
        lo, hi = max(0, s - D + 1), min(s, D - 1)
        c, d = hi - lo + 1, s - (D - 1)
        b = blk[off : off + c * c].reshape(c, c)
        for a in range(c):
            for k in range(c):
                N, n = lo + a, lo + k
                mat[N, n, N - d, n - d] = b[a, k]
        off += c * c
    return mat, off


def _fd4(f, x, h=1e-3):
    return (-f(x + 2 * h) + 8 * f(x + h) - 8 * f(x - h) + f(x - 2 * h)) / (12 * h)


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("D", [1, 2, 6, 15, 20])
def test_two_mode_gate_tables(dtype, D):
    """S2, MZ and CK tables against the oracle matrices; the first-parameter
    derivative blocks against difference quotients of the oracle, and the
    second-parameter phase rules the adjoint kernels use (dS2/dphi =
    i (out1 - in1) S2, dMZ/dphi_ex = i in1 MZ) against the same quotients."""
    # FP64 recurrences on both sides carry up to ~3e-12 each at cutoff 20
    tol = (1e-13 if D < 15 else 1e-11) if dtype == C128 else 1e-6
    dtol = (1e-8 if D < 15 else 1e-6) if dtype == C128 else 1e-5
    i1 = np.arange(D)[None, :, None, None]  # in1 index of mat[out1, in1, out2, in2]
    o1 = np.arange(D)[:, None, None, None]
    cases = [
        ("S2", fo.two_mode_squeezing, _s2_dense_from_blocks, [(0.3, 0.7), (-0.4, 2.1), (0.9, -1.3)]),
        ("MZ", fo.mzgate, _bs_dense_from_blocks, [(0.3, 0.7), (-2.4, 2.1), (np.pi, -1.3)]),
        ("CK", lambda k, _p, D: fo.cross_kerr(k, D), _bs_dense_from_blocks, [(0.3, 0.0), (-1.7, 0.0)]),
    ]
    for kind, fn, unpack, params in cases:
        for p0, p1 in params:
            t = to_np(runtime.build_gate(kind, p0, p1, D, dtype, DEV))
            G, nb = unpack(t, D)
            ref = fn(p0, p1, D)
            assert np.abs(G - ref).max() < tol, (kind, p0, p1)
            if kind == "CK":  # exact: a difference quotient of e^{i kappa n1 n2} loses ~(n1 n2)^5 h^4
                d0 = 1j * (o1 * np.arange(D)[None, None, :, None]) * ref
            else:
                d0 = _fd4(lambda x: fn(x, p1, D), p0)  # noqa: B023
            scale = max(1.0, np.abs(d0).max())
            assert np.abs(unpack(t[nb:], D)[0] - d0).max() < dtol * scale, (kind, "d0")
            if kind != "CK":
                d1 = _fd4(lambda x: fn(p0, x, D), p1)  # noqa: B023
                rule = 1j * ((o1 - i1) if kind == "S2" else i1) * ref
                assert np.abs(rule - d1).max() < 1e-6 * max(1.0, np.abs(d1).max()), (kind, "d1 rule")


@pytest.mark.parametrize("dtype", [C128, C64])
def test_two_mode_squeezing_tables_match_mpmath(dtype):
    """S2 tables against the 50-digit closed form of tests/golden/gates_mpmath.npz."""
    g = np.load(os.path.join(GOLDEN, "gates_mpmath.npz"))
    keys = [k for k in g.files if k.startswith("S2_") and not k.endswith("_params")]
    assert len(keys) == 4
    for key in keys:
        r, phi, D = g[key + "_params"]
        t = to_np(runtime.build_gate("S2", r, phi, int(D), dtype, DEV))
        tol = (1e-13 if D < 15 else 5e-12) if dtype == C128 else 1e-6
        assert np.abs(_s2_dense_from_blocks(t, int(D))[0] - g[key]).max() < tol, key


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("D", [1, 2, 6, 15, 20, 32])
def test_cubic_phase_tables(dtype, D):
    tol = 1e-12 if dtype == C128 else 2e-6
    for g in [0.3, -1.1, 2.5, 6.0]:
        t = to_np(runtime.build_gate("V", g, 0.0, D, dtype, DEV))
        DD = D * D
        V = fo.cubic_phase(g, D)
        assert np.abs(t[:DD].reshape(D, D) - V).max() < tol, g
        dV = _fd4(lambda x: fo.cubic_phase(x, D), g, h=1e-4)
        assert np.abs(t[DD : 2 * DD].reshape(D, D) - dV).max() < 1e-6 * max(1.0, np.abs(dV).max()), g
        assert np.abs(t[2 * DD :]).max() == 0.0


_NEW_GATE_CASES = [
    # (pure, n, D): small states (element-wise kernels), inner runs >= 16 on
    # the outer axes (per-fiber / matrix-core kernels), >= 2^16 states with the
    # innermost pair (LDS slab tiles)
    (True, 3, 5),
    (False, 2, 5),
    (True, 4, 6),
    (False, 3, 5),
    (True, 4, 16),
    (False, 2, 16),
]


def _new_gate_ops(n, rng, pure):
    ops = []
    for m in range(n):
        ops.append(("D", (m,), [float(rng.uniform(0.1, 0.4)), float(rng.uniform(0, 2 * np.pi))]))
    pairs = [(0, 1), (n - 1, n - 2)] + ([(n - 2, n - 1), (0, n - 1)] if n > 2 else [])
    for a, b in pairs:
        ops.append(("S2", (a, b), [float(rng.uniform(-0.3, 0.3)), float(rng.uniform(0, 2 * np.pi))]))
        ops.append(("MZ", (a, b), [float(rng.uniform(0.3, 2.8)), float(rng.uniform(0, 2 * np.pi))]))
        ops.append(("CK", (b, a), [float(rng.uniform(-1, 1))]))
    for m in range(n):
        ops.append(("V", (m,), [float(rng.uniform(-0.4, 0.4))]))
    if not pure:
        ops.append(("LOSS", (0,), [0.8]))
        ops.append(("S2", (0, 1), [0.2, 0.3]))
    return ops


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure,n,D", _NEW_GATE_CASES)
def test_new_gates_match_oracle_and_fd(dtype, pure, n, D):
    """S2gate, MZgate, CKgate and Vgate on every kernel route: forward through
    the backend against the oracle; gradients of a photon-number cost against
    central differences of the complex128 oracle on a seeded parameter subset."""
    rng = np.random.default_rng(400 + 10 * n + D)
    ops = _new_gate_ops(n, rng, pure)
    ref, rpure = fo.run_ops(ops, n, D, pure=pure)
    be = run_backend(ops, n, D, dtype, pure=pure)
    st = be.state()
    got = to_np(st.ket() if rpure else st.dm())
    assert np.abs(got - ref).max() < TOL[dtype]
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, not pure, dtype, DEV, fuse=0)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    params = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    st0 = torch.as_tensor(fo.vacuum(n, D, pure=pure), dtype=dtype, device=DEV)
    out = runtime.CircuitFunction.apply(params, st0, cc, None)
    assert np.abs(to_np(out) - ref).max() < TOL[dtype]
    idx = np.indices([D] * n).sum(axis=0)
    weights = (idx - 1.0) ** 2
    probs = runtime.all_fock_probs(out, n, D, not rpure).to(torch.float64)
    (probs * torch.as_tensor(weights, device=DEV)).sum().backward()
    g = params.grad.cpu().numpy()
    gflat = np.array([g[i, j] for i, (_, _, p) in enumerate(ops) for j in range(len(p))])
    nsel = len(gflat) if D ** (n if pure else 2 * n) <= 4096 else 10
    sel = np.sort(rng.choice(len(gflat), size=min(nsel, len(gflat)), replace=False))
    g_fd = _fd_subset(ops, n, D, pure, weights, sel)
    rel = np.linalg.norm(gflat[sel] - g_fd) / np.linalg.norm(g_fd)
    assert rel < (1e-8 if dtype == C128 else 1e-5), rel


def test_new_gates_reference_known_answers():
    """The reference's own known-answer tests through the HIP backend:
    two-mode squeezed vacuum on every pair of 4 modes
    (test_twomode_squeezing_operation.py:44-79), MZ on |1,1>
    (test_mzgate_operation.py:41-60), cross-Kerr and cubic phase on uniform
    states (test_nongaussian_gates.py:30-102)."""
    D = 5
    k = np.arange(D)
    for r, p in [(0.1, 0.0), (0.15, 2.0943951023931953)]:
        amp = np.zeros((D, D), dtype=complex)
        amp[k, k] = (np.exp(1j * p) * np.tanh(r)) ** k / np.cosh(r)
        for modes in itertools.combinations(range(4), 2):
            be = run_backend([("S2", modes, (r, p))], 4, D, C128)
            red = to_np(be.state().reduced_dm(list(modes)))
            assert np.abs(red - np.einsum("ac,bd->abcd", amp, amp.conj())).max() < 1e-13, modes
    one = np.zeros((6, 6), dtype=complex)
    one[1, 1] = 1.0
    for pin in [0.3, np.pi / 2, 2.5]:
        be = run_backend([("MZ", (0, 1), (pin, 0.7))], 2, 6, C128, init=one)
        pr = to_np(be.state().all_fock_probs())
        assert abs(pr[1, 1] - np.cos(pin) ** 2) < 1e-13 and abs(pr[0, 2] - 0.5 * np.sin(pin) ** 2) < 1e-13
    for kappa in [0.5, 2.0]:
        be = run_backend([("CK", (0, 1), (kappa,))], 2, 6, C128, init=np.ones((6, 6), dtype=complex) / 6)
        assert np.abs(to_np(be.state().ket()) - np.exp(1j * kappa * np.outer(np.arange(6), np.arange(6))) / 6).max() < 1e-14
    a = fo.annihilation(8).real
    x3 = np.linalg.matrix_power(a + a.T, 3)
    from scipy.linalg import expm

    for gamma in [1.0, 4.0]:
        ket = np.ones(8, dtype=complex) / np.sqrt(8)
        be = run_backend([("V", (0,), (gamma,))], 1, 8, C128, init=ket)
        assert np.abs(to_np(be.state().ket()) - expm(1j * gamma * x3 / 6) @ ket).max() < 1e-12


# ------------------------------------------------------------------ measurements, preparations
def _load(be, state, pure):
    n = state.ndim if pure else state.ndim // 2
    be.begin_circuit(n, cutoff_dim=state.shape[0], pure=pure, dtype=be._dtype, device=DEV)
    be._state = torch.as_tensor(state, dtype=be._dtype, device=DEV).clone()
    be._pure = pure


def _backend_with(state, pure, dtype):
    be = HIPFockBackend()
    be._dtype = dtype
    _load(be, state, pure)
    return be


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure", [True, False])
def test_marginal_projection_and_matrix_kernels(dtype, pure):
    """Marginal distributions on every mode subset, |0><x| projections and a
    generic single-mode matrix against the oracle."""
    n, D = 3, 5
    st = rand_state(n, D, pure, 77)
    t = torch.as_tensor(st, dtype=dtype, device=DEV)
    probs = fo.all_fock_probs(st, n, pure)
    for k in range(0, n + 1):
        for modes in itertools.combinations(range(n), k):
            got = to_np(runtime.marginal_probs(t, n, D, not pure, list(modes)))
            ref = probs.sum(axis=tuple(m for m in range(n) if m not in modes))
            assert np.abs(got - ref).max() < (1e-14 if dtype == C128 else 1e-6), modes
    c = fo.Circuit(n, D, pure=pure)
    for modes, x in [([1], [2]), ([2, 0], [1, 3]), ([0, 1, 2], [4, 0, 1])]:
        got = to_np(runtime.project_reset(t, n, D, not pure, modes, x, 0.5))
        ref = 0.5 * c.project_reset(modes, x, state=st)
        assert np.abs(got - ref).max() < TOL[dtype], modes
    rng = np.random.default_rng(3)
    U = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    for mode in range(n):
        got = to_np(runtime.apply_matrix(t, n, D, not pure, mode, torch.as_tensor(U, device=DEV)))
        ref = c.apply_gate_BLAS(U, [mode], state=st)
        assert np.abs(got - ref).max() < TOL[dtype], mode


@pytest.mark.parametrize("pure", [True, False])
def test_measure_fock_matches_oracle_with_the_same_seed(pure):
    """Same numpy seed, same outcome and conditional state as the oracle, for
    sampled and post-selected measurements; a None in `select` is rejected by
    both (circuit.py:628-629)."""
    n, D = 3, 5
    st = rand_state(n, D, pure, 78)
    for modes, select in [([2, 1], [None, 3]), ([0, 2], [1, None])]:
        with pytest.raises(NotImplementedError):
            _backend_with(st, pure, C128).measure_fock(modes, select=select)
        with pytest.raises(NotImplementedError):
            fo.Circuit(n, D, pure=pure).measure_fock(modes, select=select)
    cases = [([0], None), ([2, 0], None), ([0, 1, 2], None), ([1], [2]), ([2, 0], [3, 1])]
    for seed, (modes, select) in enumerate(cases):
        be = _backend_with(st, pure, C128)
        c = fo.Circuit(n, D, pure=pure)
        c._state = st.copy()
        np.random.seed(100 + seed)
        got = be.measure_fock(modes, select=select)
        np.random.seed(100 + seed)
        ref = c.measure_fock(modes, select=select)
        assert got.tolist() == ref.tolist(), (modes, select)
        assert np.abs(to_np(be._state) - c._state).max() < 1e-12, (modes, select)


def test_measure_fock_reference_known_answers():
    """test_fock_measurement.py:107-150 and test_postselection.py:52-68 through the
    backend: Fock products give their photon numbers and the post-selected
    beamsplitter output conserves photon number."""
    D = 5
    np.random.seed(4)
    for pure in (True, False):
        for idx, modes in enumerate([[0], [1, 2], [2, 0, 1]]):
            ns = [idx % D, (idx + 2) % D, (idx + 4) % D]
            be = HIPFockBackend()
            be.begin_circuit(3, cutoff_dim=D, pure=pure, dtype=C128, device=DEV)
            for m, k in enumerate(ns):
                be.prepare_fock_state(k, m)
            assert be.measure_fock(modes).tolist() == [[ns[m] for m in modes]]
            assert abs(be.state().trace() - 1) < 1e-12
        be = HIPFockBackend()
        be.begin_circuit(2, cutoff_dim=D, pure=pure, dtype=C128, device=DEV)
        for nphot in range(D - 1):
            be.reset(pure=pure)
            be.prepare_fock_state(nphot, 0)
            be.prepare_fock_state(D - 1 - nphot, 1)
            be.beamsplitter(np.pi / 4, 0, 0, 1)
            try:
                r1 = be.measure_fock([0], select=[D // 2])
            except ZeroDivisionError:
                continue
            r2 = be.measure_fock([1])
            assert r1[0][0] + r2[0][0] == D - 1


@pytest.mark.parametrize("dtype", [C128, C64])
def test_homodyne_density_matches_oracle(dtype):
    D = 12
    rng = np.random.default_rng(9)
    A = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    rho = A @ A.conj().T
    rho /= np.trace(rho)
    c = fo.Circuit(1, D)
    for phi in [0.0, 0.8, -2.2]:
        q, ref = c.homodyne_pdf(rho, phi, 10, 20001)
        got = to_np(runtime.homodyne_pdf(torch.as_tensor(rho, dtype=dtype, device=DEV), D, phi, 10, 20001, 2.0))
        assert np.abs(got - ref).max() < (1e-15 if dtype == C128 else 1e-9), phi
        assert abs(got.sum() - 1) < 1e-6


@pytest.mark.parametrize("pure", [True, False])
def test_measure_homodyne_matches_oracle_with_the_same_seed(pure):
    n, D = 2, 8
    st, _ = fo.run_ops([("S", (0,), (0.4, 0.3)), ("D", (1,), (0.5, 1.0)), ("BS", (0, 1), (0.6, 0.2))], n, D, pure=pure)
    for seed, (phi, mode, select) in enumerate([(0.0, 0, None), (0.7, 1, None), (1.1, 0, 0.35)]):
        be = _backend_with(st, pure, C128)
        c = fo.Circuit(n, D, pure=pure)
        c._state = st.copy()
        np.random.seed(200 + seed)
        got = be.measure_homodyne(phi, mode, select=select, num_bins=20001)
        np.random.seed(200 + seed)
        ref = c.measure_homodyne(phi, mode, select=select, num_bins=20001)
        assert got.tolist() == ref.tolist(), (phi, mode)
        assert np.abs(to_np(be._state) - c._state).max() < 1e-11, (phi, mode)
    be = HIPFockBackend()
    be.begin_circuit(1, cutoff_dim=D, pure=pure, dtype=C128, device=DEV)
    be.squeeze(np.arcsinh(1.0), 0.0, 0)
    be.displacement(0.8, 0.0, 0)
    be.measure_homodyne(0, 0, num_bins=4001)
    assert be.is_vacuum(1e-12)  # test_homodyne.py:27-38


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure", [True, False])
def test_multimode_preparation_matches_oracle(dtype, pure):
    """prepare_ket_state / prepare_dm_state on mode subsets in any order, and on
    all modes permuted, against the oracle's prepare_multimode."""
    n, D = 4, 4
    base, _ = fo.run_ops([("D", (m,), (0.3, 0.5 * m)) for m in range(n)] + [("BS", (0, 2), (0.5, 0.1))], n, D, pure=pure)
    rng = np.random.default_rng(10)
    ket2 = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    ket2 /= np.linalg.norm(ket2)
    A = rng.normal(size=(D * D, D * D)) + 1j * rng.normal(size=(D * D, D * D))
    dm2 = (A @ A.conj().T / np.trace(A @ A.conj().T)).reshape(D, D, D, D).transpose(0, 2, 1, 3)
    full = rng.normal(size=[D] * n) + 1j * rng.normal(size=[D] * n)
    for state, modes, ket in [
        (ket2, [3, 1], True),
        (ket2.reshape(-1), [0, 2], True),
        (dm2, [1, 2], False),
        (dm2.reshape(D * D, D * D), [2, 0], False),
        (full, [2, 0, 3, 1], True),
    ]:
        be = _backend_with(base, pure, dtype)
        c = fo.Circuit(n, D, pure=pure)
        c._state = base.copy()
        if ket:
            be.prepare_ket_state(state, modes)
        else:
            be.prepare_dm_state(state, modes)
        c.prepare_multimode(state, modes)
        assert be._pure == c._pure, modes
        assert np.abs(to_np(be._state) - c._state).max() < TOL[dtype], modes


def test_engine_collects_measurement_samples():
    prog = frontend.Program(3)
    with prog.context as q:
        frontend.Fock(2) | q[0]
        frontend.Fock(1) | q[2]
        frontend.MeasureFock() | (q[2], q[0])
        frontend.MeasureX | q[1]
    be = HIPFockBackend()
    np.random.seed(1)
    res = frontend.Engine(be, backend_options={"cutoff_dim": 5, "pure": True}).run(prog)
    assert res.samples.shape == (1, 3)
    assert res.samples[0][0] == 2 and res.samples[0][2] == 1
    assert q[0].val == 2 and q[2].val == 1


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure", [True, False])
@pytest.mark.parametrize("mesh", ["rectangular", "rectangular_phase_end", "rectangular_symmetric", "rectangular_compact",
                                  "triangular_compact", "sun_compact"])
def test_interferometer_matches_exact_passive_action(dtype, pure, mesh):
    """Interferometer(U) (rectangular meshes on the BS / R kernels, the
    symmetric mesh on the MZ kernels) on a 4-mode state with fewer than D
    photons equals the exact creation-operator action of U, for the register in
    order and permuted."""
    from scipy.stats import unitary_group

    n, D = 4, 5
    rng = np.random.default_rng(12)
    psi = np.zeros([D] * n, dtype=complex)
    for occ in itertools.product(range(D), repeat=n):
        if sum(occ) < D:
            psi[occ] = rng.normal() + 1j * rng.normal()
    psi /= np.linalg.norm(psi)
    U = unitary_group.rvs(n, random_state=5)
    for order in ([0, 1, 2, 3], [2, 0, 3, 1]):
        prog = frontend.Program(n)
        with prog.context as q:
            frontend.Ket(psi if pure else fo.mix(psi, n)) | q
            frontend.Interferometer(U, mesh=mesh) | tuple(q[m] for m in order)
        be = HIPFockBackend()
        res = frontend.Engine(be, backend_options={"cutoff_dim": D, "pure": pure, "dtype": dtype}).run(prog)
        # U acts on the register in `order`: permute U onto the mode axes
        P = np.zeros((n, n))
        P[order, range(n)] = 1
        ref = fo.passive_fock_action(P @ U @ P.T, psi, n, D)
        got = to_np(res.state.ket() if pure else res.state.dm())
        want = ref if pure else fo.mix(ref, n)
        assert np.abs(got - want).max() < TOL[dtype], order


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("pure,n,D", [(True, 3, 5), (True, 2, 8), (False, 2, 4), (False, 3, 5)])
def test_diagonal_runs_match_oracle_and_fd(dtype, pure, n, D):
    """Consecutive R / K gates (repeated modes, more than one run's capacity)
    run as one-pass diagonal runs: forward against the oracle, every gradient
    against central differences; row lengths with and without whole groups of
    four (vector and scalar paths)."""
    rng = np.random.default_rng(500 + 10 * n + D)
    ops = [("D", (m,), [float(rng.uniform(0.1, 0.4)), float(rng.uniform(0, 6))]) for m in range(n)]
    for j in range(11):
        ops.append(("R" if j % 2 == 0 else "K", (j % n,), [float(rng.uniform(-1, 1))]))
    ops.append(("BS", (0, n - 1), [0.5, 0.3]))
    ops += [("K", (m,), [float(rng.uniform(-0.5, 0.5))]) for m in range(n)]
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, not pure, dtype, DEV, fuse=0)
    ref, rpure = fo.run_ops(ops, n, D, pure=pure)
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    params = torch.tensor(rows, dtype=torch.float64, device=DEV, requires_grad=True)
    st0 = torch.as_tensor(fo.vacuum(n, D, pure=pure), dtype=dtype, device=DEV)
    out = runtime.CircuitFunction.apply(params, st0, cc, None)
    assert np.abs(to_np(out) - ref).max() < TOL[dtype]
    idx = np.indices([D] * n).sum(axis=0)
    weights = (idx - 1.0) ** 2
    (runtime.all_fock_probs(out, n, D, not rpure).to(torch.float64) * torch.as_tensor(weights, device=DEV)).sum().backward()
    g = params.grad.cpu().numpy()
    gflat = np.array([g[i, j] for i, (_, _, p) in enumerate(ops) for j in range(len(p))])
    _, g_fd = _fd_grad(ops, n, D, pure, None, weights)
    rel = np.linalg.norm(gflat - g_fd) / np.linalg.norm(g_fd)
    assert rel < (1e-8 if dtype == C128 else 1e-5), rel


# ------------------------------------------------------------------ maximum cutoffs
@pytest.mark.parametrize("dtype", [C128, C64])
def test_maximum_cutoff_tables_and_ops(dtype):
    """The largest cutoff of each gate family (64 single-mode, 40 two-mode, 32
    cubic phase): tables against the oracle and gates on 1- and 2-mode states,
    pure and mixed (the element-wise kernels past the matrix-core range)."""
    tol_t = 1e-10 if dtype == C128 else 2e-6
    # at cutoff 64 the FP64 recurrences (library and oracle alike) carry up to ~2e-9
    # (squeezing): compare with the 50-digit closed forms of tests/golden
    g = np.load(os.path.join(GOLDEN, "gates_mpmath.npz"))
    for key in ("D_4", "S_4"):
        p0, p1, Dg = g[key + "_params"]
        assert int(Dg) == 64
        t = to_np(runtime.build_gate(key[0], p0, p1, 64, dtype, DEV))
        assert np.abs(t[: 64 * 64].reshape(64, 64) - g[key]).max() < (5e-9 if dtype == C128 else 2e-6), key
    t = to_np(runtime.build_gate("V", 0.4, 0.0, 32, dtype, DEV))
    assert np.abs(t[: 32 * 32].reshape(32, 32) - fo.cubic_phase(0.4, 32)).max() < tol_t
    # the two-mode recurrences (library and oracle alike) lose accuracy with the
    # photon number: against 50-digit closed forms at cutoff 40 the oracle's
    # beamsplitter is off by 1e-14 at N1 + N2 = 20, 1e-10 at 50 and 7e-7 at 78,
    # its S2 by 7e-12 at out1 + in2 = 40 and 8e-7 at 78.  Entries below 40
    # photons are held to the FP64 floor, the rest to the conditioning.
    bs_tol = 5e-6 if dtype == C128 else 1e-5
    low_tol = 1e-10 if dtype == C128 else 2e-6
    o1, i1, o2, i2 = np.indices((40,) * 4)
    for kind, fn, unpack, low in [
        ("BS", lambda a, b, D: fo.beamsplitter(a, b, D), _bs_dense_from_blocks, (o1 + o2) < 40),
        ("S2", fo.two_mode_squeezing, _s2_dense_from_blocks, (o1 + i2) < 40),
        ("CK", lambda a, b, D: fo.cross_kerr(a, D), _bs_dense_from_blocks, (o1 + o2) < 80),
    ]:
        t = to_np(runtime.build_gate(kind, 0.6, 0.4, 40, dtype, DEV))
        d = np.abs(unpack(t, 40)[0] - fn(0.6, 0.4, 40))
        assert d.max() < bs_tol, kind
        assert d[low].max() < low_tol, kind
    # MZ at cutoff 40: complete photon-number blocks unitary, one-photon block the mode unitary
    t = to_np(runtime.build_gate("MZ", 0.3, 0.7, 40, dtype, DEV))
    M, _ = _bs_dense_from_blocks(t, 40)
    for s in (1, 7, 39):
        ks = np.arange(s + 1)
        blk = M[ks[:, None], ks[None, :], (s - ks)[:, None], (s - ks)[None, :]]
        assert np.abs(blk @ blk.conj().T - np.eye(s + 1)).max() < (1e-11 if dtype == C128 else 1e-5), s
    assert np.abs(M[:30, :30, :30, :30] - fo.mzgate(0.3, 0.7, 30)).max() < bs_tol
    tol64 = 1e-8 if dtype == C128 else TOL[dtype] * 10  # cutoff-64 table conditioning (above)
    for pure in (True, False):
        st1 = rand_state(1, 64, pure, 31)
        for name, p in [("R", (0.7,)), ("K", (0.2,)), ("D", (0.5, 1.1)), ("S", (0.3, 0.4))] + ([] if pure else [("LOSS", (0.7,))]):
            ops = [(name, (0,), p)]
            got = run_backend(ops, 1, 64, dtype, pure=pure, init=st1).state()
            ref, rpure = fo.run_ops(ops, 1, 64, state=st1)
            assert np.abs(to_np(got.ket() if rpure else got.dm()) - ref).max() < tol64, (name, pure)
        ops = [("V", (0,), (0.3,))]
        st32 = rand_state(1, 32, pure, 32)
        got = run_backend(ops, 1, 32, dtype, pure=pure, init=st32).state()
        ref, rpure = fo.run_ops(ops, 1, 32, state=st32)
        assert np.abs(to_np(got.ket() if rpure else got.dm()) - ref).max() < TOL[dtype], ("V", pure)
        st2 = rand_state(2, 40, pure, 33)
        for name, p in [("BS", (0.6, 0.4)), ("S2", (0.2, 0.5)), ("CK", (0.4,))]:
            for modes in [(0, 1), (1, 0)]:
                ops = [(name, modes, p)]
                got = run_backend(ops, 2, 40, dtype, pure=pure, init=st2).state()
                ref, rpure = fo.run_ops(ops, 2, 40, state=st2)
                tol = (1e-7 if dtype == C128 else 2e-5) if name in ("BS", "S2") else TOL[dtype] * 10
                assert np.abs(to_np(got.ket() if rpure else got.dm()) - ref).max() < tol, (name, modes, pure)


# ------------------------------------------------------------------ single-mode preparations
@pytest.mark.parametrize("dtype", [C128, C64])
def test_single_mode_preparations_match_closed_forms(dtype):
    """Coherent, squeezed, displaced squeezed, thermal, GKP and cat states
    prepared by the backend: on one mode they equal the states built from the
    ladder operators with matrix exponentials at cutoff 160 (then truncated;
    the displaced squeezed state to the 5e-10 that the reference's formula
    carries), the thermal closed form and the host GKP expansion (pinned on the
    CPU against its operator definition); on mode 1 of a two-mode register the
    reduced state of that mode is the prepared one and mode 0 stays vacuum."""
    from scipy.linalg import expm

    from strawberryfields_amd import fockstates as fs

    D, L = 14, 160
    a = np.diag(np.sqrt(np.arange(1, L)), 1)
    vac = np.zeros(L)
    vac[0] = 1

    def dsq(rd, pd, rs, ps):
        al, z = rd * np.exp(1j * pd), rs * np.exp(1j * ps)
        return (expm(al * a.T - np.conj(al) * a) @ expm(0.5 * (np.conj(z) * a @ a - z * a.T @ a.T)) @ vac)[:D]

    tol = TOL[dtype]
    cases = [
        ("prepare_coherent_state", (0.7, 0.4), dsq(0.7, 0.4, 0, 0), tol),
        ("prepare_squeezed_state", (0.5, -0.3), dsq(0, 0, 0.5, -0.3), tol),
        ("prepare_displaced_squeezed_state", (0.6, 1.1, 0.3, 0.2), dsq(0.6, 1.1, 0.3, 0.2), max(tol, 5e-10)),
        ("prepare_gkp", ([0.7, 0.3], 0.35, 1e-12), fs.square_gkp(0.7, 0.3, 0.35, 1e-12, D), tol),
    ]
    for method, args, want, t in cases:
        be = HIPFockBackend()
        be.begin_circuit(1, cutoff_dim=D, pure=True, dtype=dtype)
        if method == "prepare_gkp":
            be.prepare_gkp(*args, mode=0)
        else:
            getattr(be, method)(*args, 0)
        got = to_np(be.state().ket())
        assert np.abs(got - want).max() < t, method
        be = HIPFockBackend()
        be.begin_circuit(2, cutoff_dim=D, pure=True, dtype=dtype)
        if method == "prepare_gkp":
            be.prepare_gkp(*args, mode=1)
        else:
            getattr(be, method)(*args, 1)
        st = be.state()
        assert np.abs(to_np(st.reduced_dm(1)) - np.outer(want, want.conj())).max() < t, method
        r0 = to_np(st.reduced_dm(0))
        v0 = np.zeros((D, D))
        v0[0, 0] = np.vdot(want, want).real
        assert np.abs(r0 - v0).max() < t, method
    be = HIPFockBackend()
    be.begin_circuit(1, cutoff_dim=D, pure=True, dtype=dtype)
    be.prepare_thermal_state(0.6, 0)
    k = np.arange(D)
    assert np.abs(to_np(be.state().dm()) - np.diag(0.6**k / 1.6 ** (k + 1))).max() < tol
    with pytest.raises(NotImplementedError):
        be.prepare_gkp([0, 0], 0.2, 1e-12, representation="complex", mode=0)
    prog = frontend.Program(1)
    with prog.context as q:
        frontend.Catstate(0.9, 0.2, 0) | q[0]
    res = frontend.Engine(HIPFockBackend(), backend_options={"cutoff_dim": D, "dtype": dtype}).run(prog)
    al = 0.9 * np.exp(0.2j)
    want = dsq(0.9, 0.2, 0, 0) + dsq(-0.9, 0.2, 0, 0)
    want /= np.sqrt(2 * (1 + np.exp(-2 * abs(al) ** 2)))
    assert np.abs(to_np(res.state.ket()) - want).max() < tol
