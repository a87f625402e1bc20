"""CPU-only checks of the library boundary and the host logic (no GPU calls).

* libsfamd.so loads and exports every function include/sfamd.h declares;
* host-side planning and validation return the documented status codes;
* the front-end mirror keeps the reference's Gate.apply rules;
* the circuit generators produce the documented workloads.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from strawberryfields_amd import _lib, circuits, frontend
from strawberryfields_amd.base import ModeMap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "sfamd.h")).read()
    return sorted(set(re.findall(r"\b(sfamd_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name


def test_status_strings_and_version():
    lib = _lib.load()
    assert lib.sfamd_version() >= 1
    assert lib.sfamd_status_string(0) == b"success"
    assert b"cutoff" in lib.sfamd_status_string(-2)


def test_gate_table_sizes():
    lib = _lib.load()
    D = 7
    assert lib.sfamd_gate_table_size(_lib.OP_ROTATION, D) == D
    assert lib.sfamd_gate_table_size(_lib.OP_DISPLACEMENT, D) == 3 * D * D
    nb = sum((min(s, D - 1) - max(0, s - D + 1) + 1) ** 2 for s in range(2 * D - 1))
    assert nb == D * (2 * D * D + 1) // 3
    assert lib.sfamd_gate_table_size(_lib.OP_BEAMSPLITTER, D) == 2 * nb
    assert lib.sfamd_gate_table_size(_lib.OP_LOSS, D) == D**3
    assert lib.sfamd_gate_table_size(_lib.OP_BEAMSPLITTER, 41) == -1
    assert lib.sfamd_gate_table_size(_lib.OP_SQUEEZE, 65) == -1
    for k in (_lib.OP_TWO_MODE_SQUEEZE, _lib.OP_MZGATE, _lib.OP_CROSS_KERR):
        assert lib.sfamd_gate_table_size(k, D) == 2 * nb
        assert lib.sfamd_gate_table_size(k, _lib.MAX_CUTOFF_BS) > 0
        assert lib.sfamd_gate_table_size(k, _lib.MAX_CUTOFF_BS + 1) == -1
    assert lib.sfamd_gate_table_size(_lib.OP_CUBIC_PHASE, D) == 3 * D * D
    assert lib.sfamd_gate_table_size(_lib.OP_CUBIC_PHASE, _lib.MAX_CUTOFF_V + 1) == -1
    assert lib.sfamd_gate_table_size(11, D) == -1


def _create(ops, n, D, mixed=0, dtype=_lib.C64, fuse=0):
    lib = _lib.load()
    arr = (_lib.OpDesc * max(1, len(ops)))()
    for i, (k, a, b) in enumerate(ops):
        arr[i].kind, arr[i].mode0, arr[i].mode1 = k, a, b
    h = ctypes.c_void_p()
    st = lib.sfamd_prog_create(arr, len(ops), n, D, mixed, dtype, fuse, ctypes.byref(h))
    return st, h


def test_program_planning_and_validation():
    lib = _lib.load()
    D, n = 5, 3
    ops = [
        (_lib.OP_DISPLACEMENT, 0, -1),
        (_lib.OP_BEAMSPLITTER, 2, 0),
        (_lib.OP_LOSS, 1, -1),
        (_lib.OP_KERR, 2, -1),
    ]
    st, h = _create(ops, n, D)
    assert st == 0
    info = _lib.ProgInfo()
    assert lib.sfamd_prog_info(h, ctypes.byref(info)) == 0
    assert info.nops == 4 and info.final_mixed == 1 and info.nsteps == 4
    assert info.final_elems == D ** (2 * n)
    assert lib.sfamd_prog_state_elems(h, 0) == D**n
    assert lib.sfamd_prog_state_elems(h, 2) == D**n
    assert lib.sfamd_prog_state_elems(h, 3) == D ** (2 * n)
    assert info.ws_elems == 4 * D ** (2 * n)
    expect = 3 * D * D + D * (2 * D * D + 1) // 3 * 2 + D**3 + D
    assert info.table_elems == expect
    assert lib.sfamd_prog_destroy(h) == 0
    # invalid programs are rejected on the host with the documented codes
    assert _create([(_lib.OP_BEAMSPLITTER, 1, 1)], n, D)[0] == -1
    assert _create([(_lib.OP_ROTATION, 3, -1)], n, D)[0] == -1
    assert _create([(_lib.OP_ROTATION, 0, -1)], n, 65)[0] == -2
    assert _create([(_lib.OP_BEAMSPLITTER, 0, 1)], n, 41)[0] == -2
    assert _create([(11, 0, -1)], n, D)[0] == -3
    for k in (_lib.OP_TWO_MODE_SQUEEZE, _lib.OP_MZGATE, _lib.OP_CROSS_KERR):
        assert _create([(k, 0, 2)], n, D)[0] == 0
        assert _create([(k, 0, -1)], n, D)[0] == -1
        assert _create([(k, 1, 1)], n, D)[0] == -1
        assert _create([(k, 0, 1)], n, 41)[0] == -2
    assert _create([(_lib.OP_CUBIC_PHASE, 2, -1)], n, D)[0] == 0
    assert _create([(_lib.OP_CUBIC_PHASE, 0, -1)], n, 33)[0] == -2
    assert _create([(_lib.OP_MIX, -1, -1)], n, D, mixed=1)[0] == -4
    assert _create([], n, D, dtype=7)[0] == -1


def test_null_pointer_arguments_rejected_without_gpu_work():
    lib = _lib.load()
    assert lib.sfamd_apply_1mode(_lib.OP_ROTATION, None, None, None, 2, 4, 0, 0, None, 0, None) == -1
    assert lib.sfamd_norm(None, 2, 4, 0, None, None, 0, None) == -1
    assert lib.sfamd_graph_begin(None) == -1


def test_measurement_entry_points_validate_on_the_host():
    """Bad arguments of the measurement / preparation entry points are
    rejected before any launch (dummy pointers are never dereferenced)."""
    lib = _lib.load()
    a, b = ctypes.c_void_p(16), ctypes.c_void_p(32)
    I = ctypes.c_int
    assert lib.sfamd_marginal_partials(3, 5, 2) == 25
    assert lib.sfamd_marginal_partials(3, 5, 4) == -1
    # unsorted modes, mode out of range
    assert lib.sfamd_marginal_probs(a, 3, 5, 0, (I * 2)(2, 0), 2, b, b, 0, None) == -1
    assert lib.sfamd_marginal_probs(a, 3, 5, 0, (I * 1)(3), 1, b, b, 0, None) == -1
    # in == out, outcome outside the cutoff, repeated mode
    assert lib.sfamd_project_reset(a, a, 3, 5, 0, (I * 1)(0), (I * 1)(1), 1, 1.0, 0, None) == -1
    assert lib.sfamd_project_reset(a, b, 3, 5, 0, (I * 1)(0), (I * 1)(5), 1, 1.0, 0, None) == -1
    assert lib.sfamd_project_reset(a, b, 3, 5, 1, (I * 2)(1, 1), (I * 2)(0, 0), 2, 1.0, 0, None) == -1
    # repeated mode, ket embedding of a strict subset, missing rest
    assert lib.sfamd_embed_modes(a, a, b, 3, 4, (I * 2)(1, 1), 2, 1, 0, None) == -1
    assert lib.sfamd_embed_modes(a, a, b, 3, 4, (I * 2)(0, 1), 2, 0, 0, None) == -1
    assert lib.sfamd_embed_modes(None, a, b, 3, 4, (I * 2)(0, 1), 2, 1, 0, None) == -1
    # cutoff above the homodyne limit, zero bins
    assert lib.sfamd_homodyne_pdf(a, 65, 0.0, 10.0, 100, 2.0, b, 0, None) == -1
    assert lib.sfamd_homodyne_pdf(a, 5, 0.0, 10.0, 0, 2.0, b, 0, None) == -1
    assert lib.sfamd_apply_matrix(a, a, None, 2, 4, 0, 0, b, 0, None) == -1


# ------------------------------------------------------------------ front end
class RecordingBackend:
    short_name = "rec"

    def __init__(self):
        self.calls = []

    def begin_circuit(self, num_subsystems, **kw):
        self.calls.append(("begin", num_subsystems, kw))
        self.cutoff = kw.get("cutoff_dim")

    def get_cutoff_dim(self):
        return self.cutoff

    def __getattr__(self, name):
        def f(*args, **kw):
            self.calls.append((name,) + args + ((kw,) if kw else ()))

        return f

    def state(self, modes=None):
        return "state"


def test_gate_apply_rules():
    prog = frontend.Program(2)
    with prog.context as q:
        frontend.Dgate(0.0, 0.3) | q[0]  # skipped: first parameter zero
        frontend.Dgate(0.2, 0.3).H | q[0]  # dagger negates r
        frontend.Sgate(0.1) | q[1]
        frontend.BSgate(0.4, 0.1) | (q[1], q[0])
        frontend.LossChannel(0.9) | q[0]
    be = RecordingBackend()
    eng = frontend.Engine(be, backend_options={"cutoff_dim": 4})
    res = eng.run(prog)
    assert res.state == "state"
    assert be.calls[0] == ("begin", 2, {"cutoff_dim": 4})
    assert be.calls[1] == ("displacement", -0.2, 0.3, 0)
    assert be.calls[2] == ("squeeze", 0.1, 0.0, 1)
    assert be.calls[3] == ("beamsplitter", 0.4, 0.1, 1, 0)
    assert be.calls[4] == ("loss", 0.9, 0)


def test_two_mode_and_cubic_gate_calls():
    """S2gate, MZgate, CKgate, Vgate reach the backend calls of
    strawberryfields/ops.py (Vgate with gamma sqrt(hbar / 2), hbar = 2); the dagger
    negates the first parameter, and a zero first parameter skips the gate
    (ops.py:485-501), MZgate(0, phi_ex) included, as in the reference."""
    prog = frontend.Program(3)
    with prog.context as q:
        frontend.S2gate(0.3, 0.2) | (q[2], q[0])
        frontend.MZgate(0.5, 0.7).H | (q[0], q[1])
        frontend.MZgate(0.0, 0.7) | (q[0], q[1])
        frontend.CKgate(0.4) | (q[1], q[2])
        frontend.Vgate(0.6) | q[1]
        frontend.S2gate(0.3) | (q[0], q[1])
    be = RecordingBackend()
    frontend.Engine(be, backend_options={"cutoff_dim": 4}).run(prog)
    assert be.calls[1:] == [
        ("two_mode_squeeze", 0.3, 0.2, 2, 0),
        ("mzgate", -0.5, 0.7, 0, 1),
        ("cross_kerr_interaction", 0.4, 1, 2),
        ("cubic_phase", 0.6, 1),
        ("two_mode_squeeze", 0.3, 0.0, 0, 1),
    ]


def test_free_parameters_bind_at_run():
    prog = frontend.Program(1)
    a = prog.params("a")
    with prog.context as q:
        frontend.Rgate(a) | q[0]
    be = RecordingBackend()
    frontend.Engine(be, backend_options={"cutoff_dim": 3}).run(prog, args={"a": 0.25})
    assert be.calls[-1] == ("rotation", 0.25, 0)


def test_modemap_delete_and_add():
    mm = ModeMap(4)
    mm.delete(1)
    assert mm.show() == [0, None, 1, 2]
    mm.add(1)
    assert mm.show() == [0, None, 1, 2, 3]
    assert not mm.valid([5])
    with pytest.raises(ValueError):
        mm.delete(7)


# ------------------------------------------------------------------ workloads
def test_workload_generators():
    c1 = circuits.config1()
    assert len(c1["ops"]) == 15 and c1["n"] == 3 and c1["D"] == 8
    c2 = circuits.config2()
    assert len(c2["ops"]) == 256
    c3 = circuits.config3()
    assert len(c3["ops"]) == 240 and circuits.n_params(c3["ops"]) == 408
    t = c3["target"].reshape(-1)
    assert abs(np.vdot(t, t) - 1) < 1e-12
    idx = np.indices([8] * 6).reshape(6, -1).sum(0)
    assert np.all(t[idx > 6] == 0)
    c4 = circuits.config4()
    assert len(c4["circuits"]) == 256
    assert all(c["D"] ** c["n"] <= 2**20 for c in c4["circuits"])
    c5 = circuits.config5()
    assert len(c5["ops"]) == 144 and sum(1 for o in c5["ops"] if o[0] == "LOSS") == 18
    # same seed -> same parameters
    assert circuits.config3()["ops"] == c3["ops"]


def _plan(ops, n, D, fuse, dtype=_lib.C64):
    lib = _lib.load()
    codes = [(_lib.OP_KINDS[name], m[0], m[1] if len(m) > 1 else -1) for name, m, _ in ops]
    st, h = _create(codes, n, D, dtype=dtype, fuse=fuse)
    assert st == 0
    info = _lib.ProgInfo()
    lib.sfamd_prog_info(h, ctypes.byref(info))
    buf = (ctypes.c_int32 * len(ops))()
    steps = []
    for i in range(info.nsteps):
        k = lib.sfamd_prog_step_ops(h, i, buf, len(ops))
        steps.append([buf[j] for j in range(k)])
    lib.sfamd_prog_destroy(h)
    return steps


@pytest.mark.parametrize("fuse", [0, 2, 3, 4])
def test_stage_planner_keeps_every_gate_and_mode_order(fuse):
    """The fused plan covers each op once, keeps per-mode order and local sets."""
    for cfg, dtype in [(circuits.config3(), _lib.C64), (circuits.config3(), _lib.C128), (circuits.config2(), _lib.C64)]:
        ops, n, D = cfg["ops"], cfg["n"], cfg["D"]
        steps = _plan(ops, n, D, fuse, dtype)
        flat = [i for st in steps for i in st]
        assert sorted(flat) == list(range(len(ops)))
        for m in range(n):  # gates on one mode keep their program order
            order = [i for i in flat if m in ops[i][1]]
            assert order == sorted(order)
        for st in steps:
            modes = set(mm for i in st for mm in ops[i][1])
            diag_run = len(st) > 1 and all(ops[i][0] in ("R", "K") for i in st)
            assert diag_run or len(modes) <= max(fuse, 2)
        if fuse == 0:
            # one op per step, except runs of consecutive R / K gates (one pass each)
            for st in steps:
                assert len(st) == 1 or (all(ops[i][0] in ("R", "K") for i in st) and st == list(range(st[0], st[-1] + 1)))
    c3 = circuits.config3()
    if fuse == 3:
        assert len(_plan(c3["ops"], 6, 8, 3)) < 60


def test_stage_planner_leaves_mixed_ops_unfused():
    rng = np.random.default_rng(1)
    ops = circuits.layered(3, 1, rng, loss=True)
    steps = _plan(ops, 3, 6, 3)
    first_loss = min(i for i, o in enumerate(ops) if o[0] == "LOSS")
    for st in steps:
        if len(st) > 1 and not all(ops[i][0] in ("R", "K") for i in st):
            assert all(i < first_loss for i in st)
    # from the first loss on, only runs of consecutive R / K gates share a step
    tail = [st for st in steps if st[0] >= first_loss]
    for st in tail:
        assert len(st) == 1 or (all(ops[i][0] in ("R", "K") for i in st) and st == list(range(st[0], st[-1] + 1)))
    assert any(len(st) > 1 for st in tail)


# ------------------------------------------------------------------ interferometer
def _mode_unitary(gates, N):
    G = np.identity(N, dtype=complex)
    for name, modes, p in gates:
        M = np.identity(N, dtype=complex)
        if name == "R":
            M[modes[0], modes[0]] = np.exp(1j * p[0])
        elif name == "MZ":
            # closed form of the Mach-Zehnder mode unitary (decompositions.py:391-396)
            m, n = modes
            si, co, e = np.sin(p[0] / 2), np.cos(p[0] / 2), np.exp(1j * p[1])
            M[m, m], M[m, n], M[n, m], M[n, n] = si * e, co, co * e, -si
            M[[m, m, n, n], [m, n, m, n]] *= 1j * np.exp(1j * p[0] / 2)
        else:
            m, n = modes
            c, s = np.cos(p[0]), np.sin(p[0])
            M[m, m], M[m, n], M[n, m], M[n, n] = c, -np.exp(-1j * p[1]) * s, np.exp(1j * p[1]) * s, c
        G = M @ G
    return G


def test_rectangular_mesh_reconstructs_the_unitary():
    """The Rgate / BSgate list of the rectangular mesh multiplies back to U
    (BS mode unitary [[cos, -e^{-i phi} sin], [e^{i phi} sin, cos]]) for Haar
    unitaries, permutations (the zero and swap branches) and diagonal phases;
    N (N - 1) / 2 beamsplitters; the identity compiles to nothing."""
    from scipy.stats import unitary_group
    from strawberryfields_amd.interferometer import interferometer_gates

    for N in [2, 3, 4, 5, 8]:
        U = unitary_group.rvs(N, random_state=10 + N)
        g = interferometer_gates(U)
        assert np.abs(_mode_unitary(g, N) - U).max() < 1e-12, N
        assert sum(1 for x in g if x[0] == "BS") == N * (N - 1) // 2
        assert np.abs(_mode_unitary(interferometer_gates(U, drop_identity=False), N) - U).max() < 1e-12
    for U in [np.eye(4)[[2, 0, 3, 1]], np.diag(np.exp(1j * np.array([0.3, -1.2, 2.0]))), np.eye(3)[[1, 0, 2]] * 1j]:
        assert np.abs(_mode_unitary(interferometer_gates(U), len(U)) - U).max() < 1e-12
    assert interferometer_gates(np.eye(5)) == []
    with pytest.raises(ValueError):
        interferometer_gates(np.ones((3, 3)))


def test_other_meshes_follow_their_decomposition_identities():
    """The decompositions behind the other meshes satisfy the identities the
    reference's tests state for them (tests/frontend/test_decompositions.py:
    225-330, 342-383), on Haar unitaries of size 2..20, the identity and the
    anti-identity (zero and swap branches): phase-end U = D prod T; Mach-Zehnder
    U = prod M^{-1} D prod M with phases in [0, 2 pi); symmetric U = D prod M;
    triangular U = prod T^{-1} D.  The compiled gate lists of the
    rectangular_phase_end and rectangular_symmetric meshes multiply back to U
    (MZ checked against its closed-form matrix); the triangular mesh compiles as
    the reference compiles it (R(phi), BS(theta, 0) per entry, which realises
    D prod T, not U: the reference's Interferometer test leaves it out)."""
    from scipy.stats import unitary_group
    from strawberryfields_amd import interferometer as it

    def prod(mats, start):
        q = start
        for m in mats:
            q = m @ q
        return q

    cases = [unitary_group.rvs(N, random_state=30 + N) for N in (2, 3, 4, 5, 8, 20)]
    cases += [np.identity(20), np.identity(20)[::-1], np.identity(3)[::-1], np.eye(4)[[2, 0, 3, 1]]]
    for U in cases:
        N = len(U)
        tl, d, none = it.rectangular_phase_end(U)
        assert none is None
        assert np.abs(np.diag(d) @ prod([it._t(*t) for t in tl], np.identity(N)) - U).max() < 1e-13
        r, d, l = it.rectangular_mz(U)
        for t in r + l:
            assert 0 <= t[2] < 2 * np.pi and 0 <= t[3] < 2 * np.pi
        q = prod([it.mach_zehnder(*t).conj().T for t in reversed(l)],
                 np.diag(d) @ prod([it.mach_zehnder(*t) for t in r], np.identity(N)))
        assert np.abs(q - U).max() < 1e-12
        tl, d, _ = it.rectangular_symmetric(U)
        for t in tl:
            assert 0 <= t[2] < 2 * np.pi and 0 <= t[3] < 2 * np.pi
        assert np.abs(np.diag(d) @ prod([it.mach_zehnder(*t) for t in tl], np.identity(N)) - U).max() < 1e-12
        tl, d, _ = it.triangular(U)
        assert len(tl) == N * (N - 1) // 2
        assert np.abs(prod([it._t_inv(*t) for t in tl], np.diag(d)) - U).max() < 1e-13
        if N <= 8 and not np.allclose(U, np.identity(N)):
            for mesh in ("rectangular_phase_end", "rectangular_symmetric"):
                g = it.interferometer_gates(U, mesh=mesh)
                assert np.abs(_mode_unitary(g, N) - U).max() < 1e-12, mesh
                assert np.abs(_mode_unitary(it.interferometer_gates(U, False, mesh=mesh), N) - U).max() < 1e-12
            assert all(x[0] == "MZ" for x in g[: N * (N - 1) // 2])
            g = it.interferometer_gates(U, False, mesh="triangular")
            realised = np.diag(d) @ prod([it._t(*t) for t in tl], np.identity(N))
            assert np.abs(_mode_unitary(g, N) - realised).max() < 1e-12
    for mesh in ("rectangular", "rectangular_phase_end", "rectangular_symmetric", "triangular"):
        assert it.interferometer_gates(np.eye(4), mesh=mesh) == []
    with pytest.raises(ValueError):
        it.interferometer_gates(cases[0], mesh="sun_compact")  # needs n >= 3


def test_compact_meshes_reconstruct_the_unitary():
    """The compact rectangular and triangular meshes of symmetric MZ blocks
    (decompositions.py:636-905, commands ops.py:2461-2499) and the SU(n)
    factorisation of sun_compact (decompositions.py:1073-1381, commands
    ops.py:2502-2526; n >= 3, with its special branches for columns holding a
    single nonzero entry, hit by the permutations): their gate lists
    (sMZ as BS(pi/4, pi/2), R, R, BS(pi/4, pi/2)) multiply back to U for Haar
    unitaries of size 2..8, permutations and diagonal phases, as the reference's
    own test of these meshes asserts (tests/frontend/test_ops_decompositions.py:
    218-262); the triangular mesh has N (N - 1) / 2 blocks; the block matrix is
    the sMZ's, e^{i sigma} [[sin d, cos d], [cos d, -sin d]] with phases sigma +- d."""
    from scipy.stats import unitary_group
    from strawberryfields_amd import interferometer as it

    s, d = 0.4, -0.9
    g = it.smz_gates((0, 1), s + d, s - d)
    assert np.abs(_mode_unitary(g, 2) - it._smz(0, s, d, 2)).max() < 1e-14
    cases = [unitary_group.rvs(N, random_state=50 + N) for N in (2, 3, 4, 5, 6, 7, 8)]
    cases += [np.eye(4)[[2, 0, 3, 1]], np.diag(np.exp(1j * np.array([0.3, -1.2, 2.0]))), np.identity(5),
              np.eye(5)[::-1], np.eye(3)[[1, 0, 2]] * 1j]
    for U in cases:
        N = len(U)
        for mesh in ("rectangular_compact", "triangular_compact") + (("sun_compact",) if N >= 3 else ()):
            g = it.interferometer_gates(U, mesh=mesh)
            assert np.abs(_mode_unitary(g, N) - U).max() < 1e-11, (mesh, N)
        ph = it.triangular_compact(U)
        assert len(ph["deltas"]) == N * (N - 1) // 2
    with pytest.raises(ValueError):
        it.rectangular_compact(np.ones((3, 3)))
    with pytest.raises(ValueError):
        it.interferometer_gates(cases[0], mesh="hexagonal")
    with pytest.raises(ValueError):
        it.rectangular_symmetric(np.ones((3, 3)))


def test_interferometer_frontend_issues_backend_calls():
    from scipy.stats import unitary_group

    U = unitary_group.rvs(3, random_state=4)
    prog = frontend.Program(3)
    with prog.context as q:
        frontend.Interferometer(U) | (q[2], q[0], q[1])
    be = RecordingBackend()
    frontend.Engine(be, backend_options={"cutoff_dim": 3}).run(prog)
    calls = be.calls[1:]
    assert calls and all(c[0] in ("rotation", "beamsplitter") for c in calls)
    # register order is respected: the mesh acts on (q2, q0, q1)
    regs = {c[-1] for c in calls if c[0] == "rotation"}
    assert regs <= {0, 1, 2}
    prog = frontend.Program(3)
    with prog.context as q:
        frontend.Interferometer(U, mesh="rectangular_symmetric") | q
    be = RecordingBackend()
    frontend.Engine(be, backend_options={"cutoff_dim": 3}).run(prog)
    assert {c[0] for c in be.calls[1:]} <= {"mzgate", "rotation"} and be.calls[1][0] == "mzgate"
    with pytest.raises(ValueError):
        frontend.Interferometer(U, mesh="hexagonal")


# ------------------------------------------------------------------ single-mode preparations
def _ladder_expm(L):
    from scipy.linalg import expm

    a = np.diag(np.sqrt(np.arange(1, L)), 1)
    ad = a.T

    def Dop(al):
        return expm(al * ad - np.conj(al) * a)

    def Sop(z):
        return expm(0.5 * (np.conj(z) * a @ a - z * ad @ ad))

    vac = np.zeros(L)
    vac[0] = 1
    return Dop, Sop, vac


def test_displaced_squeezed_and_gkp_expansions_match_the_operator_definitions():
    """fockstates.displaced_squeezed equals D(a) S(z)|0> computed with matrix
    exponentials of the ladder operators at cutoff 160 (then truncated), to
    5e-10: the reference's formula adds 1e-10 under a square root
    (fockbackend/ops.py:437), which the closed form keeps, so the agreement
    stops near that level; its zero branches (coherent state, squeezed vacuum)
    are exact.  The GKP expansion equals its definition (normalised sum over the
    lattice teeth of displaced squeezed states, ops.py:518-596) built the same
    operator way; theta = 0 / pi select the logical states, phi enters as e^{-i phi}."""
    from strawberryfields_amd import fockstates as fs

    Dop, Sop, vac = _ladder_expm(160)
    for rd, pd, rs, ps in [(0.5, 0.3, 0.4, -0.7), (1.2, -2.0, 0.2, 1.1), (0.3, 0.0, -0.5, 0.4), (2.0, 1.0, 1.0, 0.0)]:
        ex = (Dop(rd * np.exp(1j * pd)) @ Sop(rs * np.exp(1j * ps)) @ vac)[:30]
        assert np.abs(ex - fs.displaced_squeezed(rd, pd, rs, ps, 30)).max() < 5e-10
    for rd, pd, rs, ps in [(0.0, 0.0, 0.6, 0.2), (0.7, 0.4, 0.0, 0.0), (0.0, 0.0, 0.0, 0.0)]:
        ex = (Dop(rd * np.exp(1j * pd)) @ Sop(rs * np.exp(1j * ps)) @ vac)[:30]
        assert np.abs(ex - fs.displaced_squeezed(rd, pd, rs, ps, 30)).max() < 1e-14

    eps, cut = 0.3, 1e-12
    r = -0.5 * np.log(np.tanh(eps))
    zmax = int(np.ceil(np.sqrt(-0.25 / np.pi * np.log(cut) / np.tanh(eps))))

    def logical(mu):
        k = np.zeros(160, complex)
        for t in range(-zmax, zmax + 1):
            w = np.exp(-0.5 * np.pi * np.tanh(eps) * (mu + 2 * t) ** 2)
            k += w * (Dop(np.sqrt(np.pi / 2) * (2 * t + mu) / np.cosh(eps)) @ Sop(r) @ vac)
        return k[:40]

    k0, k1 = logical(0), logical(1)
    for th, ph in [(0.8, 0.5), (0.0, 0.0), (np.pi, 0.0), (np.pi / 2, -1.1)]:
        k = np.cos(th / 2) * k0 + np.sin(th / 2) * np.exp(-1j * ph) * k1
        got = fs.square_gkp(th, ph, eps, cut, 40)
        assert abs(np.linalg.norm(got) - 1) < 1e-14
        assert np.abs(k / np.linalg.norm(k) - got).max() < 1e-9, (th, ph)
    # the teeth of logical |0> come in pairs x, -x, and D(-x)S(r)|0> is the
    # parity image of D(x)S(r)|0>, so |0>_gkp has even photon numbers only
    g0 = fs.square_gkp(0.0, 0.0, eps, cut, 40)
    assert np.abs(g0[1::2]).max() < 1e-12


def test_preparation_ops_issue_their_backend_calls():
    """DisplacedSqueezed, Thermal, GKP and Catstate reach the backend as in the
    reference (ops.py:687-758, 825-968, 1106-1135); the cat-state ket is the
    normalised coherent superposition with the reference's parity phase."""
    prog = frontend.Program(2)
    with prog.context as q:
        frontend.DisplacedSqueezed(0.3, 0.1, 0.2, 0.4) | q[0]
        frontend.Thermal(0.5) | q[1]
        frontend.GKP([0.4, 0.2], epsilon=0.25) | q[0]
        frontend.Catstate(0.8, 0.3, 1) | q[1]
    be = RecordingBackend()
    frontend.Engine(be, backend_options={"cutoff_dim": 12}).run(prog)
    names = [c[0] for c in be.calls[1:]]
    assert names == ["prepare_displaced_squeezed_state", "prepare_thermal_state", "prepare_gkp", "prepare_ket_state"]
    ket = be.calls[-1][1]
    Dop, _, vac = _ladder_expm(80)
    al = 0.8 * np.exp(0.3j)
    want = ((Dop(al) @ vac) - (Dop(-al) @ vac))[:12]
    want /= np.sqrt(2 * (1 - np.exp(-2 * abs(al) ** 2)))
    assert np.abs(ket - want).max() < 1e-13


def test_smz_gate_decomposes_as_the_fock_compiler_does():
    """sMZgate(phi_in, phi_ex) reaches the backend as BS(pi/4, pi/2),
    R(phi_ex - pi/2) on the second mode, R(phi_in - pi/2) on the first,
    BS(pi/4, pi/2) (ops.py:2015-2031)."""
    prog = frontend.Program(3)
    with prog.context as q:
        frontend.sMZgate(0.7, -0.2) | (q[2], q[0])
    be = RecordingBackend()
    frontend.Engine(be, backend_options={"cutoff_dim": 3}).run(prog)
    calls = be.calls[1:]
    assert [c[0] for c in calls] == ["beamsplitter", "rotation", "rotation", "beamsplitter"]
    assert calls[0][1:] == (np.pi / 4, np.pi / 2, 2, 0) and calls[3][1:] == (np.pi / 4, np.pi / 2, 2, 0)
    assert calls[1][-1] == 0 and abs(calls[1][1] - (-0.2 - np.pi / 2)) < 1e-15
    assert calls[2][-1] == 2 and abs(calls[2][1] - (0.7 - np.pi / 2)) < 1e-15


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message and a
    non-zero status before touching a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)


def test_batch_size_option_is_rejected_clearly():
    """The TF backend's batch_size axis (tfbackend/backend.py:84-106) is not
    offered: begin_circuit raises NotImplementedError instead of silently
    running unbatched."""
    from strawberryfields_amd import HIPFockBackend

    be = HIPFockBackend()
    with pytest.raises(NotImplementedError, match="batch_size"):
        be.begin_circuit(2, cutoff_dim=4, batch_size=3)


def test_planner_handles_gates_too_large_for_the_occupancy_budget():
    """At high cutoffs the compact beamsplitter tables of an adjoint stage
    exceed the LDS a workgroup may take when a launch needs two per compute
    unit; the planner then gives that stage the full budget (or runs the gate
    unfused) instead of producing an empty stage.  Every operation appears in
    exactly one step, in program order within each step."""
    lib = _lib.load()
    kinds = {"R": _lib.OP_ROTATION, "K": _lib.OP_KERR, "D": _lib.OP_DISPLACEMENT, "S": _lib.OP_SQUEEZE,
             "BS": _lib.OP_BEAMSPLITTER}
    rng = np.random.default_rng(5)
    for n, D, fuse in [(4, 19, 2), (4, 20, 2), (4, 20, 3), (5, 16, 3), (6, 8, 3), (5, 8, 4)]:
        ops = [(kinds[name], m[0], m[1] if len(m) > 1 else -1) for name, m, _ in circuits.layered(n, 1, rng)]
        for dtype in (_lib.C64, _lib.C128):
            st, h = _create(ops, n, D, dtype=dtype, fuse=fuse)
            assert st == 0, (n, D, fuse)
            info = _lib.ProgInfo()
            assert lib.sfamd_prog_info(h, ctypes.byref(info)) == 0
            seen = []
            buf = (ctypes.c_int32 * len(ops))()
            for s in range(info.nsteps):
                k = lib.sfamd_prog_step_ops(h, s, buf, len(ops))
                assert k >= 1
                step = [buf[j] for j in range(k)]
                assert step == sorted(step)
                seen += step
            assert sorted(seen) == list(range(len(ops))), (n, D, fuse)
            assert lib.sfamd_prog_destroy(h) == 0


def test_batch_branches_are_dealt_by_work():
    """The batch evaluator's dealing of circuits over graph branches: every
    circuit exactly once, deterministic, and largest-first to the least-loaded
    branch, so the heaviest branch exceeds the lightest by at most the largest
    item (the greedy bound; a round-robin deal of the same heavy-tailed sizes
    breaks it)."""
    from strawberryfields_amd import runtime

    rng = np.random.default_rng(5)
    works = list(rng.pareto(1.5, size=256) * 10 + 1)
    for nb in (1, 3, 8, 64, 300):
        br = runtime.deal_by_work(works, nb)
        assert len(br) == nb
        assert sorted(k for b in br for k in b) == list(range(len(works)))
        loads = [sum(works[k] for k in b) for b in br]
        assert max(loads) - min(loads) <= max(works) + 1e-9, nb
        assert br == runtime.deal_by_work(works, nb)
    rr = [[k for k in range(len(works)) if k % 8 == b] for b in range(8)]
    rr_loads = [sum(works[k] for k in b) for b in rr]
    assert max(rr_loads) - min(rr_loads) > max(works)
