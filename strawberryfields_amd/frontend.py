"""Minimal mirror of the Strawberry Fields front end for the Fock gate path.

Same names and behaviour as the reference for the operations this backend
runs: ``Program`` (``strawberryfields/program.py``), ``Engine`` /
``LocalEngine`` (``engine.py:48-572``), ``Result`` (``result.py``), free
parameters (``parameters.py``: ``prog.params`` / ``eng.run(args=...)``) and the
operations ``Sgate``, ``Dgate``, ``Rgate``, ``Kgate``, ``BSgate``, ``S2gate``,
``MZgate``, ``CKgate``, ``Vgate``, ``LossChannel``, ``Vacuum``, ``Fock``,
``Coherent``, ``Squeezed``, ``Ket``, ``DensityMatrix``, ``MeasureFock``,
``MeasureHomodyne`` / ``MeasureX`` / ``MeasureP``, ``Interferometer``
(``ops.py:541-2718``).  ``Gate.apply`` keeps the reference rules: a gate whose
first parameter is zero is skipped and a daggered gate negates its first
parameter (``ops.py:485-509``).  With Strawberry Fields installed, use
``sf.Engine(backend=HIPFockBackend())`` instead; this module exists so the
backend can be driven the same way without it.
"""
import numpy as np
import torch

_current = []


# hbar of the front end (sf.hbar default); only the cubic phase gate uses it
HBAR = 2.0


class RegRef:
    def __init__(self, ind):
        self.ind = ind

    def __repr__(self):
        return "q[{}]".format(self.ind)


class FreeParameter:
    """Named placeholder bound at run time (parameters.py:FreeParameter)."""

    def __init__(self, name):
        self.name = name
        self.val = None

    def __neg__(self):
        return _Expr(lambda v: -v, self)

    def __repr__(self):
        return "{" + self.name + "}"


class _Expr:
    def __init__(self, fn, par):
        self.fn, self.par = fn, par


def par_evaluate(p):
    """Value of a gate parameter: numbers pass through, free parameters resolve."""
    if isinstance(p, FreeParameter):
        if p.val is None:
            raise ValueError("free parameter {} is not bound".format(p.name))
        return p.val
    if isinstance(p, _Expr):
        return p.fn(par_evaluate(p.par))
    return p


class Command:
    def __init__(self, op, reg):
        self.op, self.reg = op, reg


class Program:
    """Quantum circuit on ``num_subsystems`` modes (program.py:77-901, the used subset)."""

    def __init__(self, num_subsystems):
        self.num_subsystems = num_subsystems
        self.register = tuple(RegRef(i) for i in range(num_subsystems))
        self.circuit = []
        self.free_params = {}

    @property
    def context(self):
        return self

    def __enter__(self):
        _current.append(self)
        return self.register

    def __exit__(self, *exc):
        _current.pop()

    def append(self, op, reg):
        self.circuit.append(Command(op, reg))
        return reg

    def params(self, *names):
        out = []
        for n in names:
            if n not in self.free_params:
                self.free_params[n] = FreeParameter(n)
            out.append(self.free_params[n])
        return out[0] if len(out) == 1 else out

    def bind_params(self, binding):
        for k, v in binding.items():
            if k not in self.free_params:
                raise ValueError("unknown free parameter {}".format(k))
            self.free_params[k].val = v


class Operation:
    ns = None

    def __init__(self, par):
        self.p = list(par)

    def __or__(self, reg):
        if isinstance(reg, RegRef):
            reg = [reg]
        reg = list(reg)
        if self.ns is not None and self.ns != len(reg):
            raise ValueError("Wrong number of subsystems.")
        if not _current:
            raise RuntimeError("operations must be applied inside a Program context")
        return _current[-1].append(self, reg)

    def apply(self, reg, backend, **kwargs):
        self._apply([r.ind for r in reg], backend, **kwargs)


class Gate(Operation):
    def __init__(self, par):
        super().__init__(par)
        self.dagger = False

    @property
    def H(self):
        import copy

        g = copy.copy(self)
        g.dagger = not self.dagger
        return g

    def apply(self, reg, backend, **kwargs):
        z = par_evaluate(self.p[0])
        if not isinstance(z, torch.Tensor) and np.all(np.asarray(z) == 0):
            return
        if self.dagger:
            z = -z
        orig = self.p[0]
        self.p[0] = z
        try:
            self._apply([r.ind for r in reg], backend, **kwargs)
        finally:
            self.p[0] = orig

    def _vals(self):
        return [par_evaluate(x) for x in self.p]


class Dgate(Gate):
    ns = 1

    def __init__(self, r, phi=0.0):
        super().__init__([r, phi])

    def _apply(self, reg, backend, **kwargs):
        r, phi = self._vals()
        backend.displacement(r, phi, *reg)


class Sgate(Gate):
    ns = 1

    def __init__(self, r, phi=0.0):
        super().__init__([r, phi])

    def _apply(self, reg, backend, **kwargs):
        r, phi = self._vals()
        backend.squeeze(r, phi, *reg)


class Rgate(Gate):
    ns = 1

    def __init__(self, theta):
        super().__init__([theta])

    def _apply(self, reg, backend, **kwargs):
        (theta,) = self._vals()
        backend.rotation(theta, *reg)


class Kgate(Gate):
    ns = 1

    def __init__(self, kappa):
        super().__init__([kappa])

    def _apply(self, reg, backend, **kwargs):
        (kappa,) = self._vals()
        backend.kerr_interaction(kappa, *reg)


class BSgate(Gate):
    ns = 2

    def __init__(self, theta=np.pi / 4, phi=0.0):
        super().__init__([theta, phi])

    def _apply(self, reg, backend, **kwargs):
        theta, phi = self._vals()
        backend.beamsplitter(theta, phi, *reg)


class S2gate(Gate):
    """Two-mode squeezing (strawberryfields/ops.py:2033-2078)."""

    ns = 2

    def __init__(self, r, phi=0.0):
        super().__init__([r, phi])

    def _apply(self, reg, backend, **kwargs):
        r, phi = self._vals()
        backend.two_mode_squeeze(r, phi, *reg)


class MZgate(Gate):
    """Mach-Zehnder interferometer (strawberryfields/ops.py:1951-2004); as in the
    reference (ops.py:485-501), the dagger negates the first parameter only and
    a zero phi_in skips the gate."""

    ns = 2

    def __init__(self, phi_in, phi_ex):
        super().__init__([phi_in, phi_ex])

    def _apply(self, reg, backend, **kwargs):
        phi_in, phi_ex = self._vals()
        backend.mzgate(phi_in, phi_ex, *reg)


class sMZgate(Operation):
    """Symmetric Mach-Zehnder interferometer with internal phases phi_in,
    phi_ex (ops.py:2015-2031); the fock compiler decomposes it, and so does this
    mirror: BS(pi/4, pi/2), R(phi_ex - pi/2) on the second mode, R(phi_in - pi/2)
    on the first, BS(pi/4, pi/2)."""

    ns = 2

    def __init__(self, phi_in, phi_ex):
        super().__init__([phi_in, phi_ex])

    def decompose(self, reg):
        from .interferometer import smz_gates

        out = []
        for name, modes, p in smz_gates((0, 1), par_evaluate(self.p[0]), par_evaluate(self.p[1])):
            op = Rgate(p[0]) if name == "R" else BSgate(p[0], p[1])
            out.append((op, [reg[m] for m in modes]))
        return out

    def apply(self, reg, backend, **kwargs):
        regs = [r if isinstance(r, RegRef) else RegRef(r) for r in reg]
        for op, rr in self.decompose(regs):
            op.apply(rr, backend, **kwargs)


class CKgate(Gate):
    """Cross-Kerr interaction exp(i kappa n1 n2) (strawberryfields/ops.py:2219-2257)."""

    ns = 2

    def __init__(self, kappa):
        super().__init__([kappa])

    def _apply(self, reg, backend, **kwargs):
        (kappa,) = self._vals()
        backend.cross_kerr_interaction(kappa, *reg)


class Vgate(Gate):
    """Cubic phase exp(i gamma x^3 / (3 hbar)) (strawberryfields/ops.py:1735-1779);
    the backend call takes gamma sqrt(hbar / 2)."""

    ns = 1

    def __init__(self, gamma):
        super().__init__([gamma])

    def _apply(self, reg, backend, **kwargs):
        (gamma,) = self._vals()
        backend.cubic_phase(gamma * np.sqrt(HBAR / 2), *reg)


class LossChannel(Operation):
    ns = 1

    def __init__(self, T):
        super().__init__([T])

    def _apply(self, reg, backend, **kwargs):
        backend.loss(par_evaluate(self.p[0]), *reg)


class Vacuum(Operation):
    ns = 1

    def __init__(self):
        super().__init__([])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_vacuum_state(*reg)


class Fock(Operation):
    ns = 1

    def __init__(self, n=0):
        super().__init__([n])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_fock_state(int(par_evaluate(self.p[0])), *reg)


class Coherent(Operation):
    ns = 1

    def __init__(self, r=0.0, phi=0.0):
        super().__init__([r, phi])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_coherent_state(par_evaluate(self.p[0]), par_evaluate(self.p[1]), *reg)


class Squeezed(Operation):
    ns = 1

    def __init__(self, r=0.0, p=0.0):
        super().__init__([r, p])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_squeezed_state(par_evaluate(self.p[0]), par_evaluate(self.p[1]), *reg)


class DisplacedSqueezed(Operation):
    """D(r_d e^{i phi_d}) S(r_s e^{i phi_s})|0> (ops.py:687-758)."""

    ns = 1

    def __init__(self, r_d=0.0, phi_d=0.0, r_s=0.0, phi_s=0.0):
        super().__init__([r_d, phi_d, r_s, phi_s])

    def _apply(self, reg, backend, **kwargs):
        p = [par_evaluate(x) for x in self.p]
        backend.prepare_displaced_squeezed_state(p[0], p[1], p[2], p[3], *reg)


class Thermal(Operation):
    """Thermal state with mean photon number n (ops.py:1106-1135)."""

    ns = 1

    def __init__(self, n=0):
        super().__init__([n])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_thermal_state(par_evaluate(self.p[0]), *reg)


class Catstate(Operation):
    """(|a> + e^{i pi p}|-a>) / N, a = a e^{i phi}, in the complex (Fock)
    representation (ops.py:825-931): the ket is formed on the host from the
    coherent-state amplitudes alpha^n / sqrt(n!) and N = sqrt(2 (1 + cos(pi p)
    e^{-2|a|^2})) e^{|a|^2/2}, then prepared with prepare_ket_state."""

    ns = 1

    def __init__(self, a=0.0, phi=0.0, p=0, representation="complex", ampl_cutoff=1e-12, D=2):
        super().__init__([a, phi, p, representation, ampl_cutoff, D])

    def _apply(self, reg, backend, **kwargs):
        import math

        a, phi, p = par_evaluate(self.p[0]), par_evaluate(self.p[1]), self.p[2]
        alpha = a * np.exp(1j * phi)
        theta = np.pi * p
        cutoff = backend.get_cutoff_dim()
        t = np.exp(-0.5 * abs(alpha) ** 2)
        N = t / np.sqrt(2 * (1 + np.cos(theta) * t**4))
        l = np.arange(cutoff)
        sq = np.array([math.exp(0.5 * math.lgamma(k + 1)) for k in l])
        ket = (alpha**l / sq + np.exp(1j * theta) * (-alpha) ** l / sq) * N
        backend.prepare_ket_state(ket, *reg)


class GKP(Operation):
    """Finite-energy square-lattice GKP state (ops.py:934-968)."""

    ns = 1

    def __init__(self, state=None, epsilon=0.2, ampl_cutoff=1e-12, representation="real", shape="square"):
        if state is None:
            state = [0, 0]
        super().__init__([state, epsilon, ampl_cutoff, representation, shape])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_gkp(*self.p, mode=reg[0])


class Interferometer(Operation):
    """Passive linear interferometer U (ops.py:2556-2718), compiled as the fock
    compiler does: Rgate and BSgate for the rectangular, rectangular_phase_end,
    triangular and the compact meshes (their sMZ and SU(2) blocks decomposed),
    MZgate and Rgate for rectangular_symmetric; an unknown mesh raises ValueError
    as in ops.py:2638-2651."""

    def __init__(self, U, mesh="rectangular", drop_identity=True, tol=1e-6):
        from .interferometer import MESHES

        U = np.asarray(U)
        super().__init__([U])
        self.ns = U.shape[0]
        if mesh not in MESHES:
            raise ValueError(f"Unknown mesh '{mesh}'. Mesh must be one of {set(MESHES)}")
        self.mesh, self.drop_identity, self.tol = mesh, drop_identity, tol

    def decompose(self, reg):
        """[(gate, registers)] in application order."""
        from .interferometer import interferometer_gates

        make = {"R": lambda p: Rgate(p[0]), "BS": lambda p: BSgate(p[0], p[1]), "MZ": lambda p: MZgate(p[0], p[1])}
        out = []
        for name, modes, p in interferometer_gates(self.p[0], self.drop_identity, self.tol, self.mesh):
            out.append((make[name](p), [reg[m] for m in modes]))
        return out

    def apply(self, reg, backend, **kwargs):
        regs = [r if isinstance(r, RegRef) else RegRef(r) for r in reg]
        for op, rr in self.decompose(regs):
            op.apply(rr, backend, **kwargs)


class Ket(Operation):
    """Prepare a ket on one or more modes (ops.py:957-1000)."""

    def __init__(self, state):
        super().__init__([state])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_ket_state(self.p[0], reg)


class DensityMatrix(Operation):
    """Prepare a density matrix on one or more modes (ops.py:1003-1050)."""

    def __init__(self, state):
        super().__init__([state])

    def _apply(self, reg, backend, **kwargs):
        backend.prepare_dm_state(self.p[0], reg)


class Measurement(Operation):
    """Measurement base (ops.py:262-330): the backend's values are stored on the
    register references (``reg.val``) and returned to the Engine."""

    def __init__(self, par, select=None):
        super().__init__(par)
        self.select = select

    def apply(self, reg, backend, **kwargs):
        if kwargs.get("shots", 1) is None:
            return None
        values = self._apply([r.ind for r in reg], backend, **kwargs)
        for v, r in zip(np.transpose(values), reg):
            r.val = v
        return values


class MeasureFock(Measurement):
    """Photon-counting measurement (ops.py:1147-1190), optional post-selection
    and Poisson dark counts."""

    def __init__(self, select=None, dark_counts=None):
        if dark_counts and select:
            raise NotImplementedError("Post-selection cannot be used together with dark counts.")
        if dark_counts is not None and not isinstance(dark_counts, (list, tuple)):
            dark_counts = [dark_counts]
        if select is not None and not isinstance(select, (list, tuple)):
            select = [select]
        self.dark_counts = dark_counts
        super().__init__([], select)

    def _apply(self, reg, backend, shots=1, **kwargs):
        samples = np.array(backend.measure_fock(reg, shots=shots, select=self.select, **kwargs))
        if self.dark_counts is not None:
            if len(self.dark_counts) != len(reg):
                raise ValueError(
                    "The number of dark counts must be equal to the number of measured modes: "
                    "{} != {}".format(len(self.dark_counts), len(reg))
                )
            samples += np.random.poisson(self.dark_counts, samples.shape)
        return samples


class MeasureHomodyne(Measurement):
    """Homodyne measurement of x cos(phi) + p sin(phi) (ops.py:1230-1283); the
    backend call is in hbar = 2 units, so values scale by sqrt(hbar / 2)."""

    ns = 1

    def __init__(self, phi, select=None):
        super().__init__([phi], select)

    def _apply(self, reg, backend, shots=1, **kwargs):
        s = np.sqrt(HBAR / 2)
        select = self.select
        if select is not None:
            select = select / s
        phi = par_evaluate(self.p[0])
        return s * backend.measure_homodyne(phi, *reg, shots=shots, select=select, **kwargs)


MeasureX = MeasureHomodyne(0)
MeasureP = MeasureHomodyne(np.pi / 2)


class Result:
    """Outcome of Engine.run (result.py): the final state and the samples of
    the last measurement of each mode, ``samples`` shaped (shots, measured
    modes in increasing order), ``samples_dict`` mode -> list of values."""

    def __init__(self, state, samples=None, samples_dict=None):
        self.state = state
        self.samples = np.empty((0, 0)) if samples is None else samples
        self.samples_dict = {} if samples_dict is None else samples_dict


class LocalEngine:
    """Engine.run semantics of engine.py:229-572 for a backend instance."""

    def __init__(self, backend, *, backend_options=None):
        self.backend = backend
        self.backend_name = backend.short_name
        self.backend_options = dict(backend_options or {})
        self._initialized = False
        self.run_progs = []

    def reset(self, backend_options=None):
        if backend_options is not None:
            self.backend_options.update(backend_options)
        if self._initialized:
            self.backend.reset(**self.backend_options)
        self.run_progs = []

    def run(self, program, *, args=None, compile_options=None, **kwargs):
        if kwargs.get("shots", 1) not in (1, None):
            raise NotImplementedError("only shots=1 is supported")
        if not self._initialized:
            self.backend.begin_circuit(num_subsystems=program.num_subsystems, **self.backend_options)
            self._initialized = True
        if args:
            program.bind_params(args)
        samples_dict = {}
        for cmd in program.circuit:
            val = cmd.op.apply(cmd.reg, self.backend)
            if val is not None:  # engine.py:427-442
                for i, r in enumerate(cmd.reg):
                    samples_dict.setdefault(r.ind, []).append(np.asarray(val)[:, i])
        self.run_progs.append(program)
        modes = kwargs.get("modes", None)
        samples = np.empty((0, 0))
        if samples_dict:  # engine.py:473-497: last value of each mode, sorted by mode
            last = {key: v[-1] for key, v in samples_dict.items()}
            samples = np.transpose([v for _, v in sorted(last.items())])
        samples_dict = {k: [np.array(i) for i in v] for k, v in samples_dict.items()}
        return Result(self.backend.state(modes=modes), samples, samples_dict)


Engine = LocalEngine
