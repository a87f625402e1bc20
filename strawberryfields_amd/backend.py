"""The HIP Fock backend: a ``BaseFock`` plugin on the MI355X library.

``HIPFockBackend`` has the method set of ``FockBackend``
(``strawberryfields/backends/fockbackend/backend.py:29-329``) and plugs into
``sf.Engine(backend=HIPFockBackend(), backend_options={...})``
(``strawberryfields/engine.py:71-78``).  Gates are recorded as the Engine issues
them (``engine.py:427-430``) and run lazily, as one planned circuit, at the next
readout or non-gate operation; the caller sees the same synchronous semantics
because every readout flushes first.  Parameters may be Python floats, numpy
scalars or torch tensors; when any tensor requires grad the circuit runs
through ``runtime.CircuitFunction`` and gradients flow back through the
hand-written adjoint kernels.

Backend options of ``begin_circuit`` / ``reset`` (``fockbackend/backend.py:96-128``):
``cutoff_dim`` (required), ``pure`` (default True), plus ``dtype``
(``torch.complex128`` default, as the fock backend; ``torch.complex64`` as the
TF backend), ``device`` (default ``"cuda"``) and ``checkpoint_segment``
(operations per checkpoint segment in the backward; None = automatic).
"""
import math
import numbers

import numpy as np
import torch

from . import fockstates, runtime, _lib
from .base import BaseFock, ModeMap
from .states import FockStateHIP

_DTYPES = {
    "complex64": torch.complex64,
    "complex128": torch.complex128,
    np.complex64: torch.complex64,
    np.complex128: torch.complex128,
    torch.complex64: torch.complex64,
    torch.complex128: torch.complex128,
}


def _is_zero(p):
    if isinstance(p, torch.Tensor):
        return False
    return np.all(np.asarray(p) == 0)


class HIPFockBackend(BaseFock):
    """Fock-basis simulator on one MI355X (HIP kernels via libsfamd)."""

    short_name = "hipfock"
    circuit_spec = "fock"

    def __init__(self):
        super().__init__()
        self._supported["mixed_states"] = True
        self._init_modes = None
        self._modemap = None
        self._state = None
        self._pure = True
        self._n = 0
        self._D = None
        self._dtype = torch.complex128
        self._device = None
        self._segment = None
        self._pending = []  # (name, modes, params)
        self._plans = {}

    # ------------------------------------------------------------ lifecycle
    def _remap_modes(self, modes):
        was_int = isinstance(modes, int)
        if was_int:
            modes = [modes]
        map_ = self._modemap.show()
        if not self._modemap.valid(modes) or any(map_[m] is None for m in modes):
            raise ValueError("The specified modes are not valid.")
        out = self._modemap.remap(modes)
        return out[0] if was_int else out

    def begin_circuit(self, num_subsystems, **kwargs):
        cutoff_dim = kwargs.get("cutoff_dim", None)
        pure = kwargs.get("pure", True)
        if cutoff_dim is None:
            raise ValueError("Argument 'cutoff_dim' must be passed to the Fock backend")
        if not isinstance(cutoff_dim, int):
            raise ValueError("Argument 'cutoff_dim' must be a positive integer")
        if not isinstance(num_subsystems, int):
            raise ValueError("Argument 'num_subsystems' must be a positive integer")
        if not isinstance(pure, bool):
            raise ValueError("Argument 'pure' must be either True or False")
        if cutoff_dim < 1 or cutoff_dim > _lib.MAX_CUTOFF:
            raise ValueError("Argument 'cutoff_dim' must be in [1, %d]" % _lib.MAX_CUTOFF)
        if kwargs.get("batch_size", None) is not None:
            # the TF backend's parameter-batch axis (tfbackend/backend.py:84-106,
            # tfbackend/circuit.py:129-173) has no counterpart here: batches of
            # circuits run through runtime.BatchEvaluator instead
            raise NotImplementedError("The hipfock backend does not support batch_size; "
                                      "use strawberryfields_amd.runtime.BatchEvaluator for many circuits")
        dtype = kwargs.get("dtype", torch.complex128)
        if dtype not in _DTYPES:
            raise ValueError("Argument 'dtype' must be complex64 or complex128")
        self._dtype = _DTYPES[dtype]
        self._device = torch.device(kwargs.get("device", "cuda"))
        self._segment = kwargs.get("checkpoint_segment", None)
        _lib.load()  # fail loudly here if the HIP library is missing
        self._init_modes = num_subsystems
        self._modemap = ModeMap(num_subsystems)
        self._D = cutoff_dim
        self._reset_state(num_subsystems, pure)

    def _reset_state(self, n, pure):
        self._n = n
        self._pure = pure
        self._pending = []
        shape = [self._D] * (n if pure else 2 * n)
        st = torch.zeros(shape, dtype=self._dtype, device=self._device)
        st.view(-1)[0] = 1.0
        self._state = st

    def reset(self, pure=True, **kwargs):
        self._D = kwargs.get("cutoff_dim", self._D)
        self._modemap.reset()
        self._reset_state(self._init_modes, pure)

    def add_mode(self, n=1, **kwargs):
        self._flush()
        D = self._D
        old = self._state
        k = n
        if self._pure:
            new = torch.zeros([D] * (self._n + k), dtype=old.dtype, device=old.device)
            new[(Ellipsis,) + (0,) * k] = old
        else:
            new = torch.zeros([D] * (2 * (self._n + k)), dtype=old.dtype, device=old.device)
            new[(Ellipsis,) + (0,) * (2 * k)] = old
        self._state = new
        self._n += k
        self._modemap.add(n)

    def del_mode(self, modes):
        self._flush()
        rm = self._remap_modes(modes)
        if isinstance(rm, int):
            rm = [rm]
        keep = [m for m in range(self._n) if m not in rm]
        if keep:
            self._state = runtime.reduced_dm(self._state, self._n, self._D, not self._pure, keep)
        else:
            self._state = runtime.norm2(self._state, self._n, self._D, not self._pure).to(self._dtype).reshape(())
        self._pure = False
        self._n = len(keep)
        self._modemap.delete(modes)

    def get_modes(self):
        return [i for i, j in enumerate(self._modemap._map) if j is not None]

    def get_cutoff_dim(self):
        return self._D

    # ------------------------------------------------------------ gates (recorded)
    def _record(self, name, modes, params):
        self._pending.append((name, tuple(modes), tuple(params)))

    def rotation(self, phi, mode):
        self._record("R", [self._remap_modes(mode)], [phi])

    def kerr_interaction(self, kappa, mode):
        self._record("K", [self._remap_modes(mode)], [kappa])

    def displacement(self, r, phi, mode):
        self._record("D", [self._remap_modes(mode)], [r, phi])

    def squeeze(self, r, phi, mode):
        self._record("S", [self._remap_modes(mode)], [r, phi])

    def beamsplitter(self, theta, phi, mode1, mode2):
        if self._D > _lib.MAX_CUTOFF_BS:
            raise ValueError("beamsplitter supports cutoff_dim <= %d" % _lib.MAX_CUTOFF_BS)
        self._record("BS", [self._remap_modes(mode1), self._remap_modes(mode2)], [theta, phi])

    def loss(self, T, mode):
        self._record("LOSS", [self._remap_modes(mode)], [T])

    def _two_mode(self, name, modes, params):
        if self._D > _lib.MAX_CUTOFF_BS:
            raise ValueError("two-mode gates support cutoff_dim <= %d" % _lib.MAX_CUTOFF_BS)
        self._record(name, [self._remap_modes(m) for m in modes], params)

    def two_mode_squeeze(self, r, phi, mode1, mode2):
        """S2(r e^{i phi}) = exp(r (e^{i phi} a1^dag a2^dag - e^{-i phi} a1 a2))
        (fockbackend/backend.py:175-176)."""
        self._two_mode("S2", [mode1, mode2], [r, phi])

    def mzgate(self, phi_in, phi_ex, mode1, mode2):
        """Mach-Zehnder interferometer (fockbackend/backend.py:181-182)."""
        self._two_mode("MZ", [mode1, mode2], [phi_in, phi_ex])

    def cross_kerr_interaction(self, kappa, mode1, mode2):
        """exp(i kappa n1 n2) (fockbackend/backend.py:285-288)."""
        self._two_mode("CK", [mode1, mode2], [kappa])

    def cubic_phase(self, gamma, mode):
        """expm(i gamma/6 (a + a^dag)^3) on the truncated basis, the hbar = 2
        matrix of fockbackend/ops.py:297-306 (fockbackend/backend.py:279-280)."""
        if self._D > _lib.MAX_CUTOFF_V:
            raise ValueError("cubic_phase supports cutoff_dim <= %d" % _lib.MAX_CUTOFF_V)
        self._record("V", [self._remap_modes(mode)], [gamma])

    # ------------------------------------------------------------ execution
    def _plan(self, ops, mixed_in):
        key = (tuple((n, m) for n, m, _ in ops), self._n, self._D, mixed_in, self._dtype, str(self._device))
        cc = self._plans.get(key)
        if cc is None:
            cc = runtime.CompiledCircuit(
                [(n, m) for n, m, _ in ops], self._n, self._D, mixed_in, self._dtype, self._device
            )
            if len(self._plans) > 64:
                self._plans.clear()
            self._plans[key] = cc
        return cc

    def _flush(self):
        if not self._pending:
            return
        ops = self._pending
        self._pending = []
        cc = self._plan(ops, not self._pure)
        params = runtime.params_tensor([p for _, _, p in ops], self._device)
        self._state = runtime.run_circuit(cc, params, self._state, self._segment)
        self._pure = not cc.final_mixed

    # ------------------------------------------------------------ preparations
    def _single_mode_dm(self, vec_or_dm):
        s = vec_or_dm
        if s.dim() == 1:
            s = torch.outer(s, s.conj())
        return s

    def _prepare(self, st, mode):
        """Replace internal mode `mode` by the single-mode ket/DM `st` (device tensor)."""
        self._flush()
        if self._state.requires_grad:
            raise NotImplementedError("state preparation after differentiable gates is not supported")
        st = st.to(device=self._device, dtype=self._dtype)
        if self._n == 1:
            self._state = st.clone()
            self._pure = st.dim() == 1
            return
        if self._pure:
            self._state = runtime.mix(self._state, self._n, self._D)
            self._pure = False
        self._state = runtime.replace_mode(self._state, self._n, self._D, mode, self._single_mode_dm(st))

    def _fock_vec(self, k):
        v = torch.zeros(self._D, dtype=self._dtype, device=self._device)
        if k < self._D:
            v[k] = 1.0
        return v

    def prepare_vacuum_state(self, mode):
        self._prepare(self._fock_vec(0), self._remap_modes(mode))

    def prepare_fock_state(self, n, mode):
        self._prepare(self._fock_vec(n), self._remap_modes(mode))

    def prepare_coherent_state(self, r, phi, mode):
        U = runtime.gate_matrix("D", r, phi, self._D, self._dtype, self._device)
        self._prepare(U[:, 0].clone(), self._remap_modes(mode))

    def prepare_squeezed_state(self, r, phi, mode):
        U = runtime.gate_matrix("S", r, phi, self._D, self._dtype, self._device)
        self._prepare(U[:, 0].clone(), self._remap_modes(mode))

    def prepare_displaced_squeezed_state(self, r_d, phi_d, r_s, phi_s, mode):
        """Amplitudes from the closed form (fockbackend/ops.py:420-446), not
        D[:, :D] @ S[:D, 0], whose product misses the levels >= D."""
        v = fockstates.displaced_squeezed(float(r_d), float(phi_d), float(r_s), float(phi_s), self._D)
        self._prepare(torch.as_tensor(v), self._remap_modes(mode))

    def prepare_gkp(self, state, epsilon, ampl_cutoff, representation="real", shape="square", mode=None):
        """Square-lattice GKP state cos(theta/2)|0> + e^{-i phi} sin(theta/2)|1>
        (fockbackend/backend.py:297-336, circuit.py:803-812); the complex
        representation and other lattices raise NotImplementedError as there."""
        if representation == "complex":
            raise NotImplementedError("The complex description of GKP is not implemented")
        if shape != "square":
            raise NotImplementedError("Only square GKP are implemented for now")
        theta, phi = state[0], state[1]
        v = fockstates.square_gkp(float(theta), float(phi), float(epsilon), float(ampl_cutoff), self._D)
        self._prepare(torch.as_tensor(v), self._remap_modes(mode))

    def prepare_thermal_state(self, nbar, mode):
        k = np.arange(self._D)
        if nbar == 0:
            diag = (k == 0).astype(float)
        else:
            diag = nbar**k / (nbar + 1) ** (k + 1)
        st = torch.diag(torch.as_tensor(diag, dtype=torch.float64)).to(self._dtype)
        self._prepare(st, self._remap_modes(mode))

    def prepare_ket_state(self, state, modes):
        self._prepare_array(state, modes, pure=True)

    def prepare_dm_state(self, state, modes):
        self._prepare_array(state, modes, pure=False)

    def _prepare_array(self, state, modes, pure):
        """Circuit.prepare_multimode (fockbackend/circuit.py:393-473): the listed
        modes (any order) take `state`, the others keep their reduced state, and
        the result is their product (a DM unless every mode is replaced).  The
        shape decides ket or DM: [D]*k or [D^k] is a ket, [D]*2k or
        [D^k, D^k] a DM."""
        rm = self._remap_modes(modes)
        if isinstance(rm, int):
            rm = [rm]
        if len(rm) != len(set(rm)):
            raise ValueError("The specified modes cannot appear multiple times.")
        st = torch.as_tensor(np.asarray(state) if not isinstance(state, torch.Tensor) else state)
        k = len(rm)
        D = self._D
        shape = tuple(st.shape)
        ket_shapes, dm_shapes = ((D,) * k, (D**k,)), ((D,) * (2 * k), (D**k, D**k))
        if shape not in ket_shapes + dm_shapes:
            raise ValueError("Incorrect shape for state preparation")
        # the shapes coincide only at D = 1; then the method's kind decides
        is_pure = shape in ket_shapes and (pure or shape not in dm_shapes)
        st = st.to(device=self._device, dtype=self._dtype).reshape([D] * (k if is_pure else 2 * k))
        if k == 1:
            self._prepare(st.reshape([D] if is_pure else [D, D]), rm[0])
            return
        self._flush()
        if self._state.requires_grad:
            raise NotImplementedError("state preparation after differentiable gates is not supported")
        n = self._n
        if k == n:
            if rm != list(range(n)):
                st = runtime.embed_modes(None, st, n, D, rm, not is_pure)
            self._state = st.clone() if rm == list(range(n)) else st
            self._pure = is_pure
            return
        sigma = runtime.mix(st, k, D) if is_pure else st
        rest_modes = [m for m in range(n) if m not in rm]
        rest = runtime.reduced_dm(self._state, n, D, not self._pure, rest_modes)
        self._state = runtime.embed_modes(rest, sigma, n, D, rm, True)
        self._pure = False

    # ------------------------------------------------------------ measurements
    def _measured_state(self):
        self._flush()
        if self._state.requires_grad:
            raise NotImplementedError("measurement of a state that requires grad is not supported")
        return self._state

    def _marginal_at(self, st, modes, outcome):
        """Unnormalised probability of `outcome` on `modes` (any order)."""
        srt = sorted(modes)
        p = runtime.marginal_probs(st, self._n, self._D, not self._pure, srt)
        idx = tuple(int(outcome[list(modes).index(m)]) for m in srt)
        return float(p[idx].item())

    def measure_fock(self, modes, shots=1, select=None, **kwargs):
        """Fock measurement with optional post-selection (fockbackend/backend.py:290-295,
        circuit.py:623-715): the outcome is drawn with numpy's global generator from the
        marginal distribution of the measured modes (or taken from `select`), the
        measured modes are projected and reset to vacuum, and the state is
        renormalised.  As there, `select` must hold an integer for every mode."""
        if shots != 1:
            raise NotImplementedError(
                "{} backend currently does not support shots != 1 for Fock measurement".format(self.short_name)
            )
        rm = self._remap_modes(modes)
        if isinstance(rm, int):
            rm = [rm]
        if select is not None and np.any(np.array(select) == None):  # noqa: E711
            raise NotImplementedError("Post-selection lists must only contain numerical values.")
        st = self._measured_state()
        n, D, mixed = self._n, self._D, not self._pure
        if select is not None:
            if len(select) != len(rm):
                raise ValueError(
                    "When performing post-selection, the number of "
                    "selected values (including None) must match the number of measured modes"
                )
            if not all(isinstance(v, int) for v in select):
                raise TypeError("The post-select list elements either be integers or None")
            if any(v < 0 or v >= D for v in select):
                raise IndexError("post-selected photon numbers must lie in [0, cutoff_dim)")
            outcome = list(select)
        else:
            srt = sorted(rm)
            raw = np.ravel(runtime.marginal_probs(st, n, D, mixed, srt).cpu().numpy())
            dist = raw * ~np.isclose(raw, 0.0)
            if sum(dist) != 1:
                i = np.random.choice(list(range(len(dist))), p=dist / sum(dist))
            else:
                i = np.random.choice(list(range(len(dist))), p=dist)
            permuted = [i // D ** (len(srt) - 1 - j) % D for j in range(len(srt))]
            perm = np.argsort(rm)
            outcome = [0] * len(rm)
            for j in range(len(rm)):
                outcome[perm[j]] = permuted[j]
        # squared norm of the projected state: the outcome's unnormalised probability
        w = self._marginal_at(st, rm, outcome) if select is not None else float(raw[i])
        if w == 0:
            raise ZeroDivisionError("Measurement has zero probability.")
        self._state = runtime.project_reset(st, n, D, mixed, rm, outcome, 1 / (w if mixed else np.sqrt(w)))
        return np.array([outcome])

    def measure_homodyne(self, phi, mode, shots=1, select=None, **kwargs):
        """Homodyne measurement of x cos(phi) + p sin(phi) (fockbackend/backend.py:184-198,
        circuit.py:717-801), hbar = 2: the density of the mode's reduced DM on the
        linspace(-max, max, num_bins) grid (GPU), a numpy multinomial draw, then the
        projection |0><x_phi| (x_phi the truncated, displaced and rotated infinitely
        squeezed vacuum) and renormalisation.
        Keyword args: ``max`` (default 10), ``num_bins`` (default 100000)."""
        if shots != 1:
            raise NotImplementedError(
                "{} backend currently does not support shots != 1 for homodyne measurement".format(self.short_name)
            )
        m = self._remap_modes(mode)
        st = self._measured_state()
        n, D, mixed = self._n, self._D, not self._pure
        hbar = 2.0
        rho = runtime.reduced_dm(st, n, D, mixed, [m])
        if select is not None:
            if not isinstance(select, numbers.Number):
                raise TypeError("Selected measurement result must be of numeric type.")
            sample = float(select)
        else:
            q_mag = kwargs.get("max", 10)
            num_bins = kwargs.get("num_bins", 100000)
            probs = runtime.homodyne_pdf(rho, D, phi, q_mag, num_bins, hbar).cpu().numpy()
            q = np.linspace(-q_mag, q_mag, num_bins)
            probs = probs / np.sum(probs)
            probs[np.abs(probs) < 1e-10] = 0
            hist = np.random.multinomial(1, probs)
            sample = q[list(hist).index(1)]
        alpha = sample * np.sqrt(1 / hbar / 2)
        k = np.arange(D)
        vac = np.array(
            [(-0.5) ** (j // 2) * np.sqrt(math.factorial(j)) / math.factorial(j // 2) if j % 2 == 0 else 0.0 for j in k]
        )
        Dm = runtime.gate_matrix("D", abs(alpha), np.angle(alpha), D, self._dtype, self._device)
        rot = torch.as_tensor(np.exp(1j * phi * k), dtype=self._dtype, device=self._device)
        e = rot * (Dm @ torch.as_tensor(vac, dtype=self._dtype, device=self._device))
        w = float(torch.vdot(e, rho.reshape(D, D) @ e).real.item())
        if w <= 0:
            raise ZeroDivisionError("Measurement has zero probability.")
        P = torch.zeros((D, D), dtype=self._dtype, device=self._device)
        P[0] = e.conj() / np.sqrt(w)
        self._state = runtime.apply_matrix(st, n, D, mixed, m, P)
        return np.array([[sample]])

    # ------------------------------------------------------------ queries
    def is_vacuum(self, tol=0.0, **kwargs):
        self._flush()
        first = self._state.reshape(-1)[0]
        fid = first.abs().square() if self._pure else first.real
        return bool(abs(float(fid.item()) - 1) <= tol)

    def state(self, modes=None, **kwargs):
        """The current state (fockbackend/backend.py:209-264); `modes` gives a reduced DM."""
        self._flush()
        s = self._state
        pure = self._pure
        n = self._n
        if modes is None:
            modes = list(range(n))
            red = s
        else:
            if isinstance(modes, int):
                modes = [modes]
            modes = list(modes)
            if len(modes) != len(set(modes)):
                raise ValueError("The specified modes cannot be duplicated.")
            if len(modes) > n:
                raise ValueError(
                    "The number of specified modes cannot be larger than the number of subsystems."
                )
            srt = sorted(modes)
            red = runtime.reduced_dm(s, n, self._D, not pure, srt)
            pure = False
            if modes != srt:
                # same axis order as fockbackend/backend.py:252-256
                mode_perm = np.argsort(modes)
                index_perm = [2 * x + i for x in mode_perm for i in (0, 1)]
                red = red.permute(*[int(v) for v in np.argsort(index_perm)]).contiguous()
        names = ["q[{}]".format(i) for i in np.array(self.get_modes())[modes]]
        return FockStateHIP(red, len(modes), pure, self._D, names)

    # ------------------------------------------------------------ not on this path
    def _not_here(self, *args, **kwargs):
        raise NotImplementedError("not implemented by the HIP Fock backend")

    thermal_loss = measure_threshold = _not_here
