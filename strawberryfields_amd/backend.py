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
import numpy as np
import torch

from . import runtime, _lib
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
        Dm = runtime.gate_matrix("D", r_d, phi_d, self._D, self._dtype, self._device)
        Sm = runtime.gate_matrix("S", r_s, phi_s, self._D, self._dtype, self._device)
        self._prepare(Dm @ Sm[:, 0], self._remap_modes(mode))

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
        rm = self._remap_modes(modes)
        if isinstance(rm, int):
            rm = [rm]
        st = torch.as_tensor(np.asarray(state) if not isinstance(state, torch.Tensor) else state)
        k = len(rm)
        D = self._D
        if k == self._n and rm == list(range(self._n)):
            self._flush()
            shape = [D] * (k if pure else 2 * k)
            self._state = st.to(device=self._device, dtype=self._dtype).reshape(shape).clone()
            self._pure = pure
            return
        if k != 1:
            raise NotImplementedError("preparing a state on a subset of several modes is not supported")
        st = st.reshape([D] if pure else [D, D])
        self._prepare(st, rm[0])

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

    thermal_loss = measure_homodyne = measure_fock = measure_threshold = prepare_gkp = _not_here
