"""State object returned by the HIP Fock backend.

Mirrors ``BaseFockState`` (``strawberryfields/backends/states.py:477-985``):
``ket``, ``dm``, ``trace``, ``all_fock_probs``, ``fock_prob``, ``reduced_dm``,
``mean_photon``, ``diagonal_expectation``, ``number_expectation``,
``parity_expectation``, ``fidelity_coherent``, ``fidelity_vacuum``,
``fidelity``.  Arrays are torch tensors on the GPU (like the TF backend's
``FockStateTF`` returns TF tensors, ``tfbackend/states.py:25-565``).  Reductions
run in the HIP library with FP64 accumulation; ``reduced_dm`` and
``mean_photon`` contract the traced modes straight from the ket instead of
materialising the D^2n density matrix (``states.py:613-663`` builds it first).

When the state tensor carries autograd history (a circuit with parameters
that require grad), ``all_fock_probs``, ``fock_prob``, ``trace``,
``mean_photon`` and the diagonal expectations are differentiable through the
library's probability kernel and its vector-Jacobian product kernel.
"""
import numpy as np
import torch

from . import runtime


class FockStateHIP:
    """Fock-basis state of ``num_modes`` modes at cutoff ``cutoff_dim`` on the GPU."""

    EQ_TOLERANCE = 1e-10

    def __init__(self, state_data, num_modes, pure, cutoff_dim, mode_names=None, hbar=2.0):
        self._data = state_data
        self._modes = num_modes
        self._pure = pure
        self._cutoff = cutoff_dim
        self._hbar = hbar
        self._basis = "fock"
        if mode_names is None:
            self._modemap = {i: "mode {}".format(i) for i in range(num_modes)}
        else:
            self._modemap = {i: "{}".format(j) for i, j in zip(range(num_modes), mode_names)}
        self._str = "<FockStateHIP: num_modes={}, cutoff={}, pure={}, hbar={}>".format(
            num_modes, cutoff_dim, pure, hbar
        )

    def __str__(self):
        return self._str

    __repr__ = __str__

    # -- properties (states.py:71-132, 524-532)
    @property
    def data(self):
        return self._data

    @property
    def hbar(self):
        return self._hbar

    @property
    def is_pure(self):
        return self._pure

    @property
    def num_modes(self):
        return self._modes

    @property
    def mode_names(self):
        return self._modemap

    @property
    def mode_indices(self):
        return {v: k for k, v in self._modemap.items()}

    @property
    def cutoff_dim(self):
        return self._cutoff

    @property
    def requires_grad(self):
        return self._data.requires_grad

    # -- representations
    def ket(self, **kwargs):
        return self._data if self._pure else None

    def dm(self, **kwargs):
        if not self._pure:
            return self._data
        if self._data.requires_grad:
            k = self._data.reshape(-1)
            rho = torch.outer(k, k.conj()).reshape([self._cutoff] * (2 * self._modes))
            perm = [i for m in range(self._modes) for i in (m, self._modes + m)]
            return rho.permute(perm)
        return runtime.mix(self._data, self._modes, self._cutoff)

    def _probs(self):
        return runtime.all_fock_probs(self._data, self._modes, self._cutoff, not self._pure)

    # -- readout
    def trace(self, **kwargs):
        """<psi|psi> or Tr rho (states.py:556-578)."""
        if self._data.requires_grad:
            return self._probs().to(torch.float64).sum()
        return float(runtime.norm2(self._data, self._modes, self._cutoff, not self._pure))

    def all_fock_probs(self, **kwargs):
        """Probabilities of every Fock basis state, shape [D]*n (states.py:580-608)."""
        return self._probs()

    def fock_prob(self, n, **kwargs):
        """Probability of the Fock basis state n (states.py:644-655)."""
        if len(n) != self._modes:
            raise ValueError("List length should be equal to number of modes")
        if max(n) >= self._cutoff:
            raise ValueError("Can't get distribution beyond truncation level")
        if self._data.requires_grad:
            return self._probs()[tuple(n)]
        if self._pure:
            return float(self._data[tuple(n)].abs().square().item())
        return float(self._data[tuple(n[i // 2] for i in range(2 * len(n)))].real.item())

    def reduced_dm(self, modes, **kwargs):
        """Reduced density matrix of the sorted `modes`, [D]*2k (states.py:613-642)."""
        if isinstance(modes, int):
            modes = [modes]
        modes = list(modes)
        if modes != sorted(modes):
            raise ValueError("The specified modes cannot be duplicated.")
        if len(modes) != len(set(modes)):
            raise ValueError("The specified modes cannot be duplicated.")
        if len(modes) > self._modes:
            raise ValueError(
                "The number of specified modes cannot be larger than the number of subsystems."
            )
        if modes == list(range(self._modes)):
            return self.dm()
        if self._data.requires_grad:
            return runtime.ReducedDM.apply(self._data, self._modes, self._cutoff, not self._pure, modes)
        return runtime.reduced_dm(self._data, self._modes, self._cutoff, not self._pure, modes)

    def diagonal_expectation(self, modes, values):
        """<prod_{m in modes} f(n_m)> for a number-diagonal f (states.py:947-966)."""
        if len(modes) != len(set(modes)):
            raise ValueError("There can be no duplicates in the modes specified.")
        if self._data.requires_grad:
            p = self._probs().to(torch.float64)
            w = torch.ones_like(p)
            vals = torch.as_tensor(np.asarray(values, dtype=np.float64), device=p.device)
            for m in modes:
                shape = [1] * self._modes
                shape[m] = -1
                w = w * vals.reshape(shape)
            return (p * w).sum()
        return float(
            runtime.diagonal_expectation(
                self._data, self._modes, self._cutoff, not self._pure, list(modes), values
            )
        )

    def number_expectation(self, modes):
        """(mean, variance) of prod n_m over `modes` (states.py:968-974)."""
        values = np.arange(self._cutoff)
        mean = self.diagonal_expectation(modes, values)
        var = self.diagonal_expectation(modes, values**2) - mean**2
        return mean, var

    def parity_expectation(self, modes):
        return self.diagonal_expectation(modes, (-1.0) ** np.arange(self._cutoff))

    def mean_photon(self, mode, **kwargs):
        """(mean, variance) of the photon number of `mode` (states.py:657-663)."""
        return self.number_expectation([mode])

    # -- fidelities (states.py:665-730); small readouts assembled with torch
    def _coherent(self, a):
        D = self._cutoff
        k = np.arange(D)
        from math import factorial

        v = np.array([np.exp(-0.5 * abs(a) ** 2) * a**j / np.sqrt(factorial(j)) for j in k])
        return torch.as_tensor(v, dtype=self._data.dtype, device=self._data.device)

    def fidelity_coherent(self, alpha_list, **kwargs):
        if not hasattr(alpha_list, "__len__"):
            alpha_list = [alpha_list]
        if len(alpha_list) != self._modes:
            raise ValueError("The number of alpha values must match the number of modes.")
        vecs = [self._coherent(a) for a in alpha_list]
        if self._pure:
            ov = self._data
            for v in vecs:
                ov = torch.tensordot(v.conj(), ov, dims=([0], [0]))
            return ov.abs().square()
        rho = self._data
        for v in vecs:
            rho = torch.tensordot(v.conj(), rho, dims=([0], [0]))
            rho = torch.tensordot(rho, v, dims=([0], [0]))
        return rho.real

    def fidelity_vacuum(self, **kwargs):
        return self.fidelity_coherent([0.0] * self._modes)

    def fidelity(self, other_state, mode, **kwargs):
        """<other| rho_mode |other> for a single-mode ket `other_state`."""
        rho = self.reduced_dm([mode])
        o = torch.as_tensor(np.asarray(other_state), dtype=rho.dtype, device=rho.device)
        return float((o.conj() @ rho @ o).real.item())
