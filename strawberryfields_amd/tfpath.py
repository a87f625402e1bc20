"""Comparator: the TF backend's gradient path restated in torch on the same GPU.

This is a measurement comparator, not part of the simulator.  It follows the
algorithm of ``strawberryfields/backends/tfbackend`` (TensorFlow is not
available here):

* every gate becomes a dense matrix: R and K through ``exp`` of the diagonal
  with autodiff (``tfbackend/ops.py:221-243``), D, S and BS through a custom
  gradient whose backward contracts the cotangent with the matrix derivatives,
  ``dL/dp = Re sum conj(g) dG/dp`` (``tfbackend/ops.py:190-204,318-366``);
* a single-mode gate is an einsum over one axis of the state
  (``single_mode_gate``, ``tfbackend/ops.py:688-762``) and a beamsplitter a
  dense [D]^4 einsum over two axes (``two_mode_gate``, ``tfbackend/ops.py:765-854``);
* the forward keeps every intermediate state for autodiff
  (``tfbackend/circuit.py:72-76``) and the backward is the framework's tape.

The TF backend builds D, S and BS matrices on the host with thewalrus and
copies them to the device per gate.  Here they are built on the device by the
library's FP64 builder, which removes the host round trip and makes this
comparator faster than the path it stands for; speed-ups reported against it
are conservative.
"""
import string

import numpy as np
import torch

from . import runtime

_IDX = string.ascii_lowercase


def _bs_dense_index(D):
    """Flat [D]^4 positions (out1, in1, out2, in2) of the compact block entries."""
    pos, rel = [], []
    for s in range(2 * D - 1):
        lo, hi = max(0, s - D + 1), min(s, D - 1)
        for N in range(lo, hi + 1):
            for n in range(lo, hi + 1):
                pos.append(((N * D + n) * D + (s - N)) * D + (s - n))
                rel.append(n - N)
    return np.array(pos, dtype=np.int64), np.array(rel, dtype=np.float64)


class _GateMatrix(torch.autograd.Function):
    """Dense matrix of D, S (D x D) or BS ([D]^4) with a custom VJP."""

    @staticmethod
    def forward(ctx, p0, p1, kind, D, dtype, cache):
        dev = p0.device
        t = runtime.build_gate(kind, float(p0), float(p1), D, dtype, dev)  # host sync, as .numpy() in TF
        if kind in ("D", "S"):
            DD = D * D
            U = t[:DD].reshape(D, D)
            ctx.save_for_backward(t[DD : 2 * DD].reshape(D, D), t[2 * DD : 3 * DD].reshape(D, D))
            ctx.kind = kind
            return U.clone()
        pos, rel = cache
        nb = pos.numel()
        dense = torch.zeros(D**4, dtype=dtype, device=dev)
        dense[pos] = t[:nb]
        dth = torch.zeros(D**4, dtype=dtype, device=dev)
        dth[pos] = t[nb : 2 * nb]
        dph = torch.zeros(D**4, dtype=dtype, device=dev)
        dph[pos] = t[:nb] * (1j * rel.to(dtype))
        ctx.save_for_backward(dth.reshape([D] * 4), dph.reshape([D] * 4))
        ctx.kind = kind
        return dense.reshape([D] * 4)

    @staticmethod
    def backward(ctx, g):
        d0, d1 = ctx.saved_tensors
        g0 = (g.conj() * d0).sum().real
        g1 = (g.conj() * d1).sum().real
        return g0.to(torch.float64), g1.to(torch.float64), None, None, None, None


def _single_mode(mat, mode, state, n):
    """'ab,...b...->...a...' (tfbackend/ops.py:688-762, pure state)."""
    other = _IDX[2 : 2 + n]
    lhs = other[:mode] + "b" + other[mode + 1 :]
    rhs = other[:mode] + "a" + other[mode + 1 :]
    return torch.einsum("ab," + lhs + "->" + rhs, mat, state)


def _two_mode(mat, m1, m2, state, n):
    """'abcd,..b..d..->..a..c..' with the first pair on m1 (tfbackend/ops.py:765-854)."""
    other = list(_IDX[4 : 4 + n])
    lhs, rhs = list(other), list(other)
    lhs[m1], rhs[m1] = "b", "a"
    lhs[m2], rhs[m2] = "d", "c"
    return torch.einsum("abcd," + "".join(lhs) + "->" + "".join(rhs), mat, state)


class TFPathCircuit:
    """Pure-state circuit evaluated the TF-backend way (dense einsums + autograd)."""

    def __init__(self, ops, n, D, dtype, device):
        self.ops = ops
        self.n, self.D = n, D
        self.dtype = dtype
        self.device = torch.device(device)
        pos, rel = _bs_dense_index(D)
        self.bs_cache = (torch.as_tensor(pos, device=self.device), torch.as_tensor(rel, device=self.device))
        self.k = torch.arange(D, dtype=torch.float64, device=self.device)

    def forward(self, params):
        """params: float64 [nops, 2] tensor (may require grad); returns the final ket."""
        n, D, dt = self.n, self.D, self.dtype
        state = torch.zeros([D] * n, dtype=dt, device=self.device)
        state.view(-1)[0] = 1.0
        history = [state]
        for i, (name, modes) in enumerate(self.ops):
            p0, p1 = params[i, 0], params[i, 1]
            if name == "R":
                mat = torch.diag(torch.exp(1j * p0 * self.k)).to(dt)
                state = _single_mode(mat, modes[0], state, n)
            elif name == "K":
                mat = torch.diag(torch.exp(1j * p0 * self.k**2)).to(dt)
                state = _single_mode(mat, modes[0], state, n)
            elif name in ("D", "S"):
                mat = _GateMatrix.apply(p0, p1, name, D, dt, None)
                state = _single_mode(mat, modes[0], state, n)
            elif name == "BS":
                mat = _GateMatrix.apply(p0, p1, name, D, dt, self.bs_cache)
                state = _two_mode(mat, modes[0], modes[1], state, n)
            else:
                raise NotImplementedError(name)
            history.append(state)
        self.history = history
        return state
