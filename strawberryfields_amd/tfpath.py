"""Comparator: the TF backend's gradient path restated in torch on the same GPU.

This is a measurement comparator, not part of the simulator.  It follows the
algorithm of ``strawberryfields/backends/tfbackend`` (TensorFlow is not
available here):

* every gate becomes a dense matrix: R and K through ``exp`` of the diagonal
  with autodiff (``tfbackend/ops.py:221-243``), D, S and BS through a custom
  gradient whose backward contracts the cotangent with the matrix derivatives,
  ``dL/dp = Re sum conj(g) dG/dp`` (``tfbackend/ops.py:190-204,318-366``), and
  the loss channel through the [D]^4 superoperator built from T with autodiff
  (``loss_superop``, ``tfbackend/ops.py:288-315``);
* a single-mode gate contracts one axis of a ket, or the bra axis with G and
  the ket axis with conj(G) of a density matrix (``single_mode_gate``,
  ``tfbackend/ops.py:688-762``); a beamsplitter is the dense [D]^4 matrix on
  two axes (``two_mode_gate``, ``tfbackend/ops.py:765-854``); the loss
  superoperator acts on a mode's (bra, ket) axis pair after ``mix``
  (``single_mode_superop`` 971-1019, ``mix`` 102-129);
* the forward keeps every intermediate state for autodiff
  (``tfbackend/circuit.py:72-76``) and the backward is the framework's tape.

Each contraction is written as TF's einsum lowers it -- gate axes moved last,
one matrix product against the flattened state, axes moved back -- rather
than through ``torch.einsum``, whose complex path chose a slow GEMM layout
here (1.2 ms per 8x8-by-8^5 contraction).

Gate matrices: the TF backend builds D, S and BS on the host (thewalrus) and
copies each to the device per call, after reading the parameters back
(``.numpy()``).  Here the library's FP64 builder makes the matrix on the
device, and the matrix then makes the same round trip through host memory
(device -> host -> device, synchronous) so the per-gate transfer and sync of
the TF path stay in the measured step.
"""
import math

import numpy as np
import torch

from . import runtime


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


def _host_round_trip(t):
    """Device -> host -> device copy, as the host-built TF matrices travel."""
    return t.cpu().to(t.device)


class _GateMatrix(torch.autograd.Function):
    """Dense matrix of D, S (D x D) or BS ([D]^4) with a custom VJP."""

    @staticmethod
    def forward(ctx, p0, p1, kind, D, dtype, cache):
        dev = p0.device
        t = runtime.build_gate(kind, float(p0), float(p1), D, dtype, dev)  # parameters read back, as .numpy()
        t = _host_round_trip(t)
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


def loss_superop(T, D, dtype):
    """S[a, b, c, d] of the loss channel acting as |a><b| (.) |c><d|: a = b - l,
    d = c - l, weight (1-T)^l T^((b+c)/2 - l) sqrt(b! c!) / (sqrt((b-l)! (c-l)!) l!)
    (tfbackend/ops.py:288-315), differentiable in T."""
    dev = T.device
    r = np.arange(D)
    b, c, l = np.meshgrid(r, r, r, indexing="ij")
    ok = (l <= np.minimum(b, c))
    lf = np.array([math.factorial(int(v)) for v in range(D)], dtype=np.float64)
    fac = np.where(ok, np.sqrt(lf[b] * lf[c]) / (np.sqrt(lf[np.maximum(b - l, 0)] * lf[np.maximum(c - l, 0)]) * lf[l]), 0.0)
    expo = np.where(ok, (b + c) / 2.0 - l, 0.0)
    T = T.to(torch.float64)
    w = (1 - T) ** torch.as_tensor(l, dtype=torch.float64, device=dev) * T ** torch.as_tensor(expo, device=dev)
    w = w * torch.as_tensor(fac, device=dev)  # [b, c, l]
    S = torch.zeros((D, D, D, D), dtype=torch.float64, device=dev)
    bi, ci, li = (torch.as_tensor(x[ok], device=dev) for x in (b, c, l))
    S = S.index_put((bi - li, bi, ci, ci - li), w[bi, ci, li], accumulate=True)
    return S.to(dtype)


def _apply_axes(mat, state, axes):
    """out[..., i, ...] = sum_j mat[i, j] state[..., j, ...] on the listed axes
    (their joint index, first axis slowest): axes moved last, one matrix
    product, axes moved back."""
    nd = state.dim()
    rest = [a for a in range(nd) if a not in axes]
    perm = rest + list(axes)
    x = state.permute(perm).reshape(-1, mat.shape[1])
    y = (x @ mat.transpose(0, 1)).reshape([state.shape[a] for a in perm])
    inv = [0] * nd
    for i, a in enumerate(perm):
        inv[a] = i
    return y.permute(inv)


def _pair_matrix(mat4):
    """[D]^4 (out1, in1, out2, in2) -> D^2 x D^2 matrix on the joint (mode1, mode2) index."""
    D = mat4.shape[0]
    return mat4.permute(0, 2, 1, 3).reshape(D * D, D * D)


def mix(psi, n):
    """Pure -> mixed with interleaved axes [i0, j0, i1, j1, ...] (tfbackend/ops.py:102-129)."""
    D = psi.shape[0]
    rho = torch.outer(psi.reshape(-1), psi.reshape(-1).conj()).reshape([D] * (2 * n))
    perm = [x for m in range(n) for x in (m, n + m)]
    return rho.permute(perm)


class TFPathCircuit:
    """Circuit evaluated the TF-backend way (dense contractions + autograd), on a
    ket or, with `mixed`, a density matrix (loss channels supported)."""

    def __init__(self, ops, n, D, dtype, device, mixed=False):
        self.ops = ops
        self.n, self.D = n, D
        self.dtype = dtype
        self.mixed = mixed
        self.device = torch.device(device)
        pos, rel = _bs_dense_index(D)
        self.bs_cache = (torch.as_tensor(pos, device=self.device), torch.as_tensor(rel, device=self.device))
        self.k = torch.arange(D, dtype=torch.float64, device=self.device)

    def _single(self, mat, mode, state):
        if not self.mixed:
            return _apply_axes(mat, state, [mode])
        state = _apply_axes(mat, state, [2 * mode])
        return _apply_axes(mat.conj(), state, [2 * mode + 1])

    def _pair(self, mat4, m1, m2, state):
        M = _pair_matrix(mat4)
        if not self.mixed:
            return _apply_axes(M, state, [m1, m2])
        state = _apply_axes(M, state, [2 * m1, 2 * m2])
        return _apply_axes(M.conj(), state, [2 * m1 + 1, 2 * m2 + 1])

    def forward(self, params, state=None):
        """params: float64 [nops, 2] tensor (may require grad); returns the final
        state (the vacuum is the default input)."""
        n, D, dt = self.n, self.D, self.dtype
        if state is None:
            state = torch.zeros([D] * (2 * n if self.mixed else n), dtype=dt, device=self.device)
            state.view(-1)[0] = 1.0
        history = [state]
        for i, (name, modes) in enumerate(self.ops):
            p0, p1 = params[i, 0], params[i, 1]
            if name == "R":
                state = self._single(torch.diag(torch.exp(1j * p0 * self.k)).to(dt), modes[0], state)
            elif name == "K":
                state = self._single(torch.diag(torch.exp(1j * p0 * self.k**2)).to(dt), modes[0], state)
            elif name in ("D", "S"):
                state = self._single(_GateMatrix.apply(p0, p1, name, D, dt, None), modes[0], state)
            elif name == "BS":
                mat = _GateMatrix.apply(p0, p1, name, D, dt, self.bs_cache)
                state = self._pair(mat, modes[0], modes[1], state)
            elif name == "LOSS":
                if not self.mixed:
                    raise NotImplementedError("loss needs mixed=True (the TF path mixes the state first)")
                S = loss_superop(p0, D, dt)
                # rho'[.., a, d, ..] = sum_{b, c} S[a, b, c, d] rho[.., b, c, ..]
                state = _apply_axes(S.permute(0, 3, 1, 2).reshape(D * D, D * D), state, [2 * modes[0], 2 * modes[0] + 1])
            else:
                raise NotImplementedError(name)
            history.append(state)
        self.history = history
        return state

    def all_fock_probs(self, state):
        """Fock probabilities [D]^n (real part of the DM diagonal, or |psi|^2)."""
        n, D = self.n, self.D
        if not self.mixed:
            return state.abs() ** 2
        perm = [2 * m for m in range(n)] + [2 * m + 1 for m in range(n)]
        mat = state.permute(perm).reshape(D**n, D**n)
        return torch.diagonal(mat).real.reshape([D] * n)
