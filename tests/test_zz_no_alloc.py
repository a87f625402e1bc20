"""The allocation-free fidelity loss (sfamd_fidelity_loss_noalloc) against the
workspace form (sfamd_fidelity_loss) through the C ABI, on two calls in a row
(the second on a smaller input, so the static partial-sum array holds the
first call's contents), in both dtypes, with and without the cotangent."""
import numpy as np
import pytest
import torch

from strawberryfields_amd import _lib

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _pair(n, dtype, seed):
    rng = np.random.default_rng(seed)
    a = torch.tensor(rng.normal(size=n) + 1j * rng.normal(size=n), dtype=dtype, device=DEV)
    b = torch.tensor(rng.normal(size=n) + 1j * rng.normal(size=n), dtype=dtype, device=DEV)
    return a / a.norm(), b / b.norm()


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_noalloc_fidelity_loss_matches_workspace_form(dtype):
    lib = _lib.load()
    code = _lib.dtype_code(dtype)
    stream = _lib.stream_ptr(torch.device(DEV))
    for n, seed in [(262144, 1), (1000, 2), (3, 3)]:  # largest first: the array keeps stale partials
        t, psi = _pair(n, dtype, seed)
        part = torch.empty(lib.sfamd_overlap_partials(n), dtype=torch.float64, device=DEV)
        out_ref = torch.empty(3, dtype=torch.float64, device=DEV)
        g_ref = torch.empty_like(psi)
        _lib.call("sfamd_fidelity_loss", _lib.ptr(t), _lib.ptr(psi), n, _lib.ptr(part), _lib.ptr(out_ref),
                  _lib.ptr(g_ref), code, stream)
        out = torch.empty(3, dtype=torch.float64, device=DEV)
        g = torch.empty_like(psi)
        _lib.call("sfamd_fidelity_loss_noalloc", _lib.ptr(t), _lib.ptr(psi), n, _lib.ptr(out), _lib.ptr(g), code,
                  stream)
        out2 = torch.empty(3, dtype=torch.float64, device=DEV)
        _lib.call("sfamd_fidelity_loss_noalloc", _lib.ptr(t), _lib.ptr(psi), n, _lib.ptr(out2), None, code, stream)
        torch.cuda.synchronize()
        assert torch.equal(out, out_ref), (n, out, out_ref)  # same kernels, same partial order: bit-identical
        assert torch.equal(out2, out_ref)
        assert torch.equal(g, g_ref)
        ov = torch.vdot(t.to(torch.complex128), psi.to(torch.complex128))
        assert abs(complex(out[0].item(), out[1].item()) - complex(ov)) < (1e-6 if dtype == torch.complex64 else 1e-12)
