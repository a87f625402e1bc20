"""Full-size parity of the benchmarked workloads (GPU).

* The 6-mode cutoff-8 4-layer circuit (262,144 amplitudes, 240 operations):
  forward ket against the complex128 oracle, and the gradient of
  1 - |<t|psi>|^2 against central finite differences of the oracle
  (tests/golden/c3_grad.npz), through the captured evaluator with fused
  stages of 3 local modes (the benchmarked kernels) and through autograd
  with one kernel per gate, in complex64 and complex128.
* The 3-mode cutoff-20 density matrix (64e6 entries, 144 operations with
  loss): forward observables against the oracle (tests/golden/c5_forward.npz).
* The 256-circuit uneven batch: the batched evaluator on the whole batch,
  with the autotuner choosing the tiles as the benchmark does, checked on
  eight circuits from the smallest to the largest (2^20 amplitudes) against
  the oracle's ket, its total mean photon number and central differences
  (tests/golden/c4_grad.npz).

Tolerances (north star): complex64 max|d psi| <= 1e-5 and gradient relative
error <= 1e-5 against differences at h = 1e-4; complex128 max|d psi| <= 1e-12
and gradient relative error <= 1e-8 against the Richardson-extrapolated
differences (4 g(h/2) - g(h)) / 3.
"""
import os

import numpy as np
import pytest
import torch

from oracle import fock_oracle as fo
from strawberryfields_amd import circuits, runtime

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
C128, C64 = torch.complex128, torch.complex64
DEV = "cuda:0"
PSI_TOL = {C64: 1e-5, C128: 1e-12}
GRAD_TOL = {C64: 1e-5, C128: 1e-8}

_REF = {}


def _flat_grads(ops, g):
    g = g.detach().cpu().numpy().reshape(len(ops), 2)
    return np.array([g[i, j] for i, (_, _, p) in enumerate(ops) for j in range(len(p))])


def _rows(ops):
    rows = np.zeros((len(ops), 2))
    for i, (_, _, p) in enumerate(ops):
        rows[i, : len(p)] = p
    return rows


def _c3_ref():
    if not _REF:
        c3 = circuits.config3()
        psi, _ = fo.run_ops(c3["ops"], 6, 8)
        _REF.update(c3=c3, psi=psi, fx=dict(np.load(os.path.join(GOLDEN, "c3_grad.npz"))))
    return _REF


def _report(label, **kw):
    print(label, " ".join("%s=%.3g" % kv for kv in kw.items()), flush=True)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_c3_gradient_fused_evaluator_vs_finite_differences(dtype):
    r = _c3_ref()
    c3, psi_ref, fx = r["c3"], r["psi"], r["fx"]
    ops = c3["ops"]
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], 6, 8, False, dtype, DEV)
    assert cc.fuse == 3 and cc.nsteps < len(ops)  # the benchmarked fused stages
    vac = torch.zeros([8] * 6, dtype=dtype, device=DEV)
    vac.view(-1)[0] = 1
    target = torch.as_tensor(c3["target"], dtype=dtype, device=DEV)
    ev = runtime.GradEvaluator(cc, vac, target, use_graph=True)
    loss, grads = ev(torch.tensor(_rows(ops), dtype=torch.float64, device=DEV))
    torch.cuda.synchronize()
    psi = ev.psi.cpu().numpy().astype(np.complex128)
    dpsi = float(np.abs(psi - psi_ref).max())
    ov = np.vdot(psi_ref.reshape(-1), psi.reshape(-1))
    one_minus_f = 1 - abs(ov) ** 2 / (np.vdot(psi, psi).real * np.vdot(psi_ref, psi_ref).real)
    g = _flat_grads(ops, grads)[fx["idx"]]
    want = fx["fd_h1"] if dtype == C64 else fx["fd"]
    rel = float(np.linalg.norm(g - want) / np.linalg.norm(want))
    _report("c3 fused %s" % dtype, max_dpsi=dpsi, one_minus_F=one_minus_f, grad_rel=rel,
            dloss=abs(float(loss) - float(fx["loss"])))
    assert dpsi <= PSI_TOL[dtype]
    assert abs(float(loss) - float(fx["loss"])) <= PSI_TOL[dtype]
    assert rel <= GRAD_TOL[dtype], rel
    # every R, K, D, S and BS parameter slot of the sample is non-trivial
    assert np.all(np.abs(want) > 1e-7)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_c3_gradient_unfused_autograd_matches_fused(dtype):
    r = _c3_ref()
    c3, fx = r["c3"], r["fx"]
    ops = c3["ops"]
    vac = torch.zeros([8] * 6, dtype=dtype, device=DEV)
    vac.view(-1)[0] = 1
    target = torch.as_tensor(c3["target"], dtype=dtype, device=DEV)
    out = {}
    for fuse in (0, 3):
        cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], 6, 8, False, dtype, DEV, fuse=fuse)
        leaf = torch.tensor(_rows(ops), dtype=torch.float64, device=DEV, requires_grad=True)
        psi = runtime.CircuitFunction.apply(leaf, vac, cc, None)
        runtime.fidelity_loss(psi, target).backward()
        out[fuse] = _flat_grads(ops, leaf.grad)
    g0, g3 = out[0], out[3]
    rel = float(np.linalg.norm(g0 - g3) / np.linalg.norm(g0))
    want = fx["fd_h1"] if dtype == C64 else fx["fd"]
    rel_fd = float(np.linalg.norm(g0[fx["idx"]] - want) / np.linalg.norm(want))
    _report("c3 unfused vs fused %s" % dtype, rel=rel, unfused_vs_fd=rel_fd)
    assert rel <= GRAD_TOL[dtype], rel
    assert rel_fd <= GRAD_TOL[dtype], rel_fd


def test_c3_fixture_differences_are_converged():
    """The two difference steps agree to far below the tolerances they check."""
    fx = _c3_ref()["fx"]
    rel = np.linalg.norm(fx["fd_h1"] - fx["fd"]) / np.linalg.norm(fx["fd"])
    assert rel < 1e-6, rel


def test_c4_batch_on_the_whole_batch_matches_oracle():
    c4 = circuits.config4()["circuits"]
    fx = dict(np.load(os.path.join(GOLDEN, "c4_grad.npz")))
    sel = [int(k) for k in fx["circuits"]]
    tuner = runtime.Autotuner(path="")
    be = runtime.BatchEvaluator(c4, C64, DEV, tuner=tuner)
    plist = [torch.tensor(_rows(c["ops"]), dtype=torch.float64, device=DEV) for c in c4]
    losses, grads = be(plist)
    torch.cuda.synchronize()
    sizes = [c4[k]["D"] ** c4[k]["n"] for k in sel]
    assert min(sizes) == min(c["D"] ** c["n"] for c in c4)
    assert max(sizes) == max(c["D"] ** c["n"] for c in c4) == 2**20
    for k in sel:
        c = c4[k]
        n, D, ops = c["n"], c["D"], c["ops"]
        ref, _ = fo.run_ops(ops, n, D)
        psi = be.items[k]["psi"].cpu().numpy().astype(np.complex128)
        dpsi = float(np.abs(psi - ref).max())
        L = float(fx["c%d_loss" % k])
        dl = abs(float(losses[k]) - L) / max(1.0, abs(L))
        g = _flat_grads(ops, grads[k])[fx["c%d_idx" % k]]
        want = fx["c%d_fd_h1" % k]
        rel = float(np.linalg.norm(g - want) / np.linalg.norm(want))
        _report("c4 circuit %d n=%d D=%d ops=%d fuse=%d" % (k, n, D, len(ops), be.items[k]["cc"].fuse),
                max_dpsi=dpsi, dloss=dl, grad_rel=rel)
        assert dpsi <= PSI_TOL[C64], (k, dpsi)
        assert dl <= PSI_TOL[C64], (k, dl)
        assert rel <= GRAD_TOL[C64], (k, rel)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_c5_full_size_forward_matches_oracle(dtype):
    """The 3-mode cutoff-20 density-matrix workload (6 layers with loss, 144
    operations) forward at full size against the complex128 oracle's
    observables (tests/golden/c5_forward.npz): every Fock probability, the
    trace, each single-mode reduced DM and the low-photon block of the
    two-mode reduced DM of modes (0, 1)."""
    fx = dict(np.load(os.path.join(GOLDEN, "c5_forward.npz")))
    c5 = circuits.config5()
    n, D, ops = c5["n"], c5["D"], c5["ops"]
    cc = runtime.CompiledCircuit([(o[0], o[1]) for o in ops], n, D, True, dtype, DEV)
    rho0 = torch.zeros([D] * (2 * n), dtype=dtype, device=DEV)
    rho0.view(-1)[0] = 1
    rho = runtime.run_circuit(cc, torch.tensor(_rows(ops), dtype=torch.float64, device=DEV), rho0)
    probs = runtime.all_fock_probs(rho, n, D, True).to(torch.float64).cpu().numpy()
    dp = float(np.abs(probs - fx["probs"]).max())
    dtr = abs(float(runtime.norm2(rho, n, D, True)) - float(fx["trace"]))
    drdm = max(
        float(np.abs(runtime.reduced_dm(rho, n, D, True, [m]).cpu().numpy() - fx["rdm%d" % m]).max()) for m in range(n)
    )
    r01 = runtime.reduced_dm(rho, n, D, True, [0, 1]).cpu().numpy()[:10, :10, :10, :10]
    d01 = float(np.abs(r01 - fx["rdm01_low"]).max())
    _report("c5 full size %s" % dtype, max_dprob=dp, dtrace=dtr, max_drdm1=drdm, max_drdm01=d01)
    tol = PSI_TOL[dtype]
    assert dp <= tol and dtr <= tol and drdm <= tol and d01 <= tol, (dp, dtr, drdm, d01)
