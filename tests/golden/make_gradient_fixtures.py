"""Writes the full-size gradient fixtures of tests/golden (run from the repo root).

Both come from the complex128 CPU oracle (oracle/fock_oracle.py, itself pinned
by tests/test_oracle.py) on the seeded workloads of strawberryfields_amd.circuits:

* c3_grad.npz -- the 6-mode cutoff-8 4-layer circuit, loss 1 - |<t|psi>|^2:
  the loss, <t|psi>, four seeded random projections <r_j|psi> of the output
  ket (r_j regenerated from PROJ_SEED by the tests, so the 262,144-amplitude
  ket is checked without storing it), and central differences of the loss
  for 16 parameters (R, K, D, S and BS gates of the first and the last
  layer) at h = 1e-4 and h = 5e-5 with their Richardson extrapolation
  (4 g(h/2) - g(h)) / 3, which removes the O(h^2) term.
* c4_grad.npz -- eight circuits of the 256-circuit batch (the smallest, the
  largest at 2^20 amplitudes and six between, by size rank), loss = total
  mean photon number: per circuit the loss, four random projections of the
  ket and central differences (same two steps, Richardson) for six
  parameters.

Central differences follow the gradient convention of the TF backend
(tfbackend/ops.py:197-202): a real parameter's derivative of a real loss.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fock_oracle as fo  # noqa: E402
from strawberryfields_amd import circuits  # noqa: E402

PROJ_SEED = 777
H1, H2 = 1e-4, 5e-5


def projections(psi, seed=PROJ_SEED, count=4):
    """<r_j|psi> for seeded complex normal vectors r_j (j < count)."""
    rng = np.random.default_rng(seed)
    v = psi.reshape(-1)
    out = []
    for _ in range(count):
        r = rng.normal(size=v.size) + 1j * rng.normal(size=v.size)
        out.append(np.vdot(r, v))
    return np.array(out)


def flat_index(ops, op, slot):
    k = 0
    for i, (_, _, p) in enumerate(ops):
        if i == op:
            assert slot < len(p)
            return k + slot
        k += len(p)
    raise IndexError(op)


def with_flat(ops, flat):
    out, k = [], 0
    for name, modes, p in ops:
        out.append((name, modes, tuple(flat[k : k + len(p)])))
        k += len(p)
    return out


def fd(loss, flat, idx):
    """Central differences at H1 and H2 and the Richardson combination."""
    g1, g2 = np.zeros(len(idx)), np.zeros(len(idx))
    for j, i in enumerate(idx):
        for h, g in ((H1, g1), (H2, g2)):
            a, b = flat.copy(), flat.copy()
            a[i] += h
            b[i] -= h
            g[j] = (loss(a) - loss(b)) / (2 * h)
    return g1, g2, (4 * g2 - g1) / 3


def c3_fixture():
    c3 = circuits.config3()
    ops, n, D, t = c3["ops"], c3["n"], c3["D"], c3["target"]
    flat = np.array([v for _, _, p in ops for v in p], dtype=float)
    per_layer = len(ops) // 4
    # layer offsets: BS 0-14, R 15-20, S 21-26, BS 27-41, R 42-47, D 48-53, K 54-59
    picks = []
    for layer in (0, 3):
        o = layer * per_layer
        # BS theta and phi and an R of the second mesh (the first mesh of the
        # first layer acts on the vacuum, where its gradients vanish), S r and
        # phi, the second mesh's last BS phi, D r, K
        picks += [(o + 27, 0), (o + 27, 1), (o + 42, 0), (o + 22, 0), (o + 22, 1),
                  (o + 40, 1), (o + 50, 0), (o + 57, 0)]
    idx = np.array([flat_index(ops, o, s) for o, s in picks])
    kinds = [ops[o][0] for o, _ in picks]

    def loss(x):
        st, _ = fo.run_ops(with_flat(ops, x), n, D)
        return 1 - abs(np.vdot(t.reshape(-1), st.reshape(-1))) ** 2

    psi, _ = fo.run_ops(ops, n, D)
    ov = np.vdot(t.reshape(-1), psi.reshape(-1))
    g1, g2, gr = fd(loss, flat, idx)
    return dict(idx=idx, kinds=np.array(kinds), loss=1 - abs(ov) ** 2, overlap=ov, proj=projections(psi),
                norm=np.vdot(psi, psi).real, fd_h1=g1, fd_h2=g2, fd=gr)


def c4_fixture():
    c4 = circuits.config4()["circuits"]
    order = sorted(range(len(c4)), key=lambda i: (c4[i]["D"] ** c4[i]["n"], len(c4[i]["ops"]), i))
    ranks = [0, 40, 90, 140, 180, 215, 240, len(order) - 1]
    sel = [order[r] for r in ranks]
    rng = np.random.default_rng(4242)
    out = dict(circuits=np.array(sel))
    for k in sel:
        c = c4[k]
        ops, n, D = c["ops"], c["n"], c["D"]
        flat = np.array([v for _, _, p in ops for v in p], dtype=float)
        # parameters of the gates from the first squeezer on (the gates before
        # it act on the vacuum, where the gradient vanishes)
        first = flat_index(ops, next(i for i, o in enumerate(ops) if o[0] == "S"), 0)
        idx = np.sort(rng.choice(np.arange(first, len(flat)), size=6, replace=False))

        def loss(x):
            st, pure = fo.run_ops(with_flat(ops, x), n, D)
            return sum(fo.mean_photon(st, n, pure, m)[0] for m in range(n))

        psi, pure = fo.run_ops(ops, n, D)
        g1, g2, gr = fd(loss, flat, idx)
        out["c%d_idx" % k] = idx
        out["c%d_loss" % k] = np.array(loss(flat))
        out["c%d_proj" % k] = projections(psi)
        out["c%d_fd_h1" % k] = g1
        out["c%d_fd" % k] = gr
        print("c4 circuit %d (n=%d, D=%d, %d ops) done" % (k, n, D, len(ops)), flush=True)
    return out


def main(which=("c3", "c4")):
    for name, make in (("c3", c3_fixture), ("c4", c4_fixture)):
        if name in which:
            t0 = time.time()
            np.savez_compressed(os.path.join(HERE, "%s_grad.npz" % name), **make())
            print("%s_grad.npz (%.0f s)" % (name, time.time() - t0), flush=True)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("c3", "c4"))
