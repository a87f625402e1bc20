"""Writes tests/golden/c5_forward.npz (run from the repo root; ~20 min, ~6 GB).

The 3-mode cutoff-20 density-matrix workload of
strawberryfields_amd.circuits.config5 (6 layers, a loss channel per mode in
each layer, 144 operations), run forward by the complex128 CPU oracle
(oracle/fock_oracle.py, pinned by tests/test_oracle.py) from the vacuum.  The
final state has 20^6 = 64e6 entries (1 GB), so the fixture keeps observables
that fix it tightly: all Fock probabilities (8000), the trace, every
single-mode reduced density matrix and the block of photon numbers below 10
of the two-mode reduced density matrix of modes (0, 1).

Loss channels are applied as the oracle's Kraus sum (fockbackend/circuit.py:
65-87, ops.py:471-490) with the terms accumulated one at a time, so the
working set stays a few copies of the state.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fock_oracle as fo  # noqa: E402
from strawberryfields_amd import circuits  # noqa: E402


def main():
    c5 = circuits.config5()
    n, D, ops = c5["n"], c5["D"], c5["ops"]
    c = fo.Circuit(n, D, pure=False)
    t0 = time.time()
    for i, (name, modes, p) in enumerate(ops):
        if name == "LOSS":
            acc = None
            for K in fo.loss_kraus(p[0], D):
                term = c.apply_gate_BLAS(K, [modes[0]], pure=False)
                acc = term if acc is None else acc + term
                del term
            c._state = acc
        elif name == "BS":
            c.beamsplitter(p[0], p[1], modes[0], modes[1])
        elif name == "R":
            c.phase_shift(p[0], modes[0])
        elif name == "K":
            c.kerr_interaction(p[0], modes[0])
        elif name == "D":
            c.displacement(p[0], p[1], modes[0])
        elif name == "S":
            c.squeeze(p[0], p[1], modes[0])
        else:
            raise ValueError(name)
        print("op %d/%d %s (%.0f s)" % (i + 1, len(ops), name, time.time() - t0), flush=True)
    rho = c._state
    out = dict(
        probs=fo.all_fock_probs(rho, n, False),
        trace=np.array(fo.trace(rho, n, False)),
        rdm0=fo.reduced_dm(rho, n, False, [0]),
        rdm1=fo.reduced_dm(rho, n, False, [1]),
        rdm2=fo.reduced_dm(rho, n, False, [2]),
        rdm01_low=np.ascontiguousarray(fo.reduced_dm(rho, n, False, [0, 1])[:10, :10, :10, :10]),
    )
    np.savez_compressed(os.path.join(HERE, "c5_forward.npz"), **out)
    print("wrote c5_forward.npz (%.0f s)" % (time.time() - t0))


if __name__ == "__main__":
    main()
