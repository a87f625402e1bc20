"""CPU oracle for the HIP Fock-basis library.

TEST INFRASTRUCTURE ONLY: a numpy complex128 restatement of the reference
``fock`` backend used as the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it;
the product package ``strawberryfields_amd`` never does.
"""
