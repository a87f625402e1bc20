"""Builds the oracle's native part (oracle/fock_cpu.c -> oracle/libfock_cpu.so).

The library restates the reference fock backend's numba loops (see the file
header) for the faithful CPU baseline; it is test infrastructure, loaded only
by oracle/fock_oracle.py.  gcc -O3, no other dependency.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fock_cpu.c")
LIB = os.path.join(HERE, "libfock_cpu.so")


def build(verbose=False):
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fPIC", "-shared", "-o", LIB + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
