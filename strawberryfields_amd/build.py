"""Build recipe for libsfamd.so (hand-written HIP for gfx950).

Each ``csrc/*.hip`` file is compiled by ``hipcc --offload-arch=gfx950`` into an
object under ``_build/`` and the objects are linked into
``strawberryfields_amd/libsfamd.so``, a plain C-ABI shared library (see
``include/sfamd.h``) loaded with ctypes.
"""
import concurrent.futures
import os
import re
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD_DIR = os.path.join(REPO_DIR, "_build")
LIB_PATH = os.path.join(PKG_DIR, "libsfamd.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: the device code is built without packed FP32 instructions
# (v_pk_*_f32); the host pass warns that it does not know the feature.
# -dot7-insts: without the byte dot-product instructions.  With them, the
# gfx950 backend rewrote an index of the form e - k*s0 + k*s1 (k, s0, s1 known
# to fit in a byte) into v_perm_b32 + v_dot4_u32_u8, which adds k*s0 instead
# of subtracting it: the fused beamsplitter adjoint at cutoffs 12-15 read the
# wrong cotangent amplitudes.  No kernel here uses dot-product instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-dot7-insts",
    "-I" + os.path.join(REPO_DIR, "include"),
]


def _compile(src, obj, verbose):
    cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass reports the device-only target features as unknown; harmless
    err = "\n".join(
        l
        for l in res.stderr.splitlines()
        if "packed-fp32-ops" not in l and "dot7-insts" not in l and l.strip()
    )
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, err))
    if err and verbose:
        print(err, file=sys.stderr)
    return obj


_INCLUDE = re.compile(r'^\s*#\s*include\s+"([^"]+)"', re.M)


def _deps(path, seen=None):
    """The file and every header it includes with quotes, transitively."""
    seen = set() if seen is None else seen
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for inc in _INCLUDE.findall(f.read()):
            _deps(os.path.join(os.path.dirname(path), inc), seen)
    return seen


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(f) > t for f in _deps(src))


def build(verbose=False, jobs=8, changed_only=False):
    """Compile every HIP source and link libsfamd.so; returns the library path.

    changed_only: recompile only sources older than their object or than a
    header they include; the default rebuilds everything."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(BUILD_DIR, f[:-4] + ".o") for f in srcs]
    todo = [
        (os.path.join(CSRC, s), o)
        for s, o in zip(srcs, objs)
        if not changed_only or _stale(os.path.join(CSRC, s), o)
    ]
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o, verbose) for s, o in todo]
        for f in futs:
            f.result()
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True, changed_only="--changed" in sys.argv))
