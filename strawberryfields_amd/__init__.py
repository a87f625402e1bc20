"""MI355X-native Fock-basis backend for Strawberry Fields.

``HIPFockBackend`` is a ``BaseFock`` plugin whose gate application, adjoint and
readout run in hand-written HIP kernels for gfx950 (``csrc/``), loaded through
the C ABI of ``include/sfamd.h``.  ``frontend`` mirrors the parts of the
Strawberry Fields front end (Program, Engine, gates) that drive it.
"""
from .backend import HIPFockBackend
from .states import FockStateHIP
from . import frontend, circuits, runtime

__all__ = ["HIPFockBackend", "FockStateHIP", "frontend", "circuits", "runtime"]
