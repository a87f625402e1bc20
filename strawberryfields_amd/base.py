"""Backend base classes: Strawberry Fields' own when installed, else a mirror.

The drop-in boundary is the reference plugin API
``strawberryfields/backends/base.py`` (``NotApplicableError`` 20-24, ``ModeMap``
27-114, ``BaseBackend`` 117-472, ``BaseFock`` 480-621): ``sf.Engine`` accepts any
``BaseBackend`` instance (``strawberryfields/engine.py:71-78``).  When the real
package is importable the backend subclasses its classes so the real Engine
accepts it; otherwise this module provides classes with the same names,
attributes and behaviour for the mirror front end in ``frontend.py``.
"""

try:  # pragma: no cover - strawberryfields is not installed in this image
    from strawberryfields.backends.base import BaseBackend, BaseFock, ModeMap, NotApplicableError

    HAVE_SF = True
except Exception:  # ImportError and broken optional dependencies alike
    HAVE_SF = False

    class NotApplicableError(TypeError):
        """The backend cannot perform this operation (base.py:20-24)."""

    class ModeMap:
        """Map from user mode indices to backend indices (base.py:27-114)."""

        def __init__(self, num_subsystems):
            self._init = num_subsystems
            self._map = list(range(num_subsystems))

        def reset(self):
            self._map = list(range(self._init))

        def _single_mode_valid(self, mode):
            if mode is None:
                return False
            return 0 <= mode < len(self._map)

        def _reduce_to_existing_modes(self, modes):
            if isinstance(modes, int):
                modes = [modes]
            return [m for m in modes if m in self._map]

        def remap(self, modes):
            if isinstance(modes, int):
                return self._map[modes]
            return [self._map[m] for m in modes]

        def valid(self, modes):
            if modes is None:
                return False
            if isinstance(modes, int):
                modes = [modes]
            if len(modes) == 0 or len(modes) > len(self._map):
                return False
            return all(self._single_mode_valid(m) for m in modes)

        def show(self):
            return self._map

        def delete(self, modes):
            if isinstance(modes, int):
                modes = [modes]
            if not self.valid(modes):
                raise ValueError("Specified modes for deleting are invalid.")
            new_map, ctr = [], 0
            for m, _ in enumerate(self._map):
                if m in modes or self._map[m] is None:
                    new_map.append(None)
                else:
                    new_map.append(ctr)
                    ctr += 1
            self._map = new_map

        def add(self, num_modes):
            active = len([m for m in self._map if m is not None])
            self._map += list(range(active, active + num_modes))

    class BaseBackend:
        """Abstract backend (base.py:117-472)."""

        short_name = "base"
        compiler = None

        def __init__(self):
            self._supported = {}

        def __str__(self):
            return self.__class__.__name__

        def supports(self, name):
            return self._supported.get(name, False)

        def _unsupported(self, *args, **kwargs):
            raise NotImplementedError

        begin_circuit = add_mode = del_mode = get_modes = reset = _unsupported
        prepare_vacuum_state = prepare_coherent_state = prepare_squeezed_state = _unsupported
        prepare_displaced_squeezed_state = prepare_thermal_state = _unsupported
        rotation = displacement = squeeze = beamsplitter = mzgate = _unsupported
        loss = thermal_loss = measure_homodyne = measure_fock = measure_threshold = _unsupported
        is_vacuum = state = _unsupported

    class BaseFock(BaseBackend):
        """Fock-basis backend (base.py:480-621)."""

        compiler = "fock"

        def __init__(self):
            super().__init__()
            self._supported["fock_basis"] = True

        get_cutoff_dim = prepare_fock_state = prepare_ket_state = BaseBackend._unsupported
        prepare_dm_state = cubic_phase = kerr_interaction = BaseBackend._unsupported
        cross_kerr_interaction = two_mode_squeeze = BaseBackend._unsupported
