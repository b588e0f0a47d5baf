"""MI355X engine for PyRDSim's time-step loop.

``pyrdsim_amd.system`` is the drop-in for ``pyrdsim.system``; ``Engine`` is the
lower-level handle on the HIP library (one or many replicas per launch).
"""

from .system import system  # noqa: F401


def __getattr__(name):
    if name == 'Engine':
        from .engine import Engine
        return Engine
    raise AttributeError(name)
