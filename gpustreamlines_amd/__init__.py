"""MI355X-native per-seed tracking hot path of dipy/GPUStreamlines.

Drop-in names of the reference package ``cuslines`` (cuslines/__init__.py:82-88):
``Tracker``/``GPUTracker``, ``BootDirectionGetter``, ``ProbDirectionGetter``,
``PttDirectionGetter`` and ``BACKEND``.  There is one backend, HIP on gfx950,
reached through the C ABI of ``libcuslines_hip.so``; no dispatch and no CPU
fallback.
"""

from .direction_getters import (
    BootDirectionGetter,
    GPUDirectionGetter,
    ProbDirectionGetter,
    PttDirectionGetter,
)
from .tracker import GPUTracker

Tracker = GPUTracker
BACKEND = "hip"

__all__ = [
    "Tracker",
    "GPUTracker",
    "GPUDirectionGetter",
    "ProbDirectionGetter",
    "PttDirectionGetter",
    "BootDirectionGetter",
    "BACKEND",
]
