"""MI355X-native ERA / difference-map ptychography.

Drop-in for the ``Ptychography.ERA`` / ``Ptychography.DM`` call surface of
andyofmelbourne/Ptychography (``__init__.py:3-4`` there), computed by
hand-written gfx950 HIP kernels in ``libptyhip.so`` through a ctypes C ABI.
"""

from .era import ERA, update_progress
from .dm import DM
from .batch import ERA_many, DM_many
from .back_projection import Back_projection
from .hio import HIO
from .cxi import read_cxi, write_cxi
from . import sim

__all__ = ["ERA", "DM", "update_progress", "ERA_many", "DM_many", "Back_projection", "HIO", "sim", "write_cxi", "read_cxi"]
