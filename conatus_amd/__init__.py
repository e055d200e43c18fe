"""conatus_amd — MI355X-native IMEX hot path of Conatus (batched bordered-banded solves first).

Host code is Python; the compute runs in hand-written gfx950 HIP kernels of ``libcnhip.so``
(C ABI in ``include/cnhip.h``), with torch tensors as device buffers.
"""
__version__ = "0.1.0"

from . import errors  # noqa: F401
