"""scaletrack -- MI355X (gfx950) Lagrangian parcel step behind a C-ABI.

Importing this package loads libscaletrack.so and fails loudly if it is not built.
"""
from . import _lib  # noqa: F401  (loads the library)
from .tracking import (  # noqa: F401
    CHANNEL, FP32, FP64, FLUSH_ATOMIC, FLUSH_OWNER, PERIODIC, PLAIN, SORTED, WALL, Chunk, Context, Graph,
    Batch, Mesh, Time, TwoWayEulerian, construct_mesh, device_count, evolve, evolve_device,
    autotune_batch, increment_time, params, pinned_copy, pinned_empty)
from ._lib import ScaleTrackError  # noqa: F401

__version__ = "0.1.0"
