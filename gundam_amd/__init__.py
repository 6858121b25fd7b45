"""gundam_amd -- MI355X-native event-reweighting cache for GUNDAM's
Propagator reweight-and-fill path (the reference's Cache::Manager), behind the
C ABI in include/gcm.h.

    from gundam_amd import CacheManager, Workload
"""
from ._native import GcmError, device_count, lib  # noqa: F401
from .cache_manager import CacheManager, Future  # noqa: F401
from .tables import INPUT_DTYPE, Tables, Workload  # noqa: F401

__all__ = ["CacheManager", "Future", "GcmError", "INPUT_DTYPE", "Tables",
           "Workload", "device_count", "lib"]
