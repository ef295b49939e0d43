"""MI355X-native MLS-MPM step engine (host side).

The compute path is the HIP library ``lib/libmpm_amd.so`` (C ABI in
include/mpm_engine.h); this package binds it with ctypes and provides the
scene seeding used by the tests and the benchmark.
"""
from .engine import Batch, Engine, Material, sort_pairs  # noqa: F401
