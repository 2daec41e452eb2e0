"""LBM3 hot path on MI355X: D3Q27 cumulant collision with esotwist streaming,
interpolated bounceback, moment-based open boundaries and the multilevel
interface, as HIP kernels for gfx950 behind the C ABI of include/lbm3mi.h."""
from . import ctypes_defs, cases, geometry  # noqa: F401
