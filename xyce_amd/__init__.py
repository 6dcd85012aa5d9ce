"""MI355X (gfx950) device-load path of the Xyce circuit simulator.

The compute lives in libxg_load.so (HIP kernels behind the C ABI of
include/xg_load.h); this package holds the Python binding (``_lib``), the
Master-style host wrapper (``loader``), circuit descriptions and seeded
benchmark netlists (``circuits``) and the build script (``build``).
"""
