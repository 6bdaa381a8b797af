"""ttx: MI355X-native TT-SVD / TR-SVD engine (hand-written gfx950 HIP kernels behind a C ABI).

``build.build()`` compiles the in-tree ``libttx.so``; ``api.Engine`` is the Python mirror of the
reference's decomposition entry points.
"""
from . import _lib  # noqa: F401
