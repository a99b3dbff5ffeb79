"""MI355X-native Walk-on-Spheres engine (host orchestration over a HIP C ABI).

The compute path is the HIP library lib/libn2wos_amd.so built for gfx950; this
package only marshals arguments through ctypes.  Importing it does not touch
the GPU; the first call that needs the library loads it and raises if it is
missing.
"""
from . import cache, cv, nc, problems, serve, wos  # noqa: F401
from ._lib import LIB_PATH, N2wosError, check, header_functions, lib  # noqa: F401
from .geometry import (  # noqa: F401
    SPHERE_MULTI, SPHERE_REFERENCE_BUMP, SPHERE_SH_BLOB, SPHERE_SIN6, Bvh, build_host, count_degenerate_triangles,
    normalize_to_unit_radius, uv_sphere)
