"""numpy views of the C-ABI structs declared in include/obvh/obvh.h.

Layouts follow the reference types (paths relative to the reference repo):
``kr_vec3`` 12 B, ``kr_uvec4`` 16 B, ``kr_obb3`` 60 B (mid, v0, v1, v2, ext),
``kr_mat4`` 64 B column major (src/common/vecmath.h:136,156-158,172),
``kr_bvh_node`` 36 B (src/common/ads.h:131-142), ``kr_ray`` 32 B and
``kr_intersection`` 28 B (src/common/ads.h:64-76), ``kr_mat43`` 48 B
(vecmath.h:159) and ``kr_ivec4`` 16 B (vecmath.h:188).
"""
from __future__ import annotations

import numpy as np

INVALID = 0xFFFFFFFF

NODE_DTYPE = np.dtype(
    [
        ("bmin", "<f4", (3,)),
        ("bmax", "<f4", (3,)),
        ("a", "<u4"),  # leaf: face id (primitivesOffset); internal: left child
        ("b", "<u4"),  # internal: right child
        ("nprim", "<u2"),
        ("axis", "u1"),
        ("pad", "u1"),
    ]
)
assert NODE_DTYPE.itemsize == 36

TRACE_DTYPE = np.dtype(
    [
        ("slab_min_idx", "<u4", (7,)),
        ("slab_max_idx", "<u4", (7,)),
        ("pair_k", "<i4"),
        ("p2_sel", "<i4"),
        ("q0_sel", "<i4"),
        ("q1_sel", "<i4"),
        ("q0_valid", "<i4"),
        ("q1_valid", "<i4"),
        ("degenerate", "<i4"),
        ("best", "<i4"),
        ("result_aabb", "<i4"),
        ("pad", "<i4"),
    ]
)
assert TRACE_DTYPE.itemsize == 96

RAY_DTYPE = np.dtype([("origin", "<f4", (3,)), ("tmin", "<f4"), ("direction", "<f4", (3,)), ("tmax", "<f4")])
assert RAY_DTYPE.itemsize == 32

ISECT_DTYPE = np.dtype([("normal", "<f4", (3,)), ("barys", "<f4", (2,)), ("primitive", "<u4"), ("instance", "<u4")])
assert ISECT_DTYPE.itemsize == 28

MISS = 0xFFFFFFFF          # kr_intersection.primitive of a ray that hit nothing
STACK_OVERFLOW = 0xFFFFFFFF  # kr_intersection.instance of a ray whose traversal stack overflowed

OBB_FLOATS = 15  # mid, v0, v1, v2, ext
MAT_FLOATS = 16


def obb_fields(obbs: np.ndarray) -> dict[str, np.ndarray]:
    """Split an (n, 15) float32 OBB array into named (n, 3) views."""
    o = np.asarray(obbs).reshape(-1, OBB_FLOATS)
    return {"mid": o[:, 0:3], "v0": o[:, 3:6], "v1": o[:, 6:9], "v2": o[:, 9:12], "ext": o[:, 12:15]}


def empty_nodes(n: int) -> np.ndarray:
    return np.zeros(n, dtype=NODE_DTYPE)
