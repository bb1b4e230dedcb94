"""Seeded synthetic inputs for every workload of the OBB path.

The reference ships no meshes (its glTF buffers are absent), so all inputs are
generated here, deterministically from a seed (numpy PCG64):

* ``gaussian_points``   anisotropic Gaussian cloud, random rotation, shift
                        (single DiTO fit, 131 072 points by default).
* ``icosphere``         geodesic icosphere of frequency f: 20 f^2 triangles,
                        10 f^2 + 2 shared vertices, radius 1 + U(-0.05, 0.05)
                        per vertex (f = 224 gives 1 003 520 triangles).
* ``hairball``          thin open tubes (20 radial segments, 40 triangles and
                        40 vertices each) inside the unit cube.
* ``batch_sets``        many Gaussian sets with log-uniform sizes.
* ``camera_rays``       pinhole camera on a ring around the scene box,
                        one ray per pixel, row-major (coherent queries).
* ``random_rays``       origins on a sphere around the box aimed at uniform
                        points inside it (incoherent queries).

Everything is computed in float64 and rounded once to float32.
"""
from __future__ import annotations

import math

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(np.random.PCG64(seed))


def random_rotation(rng: np.random.Generator, count: int | None = None) -> np.ndarray:
    """Uniform rotation matrices from normalised 4-D Gaussian quaternions."""
    shape = (1 if count is None else count, 4)
    q = rng.standard_normal(shape)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R[0] if count is None else R


def gaussian_points(n: int = 131072, std=(4.0, 1.0, 0.25), seed: int = 1, shift: bool = True) -> np.ndarray:
    """(n, 3) float32: N(0, diag(std)^2), rotated, translated by U(-10, 10)^3."""
    rng = _rng(seed)
    p = rng.standard_normal((n, 3)) * np.asarray(std, dtype=np.float64)
    R = random_rotation(rng)
    p = p @ R.T
    if shift:
        p += rng.uniform(-10.0, 10.0, size=3)
    return np.ascontiguousarray(p.astype(np.float32))


_ICO_V = None
_ICO_F = None


def _icosahedron():
    global _ICO_V, _ICO_F
    if _ICO_V is None:
        t = (1.0 + math.sqrt(5.0)) / 2.0
        v = np.array(
            [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
             [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
             [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        f = np.array(
            [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
             [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
             [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
             [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
        _ICO_V, _ICO_F = v, f
    return _ICO_V, _ICO_F


def icosphere(f: int = 224, noise: float = 0.05, seed: int = 2):
    """Geodesic icosphere of frequency ``f``.

    Returns (vertices (V, 3) float32, faces (F, 4) uint32 with w = 0).
    V = 10 f^2 + 2, F = 20 f^2.  Each vertex is pushed to radius
    1 + U(-noise, noise).
    """
    if f < 1:
        raise ValueError("frequency must be >= 1")
    cv, cf = _icosahedron()
    # barycentric grid (i, j) with i + j <= f; weights k = f-i-j on A, i on B, j on C
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    m = (ii + jj) <= f
    gi, gj = ii[m], jj[m]
    gk = f - gi - gj
    G = gi.size
    keys = np.empty((20, G), dtype=np.int64)
    pts = np.empty((20, G, 3))
    big = (f + 1)
    for fi in range(20):
        A, B, C = cf[fi]
        w = {A: gk, B: gi, C: gj}
        pts[fi] = (gk[:, None] * cv[A] + gi[:, None] * cv[B] + gj[:, None] * cv[C]) / f
        key = np.empty(G, dtype=np.int64)
        zero = (gk == 0).astype(int) + (gi == 0).astype(int) + (gj == 0).astype(int)
        # corners: two zero weights
        cor = zero == 2
        corner_v = np.where(gk == f, A, np.where(gi == f, B, C))
        key[cor] = corner_v[cor]
        # edges: exactly one zero weight; key by (min corner, max corner, weight of min corner)
        edg = zero == 1
        for (p, q) in ((A, B), (B, C), (A, C)):
            r = ({A, B, C} - {p, q}).pop()
            sel = edg & (w[r] == 0)
            lo, hi = (p, q) if p < q else (q, p)
            key[sel] = 100 + (lo * 12 + hi) * big + w[lo][sel]
        inter = zero == 0
        key[inter] = 100 + 144 * big + fi * big * big + gi[inter] * big + gj[inter]
        keys[fi] = key
    ukeys, inv = np.unique(keys.ravel(), return_inverse=True)
    inv = inv.reshape(20, G)
    V = np.zeros((ukeys.size, 3))
    V[inv.ravel()] = pts.reshape(-1, 3)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    rng = _rng(seed)
    V *= (1.0 + rng.uniform(-noise, noise, size=V.shape[0]))[:, None]
    # local grid index of (i, j)
    lut = -np.ones((f + 1, f + 1), dtype=np.int64)
    lut[gi, gj] = np.arange(G)
    ui, uj = np.meshgrid(np.arange(f), np.arange(f), indexing="ij")
    up = (ui + uj) <= f - 1
    di, dj = ui[(ui + uj) <= f - 2], uj[(ui + uj) <= f - 2]
    ui, uj = ui[up], uj[up]
    tri_local = np.concatenate(
        [
            np.stack([lut[ui, uj], lut[ui + 1, uj], lut[ui, uj + 1]], axis=1),
            np.stack([lut[di + 1, dj], lut[di + 1, dj + 1], lut[di, dj + 1]], axis=1),
        ]
    )
    faces = np.concatenate([inv[fi][tri_local] for fi in range(20)])
    F = np.zeros((faces.shape[0], 4), dtype=np.uint32)
    F[:, :3] = faces
    return np.ascontiguousarray(V.astype(np.float32)), F


def hairball(tubes: int = 50000, radius: float = 0.002, segments: int = 20, seed: int = 3,
             length=(0.05, 0.3)):
    """Open thin tubes with random direction and length inside [0, 1]^3.

    Returns (vertices (40 * tubes, 3) float32, faces (40 * tubes, 4) uint32)
    for the default 20 segments (2 * segments triangles and vertices per tube).
    """
    rng = _rng(seed)
    L = rng.uniform(length[0], length[1], size=tubes)
    d = rng.standard_normal((tubes, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    half = 0.5 * L[:, None] * np.abs(d) + radius
    c = half + rng.uniform(0.0, 1.0, size=(tubes, 3)) * (1.0 - 2.0 * half)
    # orthonormal frame (u, w) perpendicular to d
    a = np.where(np.abs(d[:, 0:1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    u = np.cross(d, a)
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = np.cross(d, u)
    th = 2 * np.pi * np.arange(segments) / segments
    ring = np.cos(th)[None, :, None] * u[:, None, :] + np.sin(th)[None, :, None] * w[:, None, :]
    bot = c[:, None, :] - 0.5 * L[:, None, None] * d[:, None, :] + radius * ring
    top = c[:, None, :] + 0.5 * L[:, None, None] * d[:, None, :] + radius * ring
    V = np.concatenate([bot, top], axis=1).reshape(-1, 3)
    k = np.arange(segments)
    k1 = (k + 1) % segments
    local = np.concatenate(
        [np.stack([k, k1, segments + k], axis=1), np.stack([k1, segments + k1, segments + k], axis=1)]
    )
    base = (2 * segments * np.arange(tubes))[:, None, None]
    faces = (local[None, :, :] + base).reshape(-1, 3)
    F = np.zeros((faces.shape[0], 4), dtype=np.uint32)
    F[:, :3] = faces
    return np.ascontiguousarray(V.astype(np.float32)), F


def batch_sizes(nsets: int, lo: int = 64, hi: int = 1 << 20, seed: int = 4) -> np.ndarray:
    """Set sizes floor(exp(U(ln lo, ln hi))) (log-uniform), int64."""
    rng = _rng(seed)
    s = np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), size=nsets))).astype(np.int64)
    return np.clip(s, lo, hi)


def batch_sets(sizes, seed: int = 5):
    """Concatenated (sum(sizes), 3) float32 sets and (nsets + 1) uint64 offsets.

    Set s: N(0, diag(1, r1, r2)^2) with r1, r2 ~ U(1, 20) sorted descending,
    then a random rotation.
    """
    sizes = np.asarray(sizes, dtype=np.int64)
    rng = _rng(seed)
    off = np.zeros(sizes.size + 1, dtype=np.uint64)
    off[1:] = np.cumsum(sizes)
    out = np.empty((int(off[-1]), 3), dtype=np.float32)
    ratios = rng.uniform(1.0, 20.0, size=(sizes.size, 2))
    R = random_rotation(rng, sizes.size)
    for s, n in enumerate(sizes):
        std = np.sort(np.array([1.0, ratios[s, 0], ratios[s, 1]]))[::-1]
        p = rng.standard_normal((int(n), 3)) * std
        out[int(off[s]):int(off[s + 1])] = (p @ R[s].T).astype(np.float32)
    return out, off


RAY_TMAX = 1.0e30


def _rays(origins: np.ndarray, directions: np.ndarray) -> np.ndarray:
    from .layouts import RAY_DTYPE

    r = np.zeros(origins.shape[0], dtype=RAY_DTYPE)
    r["origin"] = origins.astype(np.float32)
    r["direction"] = directions.astype(np.float32)
    r["tmin"] = 0.0
    r["tmax"] = RAY_TMAX
    return r


def camera_rays(vertices: np.ndarray, width: int = 1024, height: int = 1024, angle: float = 0.6,
                fov_deg: float = 40.0, distance: float = 1.8) -> np.ndarray:
    """Pinhole camera at `distance` box radii from the box centre, on a ring
    at azimuth `angle` (radians) and 20 degrees elevation, looking at the
    centre; unit directions through pixel centres, row-major."""
    v = np.asarray(vertices, dtype=np.float64)
    lo, hi = v.min(0), v.max(0)
    c = 0.5 * (lo + hi)
    rad = 0.5 * float(np.linalg.norm(hi - lo))
    el = math.radians(20.0)
    eye = c + distance * rad * np.array([math.cos(el) * math.cos(angle), math.sin(el), math.cos(el) * math.sin(angle)])
    fwd = (c - eye) / np.linalg.norm(c - eye)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = math.tan(math.radians(fov_deg) / 2)
    ys, xs = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    px = ((xs + 0.5) / width * 2 - 1) * th * (width / height)
    py = (1 - (ys + 0.5) / height * 2) * th
    d = fwd[None, :] + px.reshape(-1, 1) * right[None, :] + py.reshape(-1, 1) * up[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _rays(np.broadcast_to(eye, d.shape), d)


def random_rays(vertices: np.ndarray, n: int, seed: int = 11) -> np.ndarray:
    """Origins uniform on the sphere of 1.5 box radii around the box centre,
    unit directions towards points uniform in the box."""
    v = np.asarray(vertices, dtype=np.float64)
    lo, hi = v.min(0), v.max(0)
    c = 0.5 * (lo + hi)
    rad = 0.5 * float(np.linalg.norm(hi - lo))
    rng = _rng(seed)
    u = rng.standard_normal((n, 3))
    o = c + 1.5 * rad * u / np.linalg.norm(u, axis=1, keepdims=True)
    t = lo + rng.uniform(size=(n, 3)) * (hi - lo)
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _rays(o, d)
