"""Pins of the CPU DiTO-14 oracle (oracle/dito.c) by closed forms, brute force
and invariants.  The reference CPU sources cannot be compiled here (its
common/vecmath.h includes the MSVC header <intrin.h>), so nothing here
compares against the reference binary; each test states what fixes the
expected answer independently of the oracle's code.
"""
import numpy as np
import pytest

from obvh_amd import synth


def obb_parts(o):
    o = np.asarray(o, dtype=np.float64).reshape(5, 3)
    return o[0], o[1:4], o[4]  # mid, axes rows, ext


def project_all(points, axes):
    return np.asarray(points, np.float64) @ np.asarray(axes, np.float64).T


def test_empty_set_is_zero_box(oracle_lib):
    # dito.cpp:498-502: zero box, identity axes; kr_points_obb floors ext.
    o = oracle_lib.dito14(np.zeros((0, 3), np.float32))
    assert np.array_equal(o, np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32))
    o = oracle_lib.points_obb(np.zeros((0, 3), np.float32))
    assert np.array_equal(o[12:], np.full(3, 0.001, np.float32))


@pytest.mark.parametrize("n", [1, 5, 100])
def test_coincident_points_give_the_point(oracle_lib, n):
    # All points equal: the furthest pair is degenerate -> AABB of one point.
    p = np.tile(np.array([[1.5, -2.0, 3.25]], np.float32), (n, 1))
    o, tr = oracle_lib.dito14(p, trace=True)
    mid, ax, ext = obb_parts(o)
    assert tr["degenerate"] == 1 and tr["result_aabb"] == 1
    assert np.array_equal(mid, p[0]) and np.array_equal(ext, np.zeros(3))
    assert np.array_equal(ax, np.eye(3))


def test_collinear_points_give_line_aligned_box(oracle_lib):
    # dito.cpp:451-476: a line-aligned box whose first axis is the line.
    rng = np.random.default_rng(7)
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    t = rng.uniform(-4, 4, size=64)
    p = (np.array([1.0, 2.0, -1.0]) + t[:, None] * d).astype(np.float32)
    o, tr = oracle_lib.dito14(p, trace=True)
    mid, ax, ext = obb_parts(o)
    assert tr["degenerate"] == 2
    assert abs(abs(ax[0] @ d) - 1) < 1e-6
    assert abs(ext[0] - (t.max() - t.min()) / 2) < 1e-5 * (t.max() - t.min())
    assert np.all(ext[1:] < 1e-5)


def test_rotated_box_corners_recover_the_box(oracle_lib):
    # 8 corners of a box with distinct half extents under a random rotation:
    # the true box is the unique minimal OBB; DiTO must return it (axes up to
    # sign and order, extents exactly up to float rounding).
    rng = np.random.default_rng(11)
    for _ in range(50):
        R = synth.random_rotation(rng)
        half = np.sort(rng.uniform(0.5, 3.0, size=3))[::-1]
        c = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * half
        shift = rng.uniform(-5, 5, 3)
        pts = (c @ R.T + shift).astype(np.float32)
        o, tr = oracle_lib.dito14(pts, trace=True)
        mid, ax, ext = obb_parts(o)
        order = np.argsort(-ext)
        assert np.allclose(ext[order], half, rtol=2e-5)
        for j, k in enumerate(order):
            assert abs(abs(ax[k] @ R[:, j]) - 1) < 1e-5
        assert np.allclose(mid, shift, atol=2e-5 * 8)


def test_slab_extremes_are_first_argmin_argmax(oracle_lib):
    # Brute force with numpy's first-occurrence argmin/argmax over the 7 slab
    # projections (float32 arithmetic, dito.cpp:146-171), with planted ties.
    rng = np.random.default_rng(3)
    p = rng.integers(-20, 20, size=(4000, 3)).astype(np.float32) * 0.5  # many exact ties
    o, tr = oracle_lib.dito14(p, trace=True)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    pr = np.stack([x, y, z, (x + y) + z, (x + y) - z, (x - y) + z, (x - y) - z], axis=1)
    assert pr.dtype == np.float32
    assert np.array_equal(tr["slab_min_idx"], pr.argmin(axis=0))
    assert np.array_equal(tr["slab_max_idx"], pr.argmax(axis=0))


@pytest.mark.parametrize("seed", range(6))
def test_box_encloses_points_and_beats_aabb(oracle_lib, seed):
    # Enclosure (FP64) and: the result is never worse than the AABB, and its
    # extents are the exact projection ranges of ALL points on its axes.
    p = synth.gaussian_points(20000, std=(3.0, 1.0, 0.3), seed=seed)
    o, tr = oracle_lib.dito14(p, trace=True)
    mid, ax, ext = obb_parts(o)
    proj = project_all(p, ax)
    lo, hi = proj.min(0), proj.max(0)
    assert np.allclose((hi - lo) / 2, ext, rtol=1e-5, atol=1e-6)
    local = project_all(p - mid, ax)
    assert np.all(np.abs(local) <= ext + 1e-4 * np.abs(ext).max())
    al = p.max(0).astype(np.float64) - p.min(0)
    q = lambda l: l[0] * l[1] + l[0] * l[2] + l[1] * l[2]
    assert q(2 * ext) <= q(al) * (1 + 1e-6)
    assert np.allclose(np.abs(ax @ ax.T), np.eye(3), atol=1e-5)


def test_triangle_box_is_best_edge_aligned_rectangle(oracle_lib):
    # Three points: the base triangle is the triangle itself; its candidates
    # are the rectangles aligned with each edge in the triangle's plane, and
    # the answer is the smallest of those (or the AABB if smaller).  Computed
    # here in float64 from geometry alone.
    rng = np.random.default_rng(5)
    for _ in range(100):
        t = rng.uniform(-3, 3, size=(3, 3)).astype(np.float32)
        o, tr = oracle_lib.dito14(t, trace=True)
        _, _, ext = obb_parts(o)
        T = t.astype(np.float64)
        nrm = np.cross(T[1] - T[0], T[2] - T[0])
        nrm /= np.linalg.norm(nrm)
        best = np.inf
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = (T[b] - T[a]) / np.linalg.norm(T[b] - T[a])
            m = np.cross(e, nrm)
            pe, pm = T @ e, T @ m
            best = min(best, (pe.max() - pe.min()) * (pm.max() - pm.min()))
        al = T.max(0) - T.min(0)
        qal = al[0] * al[1] + al[0] * al[2] + al[1] * al[2]
        got = 4 * (ext[0] * ext[1] + ext[0] * ext[2] + ext[1] * ext[2])
        assert got == pytest.approx(min(best, qal), rel=1e-4, abs=1e-6)


def test_small_sets_use_all_points(oracle_lib):
    # n <= 14: the candidate list is the input itself (dito.cpp:519-520); the
    # trace indices then index the input.
    rng = np.random.default_rng(9)
    p = rng.standard_normal((9, 3)).astype(np.float32)
    o, tr = oracle_lib.dito14(p, trace=True)
    assert 0 <= tr["p2_sel"] < 9 and 0 <= tr["q0_sel"] < 9 and 0 <= tr["q1_sel"] < 9


def test_points_obb_floors_extent(oracle_lib):
    # kr_points_obb = DiTO_14 + ext = max(ext, 0.001) (geometry.cpp:249).
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)  # planar -> one zero extent
    raw = oracle_lib.dito14(p)
    fl = oracle_lib.points_obb(p)
    assert np.array_equal(fl[:12], raw[:12])
    assert np.array_equal(fl[12:], np.maximum(raw[12:], np.float32(0.001)))


def test_batch_equals_single(oracle_lib):
    sizes = np.array([0, 1, 3, 14, 15, 100, 5000])
    pts, off = synth.batch_sets(sizes, seed=1)
    out, tr = oracle_lib.points_obb_batch(pts, off)
    for s in range(sizes.size):
        one = oracle_lib.points_obb(pts[off[s]:off[s + 1]])
        assert np.array_equal(out[s], one)


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_threaded_batch_equals_sequential(oracle_lib, threads):
    """The all-cores timing run fits the same boxes, traces included."""
    sizes = np.array([0, 1, 3, 14, 15, 100, 5000] * 9 + [20000, 7])
    pts, off = synth.batch_sets(sizes, seed=4)
    out1, tr1 = oracle_lib.points_obb_batch(pts, off)
    outn, trn = oracle_lib.points_obb_batch(pts, off, threads=threads)
    assert np.array_equal(out1.view(np.uint32), outn.view(np.uint32))
    assert tr1.tobytes() == trn.tobytes()


# ---------------------------------------------------------------- non-finite --
# dito.cpp:116-173 and :193-206 start every running extreme at point 0 and
# compare strictly, so a NaN never wins a compare and a NaN at point 0 is
# never replaced.  These fix the oracle's answer on non-finite input without
# its code: a point that is NaN in every coordinate behaves exactly like a
# copy of point 0 (which also never wins a strict compare), and a NaN point 0
# leaves every slab at (NaN, index 0).

def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize("j", [1, 7, 500, 999])
def test_nan_point_mid_set_acts_like_a_copy_of_point_0(oracle_lib, j):
    p = synth.gaussian_points(1000, seed=41)
    q = p.copy()
    q[j] = np.nan
    r = p.copy()
    r[j] = p[0]
    got, gtr = oracle_lib.points_obb(q, trace=True)
    exp, etr = oracle_lib.points_obb(r, trace=True)
    assert _bits_equal(got, exp)
    for f in ("slab_min_idx", "slab_max_idx", "pair_k", "p2_sel", "q0_sel", "q1_sel", "best", "result_aabb"):
        assert np.array_equal(gtr[f], etr[f]), f
    assert not np.isnan(got).any()


@pytest.mark.parametrize("n", [1, 10, 1000])
def test_nan_point_0_sticks(oracle_lib, n):
    p = synth.gaussian_points(n, seed=42)
    p[0] = np.nan
    o, tr = oracle_lib.points_obb(p, trace=True)
    mid, ax, ext = obb_parts(o)
    assert np.all(tr["slab_min_idx"] == 0) and np.all(tr["slab_max_idx"] == 0)
    # every comparison against NaN fails: no candidate beats the NaN AABB
    # quality, the box is the AABB branch with NaN mid and extents
    assert tr["best"] == -1 and tr["result_aabb"] == 1
    assert np.isnan(mid).all() and np.isnan(ext).all()
    assert np.array_equal(ax, np.eye(3))


def test_nan_x_of_point_0_sticks_on_x_slabs_only(oracle_lib):
    p = synth.gaussian_points(500, seed=43)
    p[0, 0] = np.nan
    _, tr = oracle_lib.points_obb(p, trace=True)
    # slabs 0 and 3..6 contain x; slabs 1 (y) and 2 (z) are the plain
    # first-index argmin / argmax over all points
    for k in (0, 3, 4, 5, 6):
        assert tr["slab_min_idx"][k] == 0 and tr["slab_max_idx"][k] == 0
    assert tr["slab_min_idx"][1] == int(np.argmin(p[:, 1])) and tr["slab_max_idx"][1] == int(np.argmax(p[:, 1]))
    assert tr["slab_min_idx"][2] == int(np.argmin(p[:, 2])) and tr["slab_max_idx"][2] == int(np.argmax(p[:, 2]))


def test_infinite_coordinate_is_the_extreme(oracle_lib):
    p = synth.gaussian_points(300, seed=44)
    p[17, 0] = np.inf
    p[40, 1] = -np.inf
    _, tr = oracle_lib.points_obb(p, trace=True)
    assert tr["slab_max_idx"][0] == 17 and tr["slab_min_idx"][1] == 40
    # point 17 projects to +inf on slabs 3..6; point 40 to -inf on x+y+z and
    # x+y-z, to +inf on x-y+z and x-y-z, where it ties with point 17 and the
    # strict compare keeps the first (17)
    for k in (3, 4, 5, 6):
        assert tr["slab_max_idx"][k] == 17, k
    assert tr["slab_min_idx"][3] == 40 and tr["slab_min_idx"][4] == 40
    assert tr["slab_min_idx"][5] not in (17, 40) and tr["slab_min_idx"][6] not in (17, 40)


def test_all_nan_set(oracle_lib):
    p = np.full((64, 3), np.nan, np.float32)
    o, tr = oracle_lib.points_obb(p, trace=True)
    assert np.all(tr["slab_min_idx"] == 0) and np.all(tr["slab_max_idx"] == 0)
    assert np.isnan(o[:3]).all() and np.isnan(o[12:]).all()
