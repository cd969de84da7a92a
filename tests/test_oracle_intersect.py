"""Pins the oracle's ray-triangle test (Embree MoellerTrumboreIntersector1,
triangle_intersector_moeller.h:75-113) with closed forms and brute force, and
its BVH against brute force."""
import math

import numpy as np
import pytest

from lantern_amd.scenes import SceneDesc
from oracle import binding as ob


def _scene(tris):
    d = SceneDesc()
    m = d.add_material(bsdf=0, albedo=(0.5, 0.5, 0.5))
    tris = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    d.add_mesh(tris.reshape(-1, 3), np.arange(3 * len(tris), dtype=np.uint32).reshape(-1, 3), material=m)
    return ob.OracleScene(d)


def _ray(o, dvec, tnear=0.0, tfar=np.inf):
    return np.array([[*o, tnear, *dvec, tfar]], np.float32)


TRI = [[0, 0, 5], [1, 0, 5], [0, 1, 5]]


def test_closed_form_hit_t_u_v():
    s = _scene([TRI])
    tuv, geom, prim = s.intersect(_ray((0, 0, 0), (0.25, 0.25, 5.0)), use_bvh=False)
    assert geom[0] == 0 and prim[0] == 0
    # hit point (0.25, 0.25, 5) = (1-u-v) v0 + u v1 + v v2 with u = v = 0.25, t = 1
    assert tuv[0, 0] == pytest.approx(1.0, rel=1e-7)
    assert tuv[0, 1] == pytest.approx(0.25, abs=1e-7)
    assert tuv[0, 2] == pytest.approx(0.25, abs=1e-7)


@pytest.mark.parametrize("target,hit", [
    ((0.5, 0.0, 5.0), True),    # on edge v0-v1 (V = 0): inclusive
    ((0.0, 0.5, 5.0), True),    # on edge v0-v2 (U = 0): inclusive
    ((0.5, 0.5, 5.0), True),    # on edge v1-v2 (U + V = |den|): inclusive
    ((0.0, 0.0, 5.0), True),    # vertex
    ((0.6, 0.6, 5.0), False),   # outside
    ((-0.01, 0.5, 5.0), False),
])
def test_edges_are_inclusive(target, hit):
    # axis-parallel rays through exactly representable points: U, V and U+V
    # are computed without rounding, so the edge cases are exact
    s = _scene([TRI])
    o = (target[0], target[1], 0.0)
    tuv, geom, _ = s.intersect(_ray(o, (0.0, 0.0, 1.0)), use_bvh=False)
    assert (geom[0] != 0xFFFFFFFF) == hit
    if hit:
        assert tuv[0, 0] == 5.0


def test_parallel_ray_misses():
    s = _scene([TRI])
    _, geom, _ = s.intersect(_ray((0.1, 0.1, 5.0), (1.0, 0.0, 0.0)), use_bvh=False)
    assert geom[0] == 0xFFFFFFFF


def test_tnear_strict_tfar_inclusive():
    s = _scene([TRI])
    # t = 1 exactly along (0.25,0.25,5)
    _, g1, _ = s.intersect(_ray((0, 0, 0), (0.25, 0.25, 5.0), tnear=1.0), use_bvh=False)
    _, g2, _ = s.intersect(_ray((0, 0, 0), (0.25, 0.25, 5.0), tfar=1.0), use_bvh=False)
    _, g3, _ = s.intersect(_ray((0, 0, 0), (0.25, 0.25, 5.0), tfar=0.999), use_bvh=False)
    assert g1[0] == 0xFFFFFFFF and g2[0] == 0 and g3[0] == 0xFFFFFFFF


def test_both_facing_orientations_hit():
    s = _scene([TRI, [[0, 0, 7], [0, 1, 7], [1, 0, 7]]])
    tuv, g, p = s.intersect(_ray((0.1, 0.1, 10), (0, 0, -1)), use_bvh=False)
    assert p[0] == 1 and tuv[0, 0] == pytest.approx(3.0)
    tuv, g, p = s.intersect(_ray((0.1, 0.1, 0), (0, 0, 1)), use_bvh=False)
    assert p[0] == 0 and tuv[0, 0] == pytest.approx(5.0)


def test_exact_tie_later_triangle_wins_in_brute_force():
    # two identical triangles: same t exactly; `T <= |den| tfar` lets the later one replace
    s = _scene([TRI, TRI])
    tuv, g, p, ties = s.intersect(_ray((0, 0, 0), (0.25, 0.25, 5.0)), use_bvh=False, ties=True)
    assert p[0] == 1 and ties[0] == 2


def _soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    return (c + rng.normal(0, 0.6, (n, 3, 3))).astype(np.float32)


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-8, 8, (n, 3))
    d = rng.normal(size=(n, 3))
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = rng.choice([0.0, 0.001], n)
    r[:, 4:7] = d
    r[:, 7] = np.where(rng.uniform(size=n) < 0.2, rng.uniform(1, 10, n), np.inf)
    return r


def test_bvh_matches_brute_force_on_soup():
    tris = _soup(3000, 1)
    s = _scene(tris)
    rays = _rays(3000, 2)
    a = s.intersect(rays, use_bvh=False, ties=True)
    b = s.intersect(rays, use_bvh=True)
    no_tie = a[3] <= 1
    assert np.array_equal(a[2][no_tie], b[2][no_tie])
    assert np.array_equal(a[0][no_tie], b[0][no_tie])
    assert (a[2] != 0xFFFFFFFF).sum() > 1000


def test_t_against_float64_geometry():
    tris = _soup(500, 3)
    s = _scene(tris)
    rays = _rays(2000, 4)
    tuv, g, p = s.intersect(rays, use_bvh=False)
    hit = p != 0xFFFFFFFF
    assert hit.sum() > 200
    # exact intersection of the ray with the triangle's plane, in float64
    T = tris.astype(np.float64)[p[hit]]
    o = rays[hit, 0:3].astype(np.float64)
    d = rays[hit, 4:7].astype(np.float64)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    t_exact = np.sum(n * (T[:, 0] - o), axis=1) / np.sum(n * d, axis=1)
    assert np.allclose(tuv[hit, 0], t_exact, rtol=1e-4, atol=1e-5)
    # barycentrics reproduce the hit point
    u, v = tuv[hit, 1].astype(np.float64), tuv[hit, 2].astype(np.float64)
    P = (1 - u - v)[:, None] * T[:, 0] + u[:, None] * T[:, 1] + v[:, None] * T[:, 2]
    assert np.allclose(P, o + t_exact[:, None] * d, atol=2e-4)
    assert np.all(u >= 0) and np.all(v >= 0) and np.all(u + v <= 1 + 1e-6)
