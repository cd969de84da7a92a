"""Quad primitives as Embree treats them (RTC_GEOMETRY_TYPE_QUAD, used for LMF
files with VerticesPerPrimative == 4, scene.cpp:508-511): triangles
(v0, v1, v3) and (v2, v3, v1) with the second reporting (1 - u, 1 - v)
(quad_intersector_moeller.h:146-153), attributes interpolated on the half
that (u, v) falls on (scene_quad_mesh.cpp:238-255), and the reference's quad
area formula (scene.cpp:540-548, quirk A12). Checked against closed forms."""
import math

import numpy as np

from lantern_amd import scenes
from oracle import binding as ob


def _one_quad(v, nrm=None, uv=None, material=None):
    d = scenes.SceneDesc()
    m = d.add_material(bsdf=0, albedo=(0.5, 0.5, 0.5)) if material is None else material
    d.add_mesh(np.asarray(v, np.float32), np.array([[0, 1, 2, 3]], np.uint32), nrm, uv, m)
    return d


def _down_rays(xy, z0=5.0):
    r = np.zeros((len(xy), 8), np.float32)
    r[:, 0:2] = xy
    r[:, 2] = z0
    r[:, 6] = -1.0
    r[:, 7] = np.inf
    return r


def test_unit_square_reports_its_own_uv_on_both_halves():
    # v0 (0,0) v1 (1,0) v2 (1,1) v3 (0,1): the quad parameterisation is (x, y)
    sc = ob.OracleScene(_one_quad([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]))
    rng = np.random.default_rng(0)
    xy = rng.uniform(0.02, 0.98, (2000, 2)).astype(np.float32)
    tuv, geom, prim = sc.intersect(_down_rays(xy), use_bvh=False)
    assert np.all(geom == 0) and np.all(prim == 0)
    np.testing.assert_allclose(tuv[:, 0], 5.0, rtol=1e-6)
    np.testing.assert_allclose(tuv[:, 1], xy[:, 0], atol=2e-6)
    np.testing.assert_allclose(tuv[:, 2], xy[:, 1], atol=2e-6)


def test_non_planar_quad_is_split_on_the_v1_v3_diagonal():
    # v2 raised: the surface is the plane of (v0, v1, v3) below the diagonal
    # x + y = 1 and the plane of (v2, v3, v1) above it (not the v0-v2 split)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 1], [0, 1, 0]], np.float32)
    sc = ob.OracleScene(_one_quad(v))
    xy = np.array([[0.2, 0.3], [0.8, 0.7], [0.9, 0.4], [0.3, 0.9]], np.float32)
    tuv, _, _ = sc.intersect(_down_rays(xy), use_bvh=False)
    z = 5.0 - tuv[:, 0]
    # upper triangle plane through (1,1,1), (0,1,0), (1,0,0): z = x + y - 1
    expect = np.where(xy.sum(1) <= 1.0, 0.0, xy.sum(1) - 1.0)
    np.testing.assert_allclose(z, expect, atol=2e-6)


def test_quad_attributes_interpolate_per_half():
    # albedo = interpolated uv (UV texture) under a furnace: pixel = (tu, tv, 0) * bg
    v = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    uvs = np.array([[0.1, 0.2], [0.9, 0.3], [0.7, 0.8], [0.2, 0.6]], np.float32)
    d = scenes.SceneDesc()
    m = d.add_material(bsdf=0, albedo=(1.0, 1.0, 1.0), texture=-2)
    v = np.asarray(v, np.float32) * 40.0 - np.array([20.0, 20.0, 0.0], np.float32)
    d.add_mesh(v[:, [0, 2, 1]], np.array([[0, 3, 2, 1]], np.uint32),
               np.tile(np.array([[0, 1, 0]], np.float32), (4, 1)), uvs, m)
    d.background = (1.0, 1.0, 1.0)
    cam = dict(phi=0.3, theta=0.4, radius=4.0, width=24, height=24, target=(0.0, 0.0, 0.0), fov=math.pi / 4,
               filter=1)
    sc = ob.OracleScene(d)
    s, b, c, _, _ = sc.render(cam, dict(spp=1, max_depth=4, rr_start=4), threads=2)
    img = s[..., :2] / c[..., None]
    # the hit points of the centre rays (Dirac filter), through the oracle's camera
    L = ob.lib()
    ys, xs = np.mgrid[0:24, 0:24]
    px = np.ascontiguousarray(xs.ravel(), np.uint32)
    py = np.ascontiguousarray(ys.ravel(), np.uint32)
    rays = np.zeros((576, 8), np.float32)
    zero = np.zeros(576, np.float32)
    L.orc_camera_rays(ob.camera(cam), 576, ob.up(px), ob.up(py), ob.fp(zero), ob.fp(zero), ob.fp(rays))
    tuv, _, _ = sc.intersect(rays, use_bvh=False)
    # quad (u, v) -> attribute on the half it falls on; corners in quad order
    # q0 = idx0 (v0), q1 = idx3 (v1), q2 = idx2 (v2), q3 = idx1 (v3)
    a0, a1, a2, a3 = uvs[0], uvs[3], uvs[2], uvs[1]
    u, w = tuv[:, 1:2].astype(np.float64), tuv[:, 2:3].astype(np.float64)
    left = (u + w) <= 1.0
    lo = (1 - u - w) * a0 + u * a1 + w * a3
    hi = (u + w - 1) * a2 + (1 - u) * a3 + (1 - w) * a1
    expect = np.where(left, lo, hi).reshape(24, 24, 2)
    np.testing.assert_allclose(img, expect, atol=2e-6)
    assert left.any() and (~left).any()


def test_quad_light_area_follows_the_reference_formula():
    # AddLMF's quad area 0.5|(v0-v1)x(v0-v2)| + 0.5|(v0-v2)x(v0-v3)| with v3 =
    # vertices[Indices[i + 2]]: the second term vanishes, so a 2 x 3 rectangle
    # counts as 3, not 6, and Le doubles
    v = np.array([[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0]], np.float32)
    d = _one_quad(v)
    d.meshes[0]["emission"] = ((1.0, 1.0, 1.0), 30.0)
    lights = ob.OracleScene(d).lights()
    assert lights[0, 3] == np.float32(3.0)
    assert lights[0, 0] == np.float32(1.0 * 30.0 * np.float32(1 / math.pi) / 3.0)


def test_quad_scene_bvh_matches_brute_force():
    d = scenes.quad_scene()
    sc = ob.OracleScene(d)
    rng = np.random.default_rng(3)
    o = rng.uniform(-4, 4, (4000, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(0.5, 4.0, 4000)
    dr = rng.normal(size=(4000, 3)).astype(np.float32)
    r = np.zeros((4000, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = dr / np.linalg.norm(dr, axis=1, keepdims=True)
    r[:, 7] = np.inf
    a = sc.intersect(r, use_bvh=True)
    b = sc.intersect(r, use_bvh=False)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
