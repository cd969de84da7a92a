"""The oracle's FP32 kernels against libm, and the camera/filter closed forms
(reconstruction_filter.cpp:34-142, pinhole_camera.cpp:27-40, 88-113)."""
import ctypes as C
import math

import numpy as np

from oracle import binding as ob


def _ulp_diff(a, b):
    a = np.float32(a).view(np.int32).astype(np.int64)
    b = np.float32(b).view(np.int32).astype(np.int64)
    return abs(a - b)


def test_logf_within_2ulp():
    L = ob.lib()
    xs = np.concatenate([np.random.default_rng(0).uniform(1e-30, 1.0, 3000), np.linspace(0.5, 2.0, 1000),
                         np.float32(2.0) ** np.arange(-120, 120)]).astype(np.float32)
    worst = max(_ulp_diff(L.orc_logf(float(x)), np.float32(math.log(float(x)))) for x in xs)
    assert worst <= 2
    assert L.orc_logf(0.0) == -math.inf
    assert L.orc_logf(1.0) == 0.0


def test_expf_within_2ulp():
    L = ob.lib()
    xs = np.linspace(-86.0, 88.0, 4001).astype(np.float32)
    worst = max(_ulp_diff(L.orc_expf(float(x)), np.float32(math.exp(float(x)))) for x in xs)
    assert worst <= 2
    assert L.orc_expf(0.0) == 1.0
    assert L.orc_expf(-200.0) == 0.0


def test_sincos_turns_accuracy_and_quadrants():
    L = ob.lib()
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for u in np.random.default_rng(1).uniform(0.0, 1.0, 4000).astype(np.float32):
        L.orc_sincos_turns(float(u), C.byref(s), C.byref(c))
        th = 2.0 * math.pi * float(u)
        worst = max(worst, abs(s.value - math.sin(th)), abs(c.value - math.cos(th)))
    assert worst < 2.5e-7
    for u, (es, ec) in [(0.0, (0.0, 1.0)), (0.25, (1.0, 0.0)), (0.5, (0.0, -1.0)), (0.75, (-1.0, 0.0))]:
        L.orc_sincos_turns(u, C.byref(s), C.byref(c))
        assert abs(s.value - es) < 1e-7 and abs(c.value - ec) < 1e-7


def test_tent_cdf_closed_form():
    cdf = np.zeros(33, np.float32)
    ob.lib().orc_filter_cdf(cdf.ctypes.data_as(C.POINTER(C.c_float)))
    assert cdf[0] == 0.0 and cdf[32] == 1.0
    assert np.all(np.diff(cdf) >= 0)
    # pdf[i] proportional to 1 - i/31 for i < 31, pdf[31] = 0 (reconstruction_filter.cpp:117-141)
    w = 1.0 - np.arange(31) / 31.0
    ref = np.concatenate([[0.0], np.cumsum(w / w.sum())])
    assert np.allclose(cdf[:32], ref, atol=2e-7)
    assert abs(cdf[31] - 1.0) < 5e-7  # float partial sums of pdf[0..30]


def test_tent_sample_symmetry_and_range():
    L = ob.lib()
    xs = np.linspace(0.0, 0.999, 2001).astype(np.float32)
    v = np.array([L.orc_filter_sample(0, float(x)) for x in xs])
    assert v.min() >= -1.0 and v.max() <= 1.0
    assert np.all(np.diff(v[xs < 0.5]) <= 0) and np.all(np.diff(v[xs >= 0.5]) >= 0)
    # x and x + 0.5 map to -u and +u (x + 0.5 is rounded, so up to a few ulp)
    for x in np.linspace(0.0, 0.49, 50).astype(np.float32):
        a = L.orc_filter_sample(0, float(x))
        b = L.orc_filter_sample(0, float(np.float32(x + np.float32(0.5))))
        assert abs(a + b) < 2e-6
    assert L.orc_filter_sample(1, 0.3) == 0.0            # Dirac
    assert abs(L.orc_filter_sample(2, 0.75) - 0.25) < 1e-7  # Box: x - 0.5


def _cam(**kw):
    c = dict(phi=math.pi / 2, theta=0.0, radius=10.0, width=64, height=64, target=(0.0, 0.0, 0.0), fov=math.pi / 4,
             filter=1)
    c.update(kw)
    return c


def test_camera_basis_and_center_ray():
    cam = _cam()
    basis = np.zeros(14, np.float32)
    ob.lib().orc_camera_basis(C.byref(ob.camera(cam)), basis.ctypes.data_as(C.POINTER(C.c_float)))
    origin = basis[0:3]
    assert np.allclose(origin, [0.0, 0.0, 10.0], atol=1e-5)       # phi = pi/2, theta = 0: on +z
    assert np.allclose(basis[9:12], [0.0, 0.0, 1.0], atol=1e-6)   # z axis points at the camera
    assert abs(basis[12] - math.tan(math.pi / 8)) < 1e-6
    # pixel centres of an even image straddle the axis; with the Dirac filter the
    # direction is (+-t/W*tanX, ..., -1) in camera space: symmetric about -z
    rays = np.zeros((2, 8), np.float32)
    px = np.array([31, 32], np.uint32)
    py = np.array([31, 32], np.uint32)
    r = np.zeros(2, np.float32)
    ob.lib().orc_camera_rays(C.byref(ob.camera(cam)), 2, ob.up(px), ob.up(py), ob.fp(r), ob.fp(r), ob.fp(rays))
    d0, d1 = rays[0, 4:7], rays[1, 4:7]
    assert np.allclose(d0 + d1, [0.0, 0.0, -2.0], atol=1e-6)
    assert abs(d0[2] + 1.0) < 1e-6       # unnormalized, z = -1 (pinhole_camera.cpp:103-105)
    assert rays[0, 3] == 0.0 and np.isinf(rays[0, 7])


def test_camera_corner_direction():
    cam = _cam(width=100, height=50)
    rays = np.zeros((1, 8), np.float32)
    px = np.array([0], np.uint32)
    py = np.array([0], np.uint32)
    r = np.zeros(1, np.float32)
    ob.lib().orc_camera_rays(C.byref(ob.camera(cam)), 1, ob.up(px), ob.up(py), ob.fp(r), ob.fp(r), ob.fp(rays))
    tx = math.tan(math.pi / 8)
    ty = math.tan(math.pi / 8 * 50 / 100)   # quirk: the angle is scaled, not the tangent
    ex = (0.5 / 100 * 2 - 1) * tx
    ey = -(0.5 / 50 * 2 - 1) * ty
    assert np.allclose(rays[0, 4:7], [ex, ey, -1.0], atol=1e-6)
