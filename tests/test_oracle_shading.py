"""The oracle's BSDFs, phase function, media and estimator against closed
forms (lambert_bsdf.h, mirror_bsdf.h, ideal_specular_dielectric.h,
vector_math.h:33-56, sampling.h:31-60, integrator.cpp:98-250).

Each check uses a form independent of the code under test: Fresnel from the
angle form of the Fresnel equations, sampling moments from their integrals,
transmittance from Beer-Lambert, and whole-path results from scenes whose
radiance is known exactly."""
import math

import numpy as np
import pytest

from lantern_amd import scenes
from oracle import binding as ob

LAMBERT, MIRROR, DIELECTRIC, GGX = 0, 1, 2, 3
LOBE_SPEC_REFL, LOBE_SPEC_TRANS, LOBE_DIFFUSE, LOBE_GLOSSY = 1, 2, 4, 8


def _fresnel_angles(n1, n2, ci):
    """Unpolarised Fresnel reflectance from the angle form
    Rs = sin^2(ti-tt)/sin^2(ti+tt), Rp = tan^2(ti-tt)/tan^2(ti+tt)."""
    ti = math.acos(ci)
    st = n1 / n2 * math.sin(ti)
    if st >= 1.0:
        return 1.0
    tt = math.asin(st)
    if ti == 0.0:
        return ((n1 - n2) / (n1 + n2)) ** 2
    rs = math.sin(ti - tt) ** 2 / math.sin(ti + tt) ** 2
    rp = math.tan(ti - tt) ** 2 / math.tan(ti + tt) ** 2
    return 0.5 * (rs + rp)


def test_fresnel_normal_incidence_and_tir():
    assert ob.fresnel(1.0, 1.5, 1.0) == pytest.approx(0.04, abs=1e-6)
    assert ob.fresnel(1.5, 1.0, 1.0) == pytest.approx(0.04, abs=1e-6)
    # beyond the critical angle asin(1/1.5) = 41.8 deg from glass to air
    assert ob.fresnel(1.5, 1.0, math.cos(math.radians(45.0))) == 1.0
    assert ob.fresnel(1.0, 1.0, 0.3) == pytest.approx(0.0, abs=1e-6)


@pytest.mark.parametrize("n1,n2", [(1.0, 1.5), (1.5, 1.0), (1.0, 1.35), (1.33, 1.0)])
def test_fresnel_matches_angle_form(n1, n2):
    for ci in np.linspace(0.02, 1.0, 60):
        assert ob.fresnel(n1, n2, float(ci)) == pytest.approx(_fresnel_angles(n1, n2, float(ci)), abs=2e-5)


def _grid_uv(n):
    g = (np.arange(n) + 0.5) / n
    u1, u2 = np.meshgrid(g, g, indexing="ij")
    return u1.ravel(), u2.ravel()


def test_lambert_sampling_moments_and_weight():
    mat = ob.material(LAMBERT, (0.3, 0.5, 0.7))
    n = np.array([0.0, 0.6, 0.8])
    wo = np.array([0.0, 1.0, 0.0])
    cos_sum, inv_pdf_sum, k = 0.0, 0.0, 0
    for u1, u2 in zip(*_grid_uv(60)):
        lobe, wi, f, pdf, _ = ob.bsdf_sample(mat, n, wo, [u1, u2])
        assert lobe == LOBE_DIFFUSE
        c = float(np.dot(wi, n))
        assert c >= -1e-6 and abs(np.linalg.norm(wi) - 1.0) < 1e-5
        assert pdf == pytest.approx(c / math.pi, rel=1e-5, abs=1e-7)
        if pdf > 1e-3:
            np.testing.assert_allclose(f / pdf, (0.3, 0.5, 0.7), rtol=1e-5)   # Eval/Pdf = albedo (quirk A9)
        cos_sum += c
        k += 1
    assert cos_sum / k == pytest.approx(2.0 / 3.0, abs=2e-3)   # E[cos] under cos/pi sampling


def test_mirror_reflects_about_the_normal():
    mat = ob.material(MIRROR, (0.9, 0.8, 0.7))
    n = np.array([0.0, 1.0, 0.0])
    wo = np.array([0.6, 0.8, 0.0])
    lobe, wi, f, pdf, _ = ob.bsdf_sample(mat, n, wo, [0.5])
    assert lobe == LOBE_SPEC_REFL
    np.testing.assert_allclose(wi, (-0.6, 0.8, 0.0), atol=1e-6)
    np.testing.assert_allclose(f, (0.9, 0.8, 0.7))
    assert pdf == 1.0


def test_dielectric_snell_and_fresnel_split():
    mat = ob.material(DIELECTRIC, (1.0, 1.0, 1.0), ior=1.5)
    n = np.array([0.0, 1.0, 0.0])
    th = math.radians(50.0)
    wo = np.array([math.sin(th), math.cos(th), 0.0])
    F = _fresnel_angles(1.0, 1.5, math.cos(th))
    lobe, wi, f, pdf, ioro = ob.bsdf_sample(mat, n, wo, [F * 0.5])
    assert lobe == LOBE_SPEC_REFL and ioro == 1.0
    lobe, wi, f, pdf, ioro = ob.bsdf_sample(mat, n, wo, [min(0.999, F + 0.01)])
    assert lobe == LOBE_SPEC_TRANS and ioro == pytest.approx(1.5)
    sin_t = math.hypot(wi[0], wi[2])
    assert wi[1] < 0 and sin_t * 1.5 == pytest.approx(math.sin(th), abs=1e-5)   # Snell
    # from inside: the normal flips and the outgoing IOR is 1 (ideal_specular_dielectric.h:39-44)
    wo_in = np.array([math.sin(0.2), -math.cos(0.2), 0.0])
    lobe, wi, f, pdf, ioro = ob.bsdf_sample(mat, n, wo_in, [0.9], iori=1.5)
    assert lobe == LOBE_SPEC_TRANS and ioro == 1.0 and wi[1] > 0
    assert math.hypot(wi[0], wi[2]) == pytest.approx(1.5 * math.sin(0.2), abs=1e-5)


def test_ggx_reciprocity_and_energy_bound():
    mat = ob.material(GGX, (1.0, 1.0, 1.0), roughness=0.3)
    n = np.array([0.0, 0.0, 1.0])
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[2], b[2] = abs(a[2]) + 0.1, abs(b[2]) + 0.1
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        fab, _ = ob.bsdf_eval(mat, n, a, b)   # includes cos(wi) like the reference BSDFs
        fba, _ = ob.bsdf_eval(mat, n, b, a)
        assert fab[0] / b[2] == pytest.approx(fba[0] / a[2], rel=1e-4)
    # quadrature over the hemisphere at normal incidence: the sampling density
    # integrates to P(tan^2 theta_h < 1) = 1/(1 + a^2) under D(h)cos(h) (the
    # rest reflects below the horizon), and since Eval/Pdf = G1(wo) G1(wi) <= 1
    # there, the directional albedo (F0 = 1) lies below that mass
    wo = np.array([0.0, 0.0, 1.0])
    nt, npf = 400, 8
    alb, pdf_mass = 0.0, 0.0
    for i in range(nt):
        t = (i + 0.5) / nt * (math.pi / 2)
        dw = (math.pi / 2 / nt) * (2 * math.pi / npf) * math.sin(t)
        for j in range(npf):
            p = (j + 0.5) / npf * 2 * math.pi
            wi = np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])
            f, pdf = ob.bsdf_eval(mat, n, wo, wi)
            alb += f[0] * dw
            pdf_mass += pdf * dw
    assert pdf_mass == pytest.approx(1.0 / (1.0 + 0.09), abs=2e-3)
    assert 0.85 < alb <= pdf_mass


def test_ggx_sample_weight_matches_eval_over_pdf():
    mat = ob.material(GGX, (0.9, 0.6, 0.5), roughness=0.3)
    n = np.array([0.0, 0.0, 1.0])
    wo = np.array([0.5, 0.0, math.sqrt(0.75)])
    for u1, u2 in zip(*_grid_uv(12)):
        lobe, wi, f, pdf, _ = ob.bsdf_sample(mat, n, wo, [u1, u2])
        assert lobe == LOBE_GLOSSY
        f2, pdf2 = ob.bsdf_eval(mat, n, wo, wi)
        np.testing.assert_allclose(f, f2, rtol=1e-6)
        assert pdf == pytest.approx(pdf2, rel=1e-6)


@pytest.mark.parametrize("g", [0.0, 0.3, -0.5, 0.8])
def test_henyey_greenstein_mean_cosine_is_g(g):
    d = np.array([0.0, 0.0, 1.0])
    u1 = (np.arange(4000) + 0.5) / 4000
    cs = np.array([ob.sample_hg(g, d, float(u), 0.37)[2] for u in u1])
    assert cs.mean() == pytest.approx(g, abs=2e-3)
    assert np.all(np.abs(cs) <= 1.0)


def _medium_scene(density, scale, bmin=(-1.0, -1.0, -1.0), bmax=(1.0, 1.0, 1.0)):
    d = scenes.SceneDesc()
    d.grids.append(dict(density=np.asarray(density, np.float32), bmin=bmin, bmax=bmax))
    d.media.append(dict(type=2, grid=0, sigma_t_scale=scale, albedo=1.0, hg_g=0.0))
    m = d.add_material(bsdf=LAMBERT, albedo=(0.5, 0.5, 0.5))
    p, n, uv, idx = scenes.grid(1.0, 1.0, 2, 2)
    p, n = scenes.transform(p, n, scenes.translate(0.0, -5.0, 0.0))
    d.add_mesh(p, idx, n, uv, m)
    return d


def _rays(n, seed, length):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.5, 0.5, (n, 3))
    dr = rng.normal(size=(n, 3))
    dr /= np.linalg.norm(dr, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = dr
    r[:, 7] = length
    return r


@pytest.mark.parametrize("method", [0, 1])
def test_transmittance_of_constant_density_is_beer_lambert(method):
    # constant density 0.5 everywhere the rays go; one corner voxel at 1.0 raises
    # the majorant so ratio tracking takes fractional steps
    dens = np.full((8, 8, 8), 0.5, np.float32)
    dens[0, 0, 0] = 1.0
    sc = ob.OracleScene(_medium_scene(dens, 3.0))
    rays = _rays(20000, 1, 0.25)   # stays inside [-0.75, 0.75]^3, clear of the corner voxel
    tr = sc.transmittance(0, rays, seed=7, method=method)
    expect = math.exp(-3.0 * 0.5 * 0.25)
    err = 4.0 * tr.std() / math.sqrt(len(tr))
    assert abs(tr.mean() - expect) < err + 1e-4
    if method == 0:
        assert np.any((tr > 0.0) & (tr < 1.0))   # fractional weights, not a 0/1 estimate
        assert tr.std() < np.sqrt(expect * (1.0 - expect))   # below the 0/1 estimate's variance


def test_transmittance_counts_only_the_part_inside_the_grid():
    dens = np.full((4, 4, 4), 1.0, np.float32)
    sc = ob.OracleScene(_medium_scene(dens, 2.0))
    n = 20000
    r = np.zeros((n, 8), np.float32)
    r[:, 0] = -3.0
    r[:, 1] = np.linspace(-0.9, 0.9, n)
    r[:, 4] = 1.0
    r[:, 3] = 0.0
    r[:, 7] = 10.0    # crosses the whole box: 2 units inside
    tr = sc.transmittance(0, r, seed=3, method=1)   # 0/1 estimates: 5 sigma of a Bernoulli mean
    assert tr.mean() == pytest.approx(math.exp(-4.0), abs=5 * math.sqrt(math.exp(-4.0) / n))
    r[:, 7] = 2.5     # ends half a unit inside
    tr = sc.transmittance(0, r, seed=3, method=1)
    assert tr.mean() == pytest.approx(math.exp(-1.0), abs=5 * math.sqrt(0.25 / n))
    r[:, 1] = 1.5     # passes above the box
    assert np.all(sc.transmittance(0, r, seed=3, method=0) == 1.0)


def test_furnace_dielectric_plane_shows_the_pulled_in_first_hit():
    # Quirk A1 (pinhole_camera.cpp:103-110, integrator.cpp:124-125): the first
    # hit point is origin + normalize(dir) * tfar with tfar measured along the
    # unnormalised primary direction, so it lies above the plane. A sample that
    # reflects escapes at once (albedo * bg, 1 bounce); one that refracts meets
    # the plane a second time on its way down (albedo^2 * bg, 2 bounces).
    a, bg = np.array((0.8, 0.9, 1.0)), np.array((2.0, 1.0, 0.5))
    d = scenes.furnace_plane(DIELECTRIC, tuple(a), tuple(bg), ior=1.5)
    sc = ob.OracleScene(d)
    s, b, c, _, _ = sc.render(d.camera, dict(spp=1, max_depth=8, rr_start=8), threads=2)
    one = np.all(np.isclose(s, a * bg, rtol=2e-6), axis=-1) & (b == 1)
    two = np.all(np.isclose(s, a * a * bg, rtol=2e-6), axis=-1) & (b == 2)
    assert np.all(one | two)
    assert 0.03 < one.mean() < 0.15     # reflected fraction = Fresnel at 0.3-0.5 rad incidence


@pytest.mark.parametrize("bsdf,albedo,ior", [(LAMBERT, (0.3, 0.5, 0.7), 1.0), (MIRROR, (0.9, 0.8, 0.6), 1.0)])
def test_furnace_plane_is_albedo_times_background(bsdf, albedo, ior):
    bg = (2.0, 1.0, 0.5)
    d = scenes.furnace_plane(bsdf, albedo, bg, ior=ior)
    sc = ob.OracleScene(d)
    s, b, c, _, _ = sc.render(d.camera, dict(spp=4, max_depth=8, rr_start=8), threads=2)
    img = s / c[..., None]
    np.testing.assert_allclose(img, np.broadcast_to(np.multiply(albedo, bg), img.shape), rtol=2e-6)
    assert np.all(b == 4) and np.all(c == 4)   # one bounce per sample, summed over 4 samples
    # max_depth 1: the path ends at the plane before reaching the environment
    s1, b1, c1, _, _ = sc.render(d.camera, dict(spp=1, max_depth=1, rr_start=8), threads=2)
    assert np.all(s1 == 0.0) and np.all(b1 == 1)


def test_fp32_estimator_tracks_fp64():
    d = scenes.nine_sphere(48, 48, 4, subdiv=2)
    sc = ob.OracleScene(d)
    p = dict(spp=4, max_depth=8, rr_start=4)
    s32, _, c, _, _ = sc.render(d.camera, p, precision=32, threads=4)
    s64, _, _, _, _ = sc.render(d.camera, p, precision=64, threads=4)
    a = s32 / c[..., None]
    r = s64 / c[..., None]
    close = np.abs(a - r) <= 1e-4 * (1.0 + np.abs(r))
    assert close.mean() > 0.98          # the same paths; FP32 rounding only flips rare decisions
    assert np.mean((a - r) ** 2 / (r ** 2 + 1e-2)) < 1e-3
