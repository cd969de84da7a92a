"""The oracle's FP64 path-replay adjoint against central finite differences.

Scene: the differentiable cube (albedo texture on a Lambert cube inside a
grid-medium smoke box), reduced in size so that 512 FP64 renders take
seconds. Loss: mean squared difference to a render of the same scene built
from seed + 1 parameters. For fixed random numbers the estimator is a smooth
function of the texels (no Russian roulette) and, with the tracking
decisions frozen at the current densities, of the densities, so a central
difference with eps = 1e-3 measures the same derivative the adjoint returns.
Bar: 256 random parameters (128 texels, 128 voxels) within 1e-3 relative."""
import numpy as np
import pytest

from lantern_amd import scenes
from oracle import binding as ob

EPS = 1e-3
SEED = scenes.DEFAULT_SEED


def _image(sc, d):
    s, _, c, _, _ = sc.render(d.camera, d.params, precision=64, threads=4)
    return s / c[..., None]


def _loss(img, target):
    return float(np.mean((img - target) ** 2))


def _pick(rng, g, n):
    """n parameters: half uniformly at random, half among those a path touched."""
    touched = np.flatnonzero(np.abs(g) > 0)
    a = rng.choice(len(g), n // 2, replace=False)
    b = rng.choice(touched, n - n // 2, replace=False) if len(touched) >= n - n // 2 else touched
    return np.unique(np.concatenate([a, b]))


@pytest.fixture(scope="module")
def setup():
    kw = dict(width=24, height=24, spp=4, tex_res=16, grid_res=8)
    d = scenes.diff_cube(seed=SEED, **kw)
    dt = scenes.diff_cube(seed=SEED + 1, **kw)
    target = _image(ob.OracleScene(dt), dt)
    sc = ob.OracleScene(d)
    img = _image(sc, d)
    pixel_grad = 2.0 * (img - target) / img.size
    gt, gv = sc.adjoint(d.camera, d.params, pixel_grad, threads=4)
    return d, sc, target, gt, gv


def test_adjoint_texels_match_central_differences(setup):
    d, sc, target, gt, _ = setup
    tex = d.textures[0].reshape(-1).copy()
    idx = _pick(np.random.default_rng(1), gt, 128)
    fd = np.zeros(len(idx))
    for k, j in enumerate(idx):
        t = tex.copy()
        t[j] += EPS
        sc.update_texture(0, t)
        lp = _loss(_image(sc, d), target)
        t[j] -= 2 * EPS
        sc.update_texture(0, t)
        lm = _loss(_image(sc, d), target)
        fd[k] = (lp - lm) / (2 * EPS)
    sc.update_texture(0, tex)
    g = gt[idx]
    scale = np.abs(gt).max()
    print(f"texels: {np.count_nonzero(g)} touched of {len(idx)}, max |g| {scale:.3e}")
    assert np.all(np.abs(g - fd) <= 1e-3 * np.abs(fd) + 1e-9 * scale)
    assert np.count_nonzero(g) >= 60


def test_adjoint_densities_match_central_differences(setup):
    d, sc, target, _, gv = setup
    dens = d.grids[0]["density"].reshape(-1).copy()
    idx = _pick(np.random.default_rng(2), gv, 128)
    fd = np.zeros(len(idx))
    for k, j in enumerate(idx):
        v = dens.copy()
        v[j] += EPS
        sc.update_grid(0, v, freeze_sampling=True)
        lp = _loss(_image(sc, d), target)
        v[j] -= 2 * EPS
        sc.update_grid(0, v, freeze_sampling=True)
        lm = _loss(_image(sc, d), target)
        fd[k] = (lp - lm) / (2 * EPS)
    sc.update_grid(0, dens, freeze_sampling=False)
    g = gv[idx]
    scale = np.abs(gv).max()
    print(f"voxels: {np.count_nonzero(g)} touched of {len(idx)}, max |g| {scale:.3e}")
    assert np.all(np.abs(g - fd) <= 1e-3 * np.abs(fd) + 1e-9 * scale)
    assert np.count_nonzero(g) >= 60


def test_frozen_sampling_at_the_current_densities_changes_nothing(setup):
    d, sc, _, _, _ = setup
    a = _image(sc, d)
    sc.update_grid(0, d.grids[0]["density"], freeze_sampling=True)
    b = _image(sc, d)
    assert np.array_equal(a, b)
