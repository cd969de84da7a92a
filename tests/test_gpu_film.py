"""Film post-processing on the GPU: the final resolve shader's tone mapping
(final_resolve_ps.glsl:24-44) against its formulas in numpy, the PFM dump
against the resolved film, and resuming a render from a saved accumulation
(lt_film_download -> lt_film_upload) against rendering in one go."""
import numpy as np
import pytest

import lantern_amd
from lantern_amd import scenes

pytestmark = pytest.mark.gpu


def _shader(rgb, op, exposure):
    c = np.concatenate([rgb, np.ones(rgb.shape[:-1] + (1,), np.float32)], -1) * np.float32(exposure)
    if op == 1:
        x = np.maximum(np.float32(0), c[..., :3] - np.float32(0.004))
        y = (x * (np.float32(6.2) * x + np.float32(0.5))) / (x * (np.float32(6.2) * x + np.float32(1.7)) + np.float32(0.06))
        c = np.concatenate([y, np.ones_like(y[..., :1])], -1)
    return np.clip(c, 0, 1).astype(np.float32)


@pytest.mark.parametrize("op,exposure", [(0, 1.0), (0, 0.7), (1, 1.0), (1, 2.5)])
def test_tonemap_matches_the_resolve_shader(gpu_device, op, exposure):
    rng = np.random.default_rng(op * 10 + int(exposure * 10))
    w, h = 37, 21
    s = rng.uniform(-0.5, 6.0, (h, w, 3)).astype(np.float32)
    c = rng.integers(0, 9, (h, w)).astype(np.uint32)
    film = lantern_amd.Film(gpu_device, w, h)
    film.upload(s, np.zeros((h, w), np.uint32), c)
    got = film.tonemap(op, exposure)
    rgb = np.where(c[..., None] > 0, s / np.maximum(c, 1)[..., None].astype(np.float32), 0).astype(np.float32)
    np.testing.assert_allclose(got, _shader(rgb, op, exposure), atol=2e-7)
    film.close()


def test_pfm_dump_and_resume(gpu_device, tmp_path):
    d = scenes.nine_sphere(40, 30, 4, subdiv=2)
    gs = lantern_amd.Scene(gpu_device, d)
    one = lantern_amd.Film(gpu_device, 40, 30)
    lantern_amd.render(gs, d.camera, dict(d.params, spp=4), one)
    a = lantern_amd.Film(gpu_device, 40, 30)
    lantern_amd.render(gs, d.camera, dict(d.params, spp=2, frame0=0), a)
    saved = a.download()
    a.close()
    b = lantern_amd.Film(gpu_device, 40, 30)
    b.upload(*saved)
    lantern_amd.render(gs, d.camera, dict(d.params, spp=2, frame0=2), b)
    for x, y in zip(one.download(), b.download()):
        assert np.array_equal(x, y)
    path = tmp_path / "img.pfm"
    b.save_pfm(path)
    raw = open(path, "rb").read()
    header = raw.split(b"\n", 3)
    assert header[0] == b"PF" and header[1] == b"40 30" and float(header[2]) < 0
    img = np.frombuffer(header[3], "<f4").reshape(30, 40, 3)[::-1]
    np.testing.assert_array_equal(img, b.resolve())
    b.close()
    one.close()
    gs.close()
