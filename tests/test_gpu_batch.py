"""lt_render_batch: several scenes in one launch sequence. Each scene's film
must equal what lt_render produces for that scene alone with stream + i
(the same paths, accumulated in the same sample order), whatever the wave
packing; one scene is also checked against the oracle."""
import numpy as np
import pytest

import lantern_amd
from lantern_amd import scenes

pytestmark = pytest.mark.gpu


def _films(dev, descs):
    return [lantern_amd.Film(dev, d.camera["width"], d.camera["height"]) for d in descs]


@pytest.mark.parametrize("wave_paths", [0, 20000, 300000])
def test_batch_equals_scene_by_scene(gpu_device, wave_paths):
    descs = scenes.batch_scenes(6, subdivs=(3, 4), size_range=(24, 80), spp=3, max_depth=8)
    ss = [lantern_amd.Scene(gpu_device, d) for d in descs]
    cams = [d.camera for d in descs]
    params = dict(descs[0].params, wave_paths=wave_paths, stream=5)
    fa, fb = _films(gpu_device, descs), _films(gpu_device, descs)
    st = lantern_amd.render_batch(ss, cams, params, fa)
    lantern_amd.render_batch(ss, cams, params, fb, sequential=True)
    assert st["samples"] == sum(d.camera["width"] * d.camera["height"] * 3 for d in descs)
    for i in range(len(descs)):
        a, b = fa[i].download(), fb[i].download()
        for x, y in zip(a, b):
            assert np.array_equal(x, y), f"scene {i}"
        # and lt_render with stream 5 + i
        f = lantern_amd.Film(gpu_device, cams[i]["width"], cams[i]["height"])
        lantern_amd.render(ss[i], cams[i], dict(params, stream=5 + i), f)
        assert np.array_equal(f.download()[0], a[0])
        f.close()
    for f in fa + fb:
        f.close()
    for s in ss:
        s.close()


def test_batch_scene_matches_oracle(gpu_device):
    from oracle.binding import OracleScene
    descs = scenes.batch_scenes(3, subdivs=(4,), size_range=(32, 48), spp=2, max_depth=8)
    ss = [lantern_amd.Scene(gpu_device, d) for d in descs]
    fa = _films(gpu_device, descs)
    lantern_amd.render_batch(ss, [d.camera for d in descs], dict(descs[0].params, stream=0), fa)
    for i, d in enumerate(descs):
        so, bo, co, _, _ = OracleScene(d).render(d.camera, dict(d.params), precision=32, stream=i)
        sg, bg, cg = fa[i].download()
        assert np.array_equal(cg, co)
        close = np.all(np.abs(sg - so) <= 1e-5 * np.maximum(np.abs(so), 1.0), axis=2)
        assert close.mean() >= 0.999
    for f in fa:
        f.close()
    for s in ss:
        s.close()


def test_batch_rejects_mixed_layouts(gpu_device):
    descs = scenes.batch_scenes(2, subdivs=(2,), size_range=(16, 16), spp=1)
    a = lantern_amd.Scene(gpu_device, descs[0], layout=0)
    b = lantern_amd.Scene(gpu_device, descs[1], layout=1)
    fs = _films(gpu_device, descs)
    with pytest.raises(lantern_amd.LanternError):
        lantern_amd.render_batch([a, b], [d.camera for d in descs], dict(descs[0].params), fs)
    for f in fs:
        f.close()
    a.close()
    b.close()
