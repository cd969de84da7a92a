"""The C++ facade end to end on the GPU: a Lantern-format JSON scene loaded by
Scene::LoadSceneFromJSON and rendered by Integrator::RenderFrame per frame
(the way source/lantern/main.cpp drives them) equals the same scene built
through the Python mirror and rendered by lt_render with the Integrator's
defaults (max depth 1500, RR from 4 bounces, seed 0x4C414E54): identical
sums, bounces and counts, because every frame is one sample per pixel with
the same per-path streams and the host FrameBuffer adds the frames in order."""
import os
import subprocess

import numpy as np
import pytest

import lantern_amd
from tests.facade_common import GOLDEN, build_driver, desc_from_json, read_output

pytestmark = pytest.mark.gpu


def test_json_scene_through_the_facade_equals_lt_render(gpu_device, tmp_path):
    driver = build_driver()
    path = os.path.join(GOLDEN, "facade_scene.json")
    out = tmp_path / "fb.bin"
    r = subprocess.run([driver, path, "4", str(out)], capture_output=True, text=True, timeout=200)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    s_f, b_f, c_f = read_output(out)
    d = desc_from_json(path)
    gs = lantern_amd.Scene(gpu_device, d)
    film = lantern_amd.Film(gpu_device, d.camera["width"], d.camera["height"])
    lantern_amd.render(gs, d.camera, dict(spp=4, max_depth=1500, rr_start=4, seed=0x4C414E54), film)
    s, b, c = film.download()
    assert np.array_equal(c_f, c) and np.all(c == 4)
    assert np.array_equal(b_f, b)
    assert np.array_equal(s_f, s)
    assert s.mean() > 0.01
    film.close()
    gs.close()
