"""Transmittance queries through a grid medium on the GPU (lt_medium_transmittance)
against the oracle's ratio and delta tracking and against Beer-Lambert.

Both sides draw the same Philox stream per ray, so the estimates are compared
per ray: bit-identical except where a last-bit difference in a distance moves
a tentative collision across a voxel face."""
import math

import numpy as np
import pytest

import lantern_amd
from oracle.binding import OracleScene
from tests.test_oracle_shading import _medium_scene, _rays
from lantern_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [0, 1])
def test_transmittance_matches_oracle_and_beer_lambert(gpu_device, method):
    dens = np.full((8, 8, 8), 0.5, np.float32)
    dens[0, 0, 0] = 1.0
    d = _medium_scene(dens, 3.0)
    gs = lantern_amd.Scene(gpu_device, d)
    rays = _rays(50000, 1, 0.25)
    tg = gs.transmittance(0, rays, seed=7, method=method)
    to = OracleScene(d).transmittance(0, rays, seed=7, method=method)
    same = tg == to
    print(f"identical estimates {same.mean():.6f}")
    assert same.mean() > 0.999
    expect = math.exp(-3.0 * 0.5 * 0.25)
    assert abs(tg.mean() - expect) < 5.0 * tg.std() / math.sqrt(len(tg)) + 1e-4
    gs.close()


def test_transmittance_fbm_grid_matches_oracle(gpu_device):
    dens = scenes.fbm_grid(32, 5)
    d = _medium_scene(dens, 5.0)
    gs = lantern_amd.Scene(gpu_device, d)
    rays = _rays(50000, 2, 1.5)
    for method in (0, 1):
        tg = gs.transmittance(0, rays, seed=11, method=method)
        to = OracleScene(d).transmittance(0, rays, seed=11, method=method)
        assert (tg == to).mean() > 0.999
        assert np.all((tg >= 0.0) & (tg <= 1.0))
    gs.close()


def test_transmittance_rejects_bad_arguments(gpu_device):
    d = _medium_scene(np.ones((2, 2, 2), np.float32), 1.0)
    gs = lantern_amd.Scene(gpu_device, d)
    with pytest.raises(lantern_amd.LanternError):
        gs.transmittance(3, _rays(4, 0, 1.0))
    with pytest.raises(lantern_amd.LanternError):
        gs.transmittance(0, _rays(4, 0, 1.0), method=5)
    gs.close()
