"""Image quality at equal samples per pixel: the GPU estimator (per-path
Philox streams) against the CPU reference mode (per-tile PCG32 streams, the
reference's seeding), both measured against a 4096-spp FP64 render of the
same estimator with independent random numbers.

relMSE = mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2), with the
rendered value sum/count (visualizer.cpp:362-365). The estimate of one image
is noisy (a few bright paths dominate it, +-15 % between seeds), so each side
is averaged over six independent renders (seeds for the GPU, frame offsets for
the reference mode's tile seeding). Bar: the GPU's mean relMSE is within 10 %
of the reference mode's, and both fall as 1/spp."""
import numpy as np
import pytest

import lantern_amd
from lantern_amd import scenes
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def _rel_mse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


@pytest.mark.parametrize("workload", ["nine_sphere", "triangle_heavy"])
def test_gpu_rel_mse_matches_reference_mode(gpu_device, workload):
    if workload == "nine_sphere":
        d = scenes.nine_sphere(96, 96, 16)
        window = None
    else:
        d = scenes.triangle_heavy(1280, 720, 16)
        window = (608, 328, 672, 392)     # 64x64 centre crop
    params = dict(d.params)
    os_ = ob.OracleScene(d)
    ref_params = dict(params, spp=4096, seed=params["seed"] ^ 0x5EED)
    rs, _, rc, _, _ = os_.render(d.camera, ref_params, precision=64, window=window)
    sl = (slice(None), slice(None)) if window is None else (slice(window[1], window[3]), slice(window[0], window[2]))
    ref = (rs / np.maximum(rc, 1)[..., None])[sl]

    gs = lantern_amd.Scene(gpu_device, d)
    film = lantern_amd.Film(gpu_device, d.camera["width"], d.camera["height"])
    means_gpu, means_cpu = [], []
    for spp in (4, 16):
        eg, ec = [], []
        for k in range(6):
            film.reset()
            lantern_amd.render(gs, d.camera, dict(params, spp=spp, seed=params["seed"] + 7919 * k), film)
            s, _, c = film.download()
            eg.append(_rel_mse((s / c[..., None])[sl], ref))
            cs, _, cc, _, _ = os_.render(d.camera, dict(params, spp=spp), precision=32, rng=ob.ORC_RNG_PCG_TILE,
                                         window=window, frame0=1000 * k)
            ec.append(_rel_mse((cs / np.maximum(cc, 1)[..., None])[sl], ref))
        means_gpu.append(float(np.mean(eg)))
        means_cpu.append(float(np.mean(ec)))
    film.close()
    gs.close()
    print(workload, "mean relMSE at 4/16 spp: gpu", means_gpu, "cpu reference mode", means_cpu)
    for g, c in zip(means_gpu, means_cpu):
        assert g <= 1.1 * c
    assert means_gpu[1] < 0.5 * means_gpu[0] and means_cpu[1] < 0.5 * means_cpu[0]
