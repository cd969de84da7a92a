"""Differentiable rendering over the C ABI (lt_render_record / lt_adjoint).

The parameters are one image texture (H x W x 3 texels) and one grid medium
(D x H x W densities) of a committed scene. `DiffRenderer.forward` copies
them into the scene on the GPU, renders while recording every sample's
radiance, and returns the image (pixel means). `DiffRenderer.backward` runs
the path-replay adjoint for dLoss/dimage and returns the two gradients.
`render` wraps both in a torch.autograd.Function. torch supplies device
memory only; all rendering and gradient work is the library's.

One forward pass per backward pass: the recorded radiance of the last
forward pass is what the adjoint replays.
"""
import ctypes as C

import torch

from . import _capi, camera_struct, params_struct
from ._capi import check, lib

__all__ = ["DiffRenderer", "render"]


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class DiffRenderer:
    def __init__(self, scene, camera, params, texture_id=0, grid_id=0):
        p = dict(params)
        if p.get("rr_start", 4) < p.get("max_depth", 1500):
            raise ValueError("the adjoint needs Russian roulette off: rr_start >= max_depth")
        self.scene = scene
        self.w, self.h, self.spp = camera["width"], camera["height"], p["spp"]
        self.cam = camera_struct(camera)
        self.prm = params_struct(p)
        self.texture_id, self.grid_id = texture_id, grid_id
        L = lib()
        self.film = C.c_void_p()
        check(L.lt_film_create(scene.dev.h, self.w, self.h, C.byref(self.film)), "lt_film_create")
        self.tex_off, self.tex_n = self._range(_capi.LT_PARAM_TEXTURE, texture_id)
        self.vox_off, self.vox_n = self._range(_capi.LT_PARAM_GRID, grid_id)
        _, self.texel_pool = self._range(_capi.LT_PARAM_TEXEL_POOL, 0)
        _, self.voxel_pool = self._range(_capi.LT_PARAM_VOXEL_POOL, 0)
        self.rad = torch.empty(self.spp * self.w * self.h * 3, dtype=torch.float32, device="cuda")
        self.stats = _capi.lt_render_stats()

    def _range(self, kind, i):
        off, n = C.c_uint32(), C.c_uint32()
        check(lib().lt_scene_param_range(self.scene.h, kind, i, C.byref(off), C.byref(n)), "lt_scene_param_range")
        return off.value, n.value

    def forward(self, texture, density):
        """texture (H, W, 3) and density (D, H, W) float32 CUDA tensors -> image (h, w, 3)."""
        texture = texture.detach().contiguous()
        density = density.detach().contiguous()
        if texture.numel() != self.tex_n or density.numel() != self.vox_n:
            raise ValueError("parameter sizes differ from the scene's texture/grid")
        torch.cuda.current_stream().synchronize()
        L = lib()
        check(L.lt_scene_update_texture(self.scene.h, self.texture_id, _ptr(texture), 1), "lt_scene_update_texture")
        check(L.lt_scene_update_grid(self.scene.h, self.grid_id, _ptr(density), 1), "lt_scene_update_grid")
        check(L.lt_film_reset(self.film), "lt_film_reset")
        check(L.lt_render_record(self.scene.h, C.byref(self.cam), C.byref(self.prm), self.film, _ptr(self.rad),
                                 C.byref(self.stats)), "lt_render_record")
        check(L.lt_device_synchronize(self.scene.dev.h), "lt_device_synchronize")
        return self.rad.view(self.spp, self.h, self.w, 3).mean(0)

    def backward(self, grad_image):
        """dLoss/dimage (h, w, 3) -> (dLoss/dtexture, dLoss/ddensity)."""
        g = grad_image.detach().to(torch.float32).contiguous()
        gt = torch.zeros(max(self.texel_pool, 1), dtype=torch.float32, device="cuda")
        gv = torch.zeros(max(self.voxel_pool, 1), dtype=torch.float32, device="cuda")
        torch.cuda.current_stream().synchronize()
        L = lib()
        check(L.lt_adjoint(self.scene.h, C.byref(self.cam), C.byref(self.prm), _ptr(self.rad), _ptr(g), _ptr(gt),
                           _ptr(gv), C.byref(self.stats)), "lt_adjoint")
        check(L.lt_device_synchronize(self.scene.dev.h), "lt_device_synchronize")
        return gt[self.tex_off:self.tex_off + self.tex_n], gv[self.vox_off:self.vox_off + self.vox_n]

    def close(self):
        if self.film:
            lib().lt_film_destroy(self.film)
            self.film = C.c_void_p()


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, texture, density, renderer):
        ctx.renderer = renderer
        ctx.shapes = (texture.shape, density.shape)
        return renderer.forward(texture, density)

    @staticmethod
    def backward(ctx, grad_image):
        gt, gv = ctx.renderer.backward(grad_image)
        return gt.view(ctx.shapes[0]), gv.view(ctx.shapes[1]), None


def render(texture, density, renderer):
    """Differentiable image of the renderer's scene for these parameters."""
    return _Render.apply(texture, density, renderer)
