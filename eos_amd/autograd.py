"""torch.autograd.Function around eos_render / eos_render_adjoint (test harness).

Parameters are torch tensors on any device:
  kd, ks      [M, 3]  diffuse and specular colours
  shininess   [M]     Phong exponents (used as floats)
  light_pos   [L, 3]  point-light positions
  light_color [L, 3]  light colours
The output is the pre-quantisation image [H, W, 3] (float32, on cuda:0).
Gradients come from the hand-written HIP adjoint with hit sets held fixed
(visibility, hit points and reflection paths of the forward pass).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as E


def _materials_from(base, kd, ks, shininess):
    m = base.copy()
    m["kd"][:, :3] = kd.detach().to("cpu", torch.float32).numpy()
    m["ks"][:, :3] = ks.detach().to("cpu", torch.float32).numpy()
    m["shininess"] = shininess.detach().to("cpu", torch.float32).numpy()
    return m


def _lights_from(base, pos, col):
    lt = base.copy()
    lt["position"][:, :3] = pos.detach().to("cpu", torch.float32).numpy()
    lt["position"][:, 3] = 1.0
    lt["color"][:, :3] = col.detach().to("cpu", torch.float32).numpy()
    return lt


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kd, ks, shininess, light_pos, light_color, scene: "E.Scene", camera, depth: int):
        scene.set_materials(_materials_from(scene.data.materials, kd, ks, shininess))
        scene.set_lights(_lights_from(scene.data.lights, light_pos, light_color))
        W, H = int(camera["width"]), int(camera["height"])
        rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        scene.render_device(camera, depth, None, rgb_dev_ptr=rgb.data_ptr(), keep_records=True)
        ctx.scene, ctx.camera, ctx.depth = scene, camera, depth
        ctx.shapes = (kd.shape[0], light_pos.shape[0])
        ctx.devices = (kd.device, light_pos.device)
        return rgb

    @staticmethod
    def backward(ctx, grad_rgb):
        scene, camera, depth = ctx.scene, ctx.camera, ctx.depth
        M, L = ctx.shapes
        g = grad_rgb.detach().to("cuda:0", torch.float32).contiguous()
        gm = torch.zeros((M, 7), dtype=torch.float64, device="cuda:0")
        gl = torch.zeros((max(L, 1), 6), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        scene.adjoint_device(camera, depth, g.data_ptr(), gm.data_ptr(), gl.data_ptr())
        gm = gm.to(ctx.devices[0], torch.float32)
        gl = gl[:L].to(ctx.devices[1], torch.float32)
        return gm[:, 0:3], gm[:, 3:6], gm[:, 6], gl[:, 0:3], gl[:, 3:6], None, None, None


def render(scene: "E.Scene", kd, ks, shininess, light_pos, light_color, camera=None, depth=None):
    camera = np.array(scene.data.camera if camera is None else camera, copy=True)
    depth = scene.data.depth if depth is None else depth
    return RenderFunction.apply(kd, ks, shininess, light_pos, light_color, scene, camera, depth)


def params_of(scene_data, device="cuda:0"):
    """(kd, ks, shininess, light_pos, light_color) tensors of a SceneData."""
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device=device)  # noqa: E731
    return (t(scene_data.materials["kd"][:, :3]), t(scene_data.materials["ks"][:, :3]),
            t(scene_data.materials["shininess"]), t(scene_data.lights["position"][:, :3]),
            t(scene_data.lights["color"][:, :3]))
