"""torch.autograd.Function around eos_render / eos_render_adjoint (test harness).

Parameters are torch tensors on any device:
  kd, ks      [M, 3]  diffuse and specular colours
  shininess   [M]     Phong exponents (used as floats)
  light_pos   [L, 3]  point-light positions
  light_color [L, 3]  light colours
The output is the pre-quantisation image [H, W, 3] (float32, on cuda:0).
Gradients come from the hand-written HIP adjoint with hit sets held fixed
(visibility, hit points and reflection paths of the forward pass).

Each forward keeps its own materials and lights.  The scene holds one set of
hit records, so a backward whose scene was rendered or changed after its
forward (two renders in one loss, a parameter update in between) uploads its
forward's parameters again and re-traces the records instead of replaying
another call's, then puts the scene's parameters back.
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
        mats = _materials_from(scene.data.materials, kd, ks, shininess)
        lights = _lights_from(scene.data.lights, light_pos, light_color)
        scene.set_materials(mats)
        scene.set_lights(lights)
        W, H = int(camera["width"]), int(camera["height"])
        rgb = torch.empty((H, W, 3), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        scene.render_device(camera, depth, None, rgb_dev_ptr=rgb.data_ptr(), keep_records=True)
        ctx.scene, ctx.camera, ctx.depth = scene, camera, depth
        ctx.mats, ctx.lights = mats, lights
        ctx.generation = scene.generation   # the records this forward left on the scene
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
        stale = scene.generation != ctx.generation
        if stale:
            current = (scene._mats.copy(), scene._lights.copy())
            scene.set_materials(ctx.mats)
            scene.set_lights(ctx.lights)
        try:
            scene.adjoint_device(camera, depth, g.data_ptr(), gm.data_ptr(), gl.data_ptr(), recompute=stale)
        finally:
            if stale:
                scene.set_materials(current[0])
                scene.set_lights(current[1])
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
