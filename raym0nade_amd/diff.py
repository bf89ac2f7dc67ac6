"""Differentiable rendering: the linear Full image of rm_render as a function of
albedo texels, per-material roughness and per-material emission, with
gradients from the path-replay adjoint (rm_adjoint).

    R = DiffRenderer(raym0nade_amd.generate(4, seed))
    tex, rough, emis = R.initial_params()          # cuda tensors
    img = R.render(tex.requires_grad_(), rough, emis)
    loss = ((img - target) ** 2).sum(); loss.backward()

Every forward call uploads the parameters (rm_scene_update), renders and
shades without the tone curve; backward replays the same samples.  One
renderer holds one frame, so each forward must be followed by its backward
before the next forward (the adjoint checks this and raises otherwise).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import Frame, GeneratedScene, Scene
from ._lib import (RM_BUF_PIXELS, RM_SHADE_FULL, RM_SHADE_LINEAR, RmGrads, RmMaterial, RmRenderArgs, check, lib)


def _p(t: torch.Tensor | None):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class DiffRenderer:
    def __init__(self, gen: GeneratedScene, args: RmRenderArgs | None = None, record: bool = True):
        self.gen = gen
        self.scene = Scene(gen)
        self.args = RmRenderArgs.from_buffer_copy(args if args is not None else gen.args)
        self.frame = Frame(self.args.width, self.args.height)
        self.frame.set_recording(record)
        d = gen.desc
        self.n_materials = d.n_materials
        self.materials = [RmMaterial.from_buffer_copy(d.materials[i]) for i in range(d.n_materials)]
        self.tex_shapes = [(d.textures[t].height, d.textures[t].width) for t in range(d.n_textures)]
        self._tex0 = [np.ctypeslib.as_array(d.textures[t].rgb, shape=(h * w * 3,)).copy()
                      for t, (h, w) in enumerate(self.tex_shapes)]

    # -- parameters
    def initial_params(self, device="cuda"):
        tex = torch.from_numpy(np.concatenate(self._tex0) if self._tex0 else np.zeros(0, np.float32)).to(device)
        rough = torch.tensor([m.roughness for m in self.materials], dtype=torch.float32, device=device)
        emis = torch.tensor([list(m.emission) for m in self.materials], dtype=torch.float32, device=device)
        return tex, rough, emis

    def texture_views(self, tex: torch.Tensor):
        """Per-texture [h, w, 3] views of a flat texel tensor."""
        out, off = [], 0
        for h, w in self.tex_shapes:
            out.append(tex[off:off + h * w * 3].view(h, w, 3))
            off += h * w * 3
        return out

    def _upload(self, tex, rough, emis, stream):
        r = rough.detach().float().cpu().numpy()
        e = emis.detach().float().cpu().numpy()
        arr = (RmMaterial * self.n_materials)()
        for i, m in enumerate(self.materials):
            arr[i] = m
            arr[i].roughness = float(r[i])
            for c in range(3):
                arr[i].emission[c] = float(e[i, c])
        tex = tex.detach().contiguous().float()
        check(lib().rm_scene_update(self.scene.handle, arr, _p(tex) if tex.numel() else None, stream))

    def forward_image(self, tex, rough, emis) -> torch.Tensor:
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        self._upload(tex, rough, emis, st)
        self.frame.render(self.scene, self.args)
        self.frame.postprocess(RM_SHADE_FULL | RM_SHADE_LINEAR, self.args.exposure)
        img = torch.empty((self.args.height, self.args.width, 3), dtype=torch.float32, device=rough.device)
        check(lib().rm_frame_copy(self.frame._h, RM_BUF_PIXELS, _p(img), st))
        torch.cuda.synchronize()
        return img

    def adjoint(self, d_image: torch.Tensor, n_tex: int, replay: bool = False):
        """Gradients for d loss / d image (cuda [h, w, 3]) of the last forward."""
        dev = d_image.device
        g_tex = torch.empty(n_tex, dtype=torch.float32, device=dev)
        g_r = torch.empty(self.n_materials, dtype=torch.float32, device=dev)
        g_e = torch.empty((self.n_materials, 3), dtype=torch.float32, device=dev)
        rep = torch.empty((self.args.height, self.args.width, 3), dtype=torch.float32, device=dev) if replay else None
        g = RmGrads(_p(g_tex) if n_tex else None, _p(g_r), _p(g_e), _p(rep))
        d_image = d_image.detach().contiguous().float()
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(lib().rm_adjoint(self.scene.handle, C.byref(self.args), self.frame._h, _p(d_image), C.byref(g), st))
        return g_tex, g_r, g_e, rep

    def render(self, tex, rough, emis) -> torch.Tensor:
        return _RenderFn.apply(self, tex, rough, emis)


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, R: DiffRenderer, tex, rough, emis):
        ctx.R = R
        ctx.n_tex = tex.numel()
        return R.forward_image(tex, rough, emis)

    @staticmethod
    def backward(ctx, grad):
        g_tex, g_r, g_e, _ = ctx.R.adjoint(grad, ctx.n_tex)
        return None, g_tex, g_r, g_e
