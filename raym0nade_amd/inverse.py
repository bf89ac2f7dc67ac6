"""Inverse rendering with the path-replay adjoint: recover albedo textures,
per-material roughness and the light's emission of a scene from a target
image rendered with perturbed parameters, by Adam on the L2 image loss.

The target is the scene with every texture scaled toward mid-grey
(0.8 t + 0.1), every roughness raised by 0.15 and the emission scaled by 1.3,
rendered at 4x the spp with independent seeds; the optimisation starts from
the scene's own parameters and draws a new seed every step.
"""
from __future__ import annotations

import time

import torch

from .diff import DiffRenderer


def inverse_render(gen, steps: int = 200, args=None, lr=(0.005, 0.01, 0.3), log=None) -> dict:
    R = DiffRenderer(gen, args)
    tex0, rough0, emis0 = R.initial_params()
    emitting = emis0.abs().sum(1) > 0
    tex_t = (tex0 * 0.8 + 0.1).clamp(0.01, 1.0)
    rough_t = (rough0 + 0.15).clamp(0.05, 1.0)
    emis_t = emis0 * 1.3
    with torch.no_grad():
        ref = DiffRenderer(gen, args)
        ref.args.spp = R.args.spp * 4
        ref.args.seed = 1_000_003
        target = ref.forward_image(tex_t, rough_t, emis_t)
    tex = tex0.clone().requires_grad_()
    rough = rough0.clone().requires_grad_()
    emis = emis0.clone().requires_grad_()
    opt = torch.optim.Adam([{"params": [tex], "lr": lr[0]}, {"params": [rough], "lr": lr[1]},
                            {"params": [emis], "lr": lr[2]}])
    losses, t0 = [], time.perf_counter()

    def err():
        with torch.no_grad():
            m = emitting
            return {"texture_rmse": float(((tex - tex_t) ** 2).mean().sqrt()),
                    "roughness_rmse": float(((rough - rough_t)[~m] ** 2).mean().sqrt()),
                    "emission_rel": float(((emis - emis_t)[m].abs() / emis_t[m].abs()).mean()) if m.any() else 0.0}

    start = err()
    for step in range(steps):
        R.args.seed = step + 1
        img = R.render(tex, rough, emis)
        loss = ((img - target) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        opt.step()
        with torch.no_grad():
            tex.clamp_(0.01, 1.0)
            rough.clamp_(0.05, 1.0)
            emis[emitting] = emis[emitting].clamp(min=1e-2)
        losses.append(float(loss))
        if log and (step % 20 == 0 or step == steps - 1):
            log(f"[inverse] step {step}: loss {losses[-1]:.5f}")
    return {"steps": steps, "seconds": round(time.perf_counter() - t0, 2), "loss_first": losses[0],
            "loss_last10_mean": sum(losses[-10:]) / len(losses[-10:]), "start": start, "end": err()}
