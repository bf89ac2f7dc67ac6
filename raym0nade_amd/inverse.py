"""Inverse rendering with the path-replay adjoint: recover albedo textures,
per-material roughness and the light's emission of a scene from a target
image rendered with perturbed parameters, by Adam on the L2 image loss.

The target is the scene with every texture scaled toward mid-grey
(0.8 t + 0.1), every roughness raised by 0.15 and the emission scaled by 1.3,
rendered at 4x the spp with independent seeds; the optimisation starts from
the scene's own parameters and draws a new seed every step.

decorrelate: the gradient of the L2 loss is 2/N sum (I - T) dI/dtheta.  With
the residual I - T and the replayed derivative taken from the same samples,
its expectation has an extra covariance term (2/N sum Cov(I, dI/dtheta)) that
pulls the parameters toward less noisy images rather than toward the target.
With decorrelate, each step renders once for the residual and once more,
with an independent seed, for the derivative (the surrogate
sum(stop_grad(2/N (I_a - T)) * I_b)), which makes the gradient unbiased for a
cost of one extra forward render per step.
"""
from __future__ import annotations

import time

import torch

from .diff import DiffRenderer


def inverse_render(gen, steps: int = 200, args=None, lr=(0.001, 0.005, 0.1), decorrelate: bool = True,
                   log=None) -> dict:
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
        opt.zero_grad()
        if decorrelate:
            with torch.no_grad():
                R.args.seed = 2 * step + 1
                resid = R.forward_image(tex, rough, emis) - target
                loss = (resid ** 2).mean()
            R.args.seed = 2 * step + 2
            img = R.render(tex, rough, emis)
            (img * (resid * (2.0 / resid.numel()))).sum().backward()
        else:
            R.args.seed = step + 1
            img = R.render(tex, rough, emis)
            loss = ((img - target) ** 2).mean()
            loss.backward()
        opt.step()
        with torch.no_grad():
            tex.clamp_(0.01, 1.0)
            rough.clamp_(0.05, 1.0)
            emis[emitting] = emis[emitting].clamp(min=1e-2)
        losses.append(float(loss))
        if log and (step % 20 == 0 or step == steps - 1):
            log(f"[inverse] step {step}: loss {losses[-1]:.5f}")
    return {"steps": steps, "decorrelate": decorrelate, "lr": list(lr),
            "seconds": round(time.perf_counter() - t0, 2), "loss_first": losses[0],
            "loss_last10_mean": sum(losses[-10:]) / len(losses[-10:]), "start": start, "end": err()}
