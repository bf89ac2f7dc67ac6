"""The batched serving workload: independent render jobs, each a seeded
variant of the Cornell box (36 triangles) or of the displaced-sphere scene
(grid^2 geodesic spheres of 20 freq^2 triangles under an HDR sky, up to about
2M triangles), at a log-uniform square resolution between 128^2 and 1024^2,
with spp in {16, 64, 256} and max depth in 8..16.  Job i is drawn from a
numpy Generator seeded with (seed, i), so the set is reproducible."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import GeneratedScene, RmRenderArgs, Scene, generate


@dataclass
class JobSpec:
    config: int          # 0 Cornell box, 2 displaced spheres
    scene_seed: int
    grid: int            # config 2: spheres per side
    freq: int            # config 2: geodesic frequency (20 freq^2 triangles per sphere)
    side: int            # square resolution
    spp: int
    max_depth: int

    @property
    def samples(self) -> int:
        return self.side * self.side * self.spp


def batch_specs(n: int = 64, seed: int = 1, min_log2: float = 7.0, max_log2: float = 10.0) -> list[JobSpec]:
    out = []
    for i in range(n):
        r = np.random.default_rng([seed, i])
        cfg = 0 if r.random() < 0.5 else 2
        grid = int(r.integers(1, 9)) if cfg == 2 else 0
        freq = int(r.integers(4, 41)) if cfg == 2 else 0
        side = int(round(2.0 ** r.uniform(min_log2, max_log2)))
        spp = int(r.choice([16, 64, 256]))
        depth = int(r.integers(8, 17))
        out.append(JobSpec(cfg, seed * 1000 + i, grid, freq, side, spp, depth))
    return out


def build_job(spec: JobSpec, sky: tuple[int, int] = (512, 256)):
    """(GeneratedScene, Scene, RmRenderArgs) of one job."""
    if spec.config == 0:
        g = generate(0, spec.scene_seed)
    else:
        g = generate(2, spec.scene_seed, grid=spec.grid, freq=spec.freq, sky_width=sky[0], sky_height=sky[1])
    S = Scene(g)
    a = RmRenderArgs.from_buffer_copy(g.args)
    a.accuracy = np.float32(a.accuracy * a.width / spec.side)
    a.width = a.height = spec.side
    a.spp = spec.spp
    a.max_depth = spec.max_depth
    a.paths_per_round = 0
    return g, S, a
