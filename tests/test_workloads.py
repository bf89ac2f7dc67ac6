"""The batched-serving job set: reproducible, and drawn from the stated
ranges (Cornell box or displaced spheres up to ~2M triangles, 128^2..1024^2,
spp in {16, 64, 256}, depth 8..16)."""
from raym0nade_amd import workloads


def test_specs_reproducible_and_in_range():
    a = workloads.batch_specs(64, seed=1)
    b = workloads.batch_specs(64, seed=1)
    assert a == b
    assert workloads.batch_specs(64, seed=2) != a
    configs = {s.config for s in a}
    assert configs == {0, 2}
    for s in a:
        assert 128 <= s.side <= 1024
        assert s.spp in (16, 64, 256)
        assert 8 <= s.max_depth <= 16
        if s.config == 2:
            tris = s.grid * s.grid * 20 * s.freq * s.freq
            assert 1 <= s.grid <= 8 and 4 <= s.freq <= 40 and tris <= 2_100_000
    sides = sorted(s.side for s in a)
    assert sides[0] < 200 and sides[-1] > 700  # log-uniform spread over the range


def test_generated_job_matches_spec():
    s = workloads.JobSpec(config=2, scene_seed=7, grid=2, freq=6, side=160, spp=16, max_depth=9)
    from raym0nade_amd import generate
    g = generate(2, 7, grid=2, freq=6, sky_width=64, sky_height=32)
    assert g.n_tris == 2 * 2 * 20 * 6 * 6 + 2
    assert s.samples == 160 * 160 * 16
