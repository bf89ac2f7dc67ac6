"""Launch-parameter surface on the host (no device needed): rm_set_tuning /
rm_get_tuning bounds, the autotuner's snapshot and restore, and its handling
of candidates whose work buffers do not fit."""
import pytest

import raym0nade_amd as rm
from raym0nade_amd import autotune
from raym0nade_amd._lib import RM_ERR_INVALID, RM_ERR_OOM, RmError


def test_tuning_keys_round_trip_and_reject_out_of_range():
    saved = autotune.current()
    try:
        for key, good, bad in [(rm.TUNE_TRACE_VARIANT, (0, 2, 3), (-1, 1, 4, 5)),
                               (rm.TUNE_LDS_STACK_ENTRIES, (1, 32), (0, 33)),
                               (rm.TUNE_RENDER_MIN_WAVES, (1, 3, 4), (2, 8)),
                               (rm.TUNE_PERSISTENT_BLOCKS_PER_CU, (1, 64), (0, 65)),
                               (rm.TUNE_TRIS_PER_STEP, (0, 1, 2, 4), (3, 8)),
                               (rm.TUNE_NEE_SPLIT, (0, 1, 2), (-1, 3)),
                               (rm.TUNE_RENDER_COUNT, (1, 0), (-1, 2)),
                               (rm.TUNE_TRACE_UTIL, (1, 0), (-1, 2)),
                               (rm.TUNE_WALK_PHASES, (64, 32, 0), (-1, 65)),
                               (rm.TUNE_RAY_SORT, (0, 1, 2, 3, 4, 5), (-1, 6))]:
            for v in good:
                rm.set_tuning(key, v)
                assert rm.get_tuning(key) == v
            for v in bad:
                with pytest.raises(RmError) as e:
                    rm.set_tuning(key, v)
                assert e.value.code == RM_ERR_INVALID
        with pytest.raises(RmError):
            rm.set_tuning(99, 1)
    finally:
        autotune.restore(saved)


def test_current_and_restore():
    saved = autotune.current()
    assert set(saved) >= {"trace_variant", "lds_stack_entries", "render_min_waves"}
    try:
        rm.set_tuning(rm.TUNE_LDS_STACK_ENTRIES, 3)
        rm.set_tuning(rm.TUNE_TRACE_VARIANT, 0)
        autotune.restore(saved)
        assert autotune.current() == saved
    finally:
        autotune.restore(saved)


def test_tune_render_skips_candidates_that_run_out_of_memory(monkeypatch):
    """A candidate whose work buffers do not fit counts as infinitely slow;
    any other library error propagates."""
    calls = []

    def fake_batch(jobs, use_graph=True):
        k = jobs[0][1].paths_per_round
        calls.append((k, rm.get_tuning(rm.TUNE_RENDER_BLOCKS_PER_CU)))
        if k == 8:
            raise RmError(RM_ERR_OOM, "hipMalloc failed: out of memory")

    monkeypatch.setattr(autotune, "render_batch", fake_batch)
    # every trial takes "1 s" except paths per round 4, which takes less
    monkeypatch.setattr(autotune, "_time", lambda fn, reps: (fn(), 0.5 if calls[-1][0] == 4 else 1.0)[1])
    saved = autotune.current()
    try:
        a = rm.RmRenderArgs()
        a.spp = 64
        choice = autotune.tune_render([(None, a, None, 0)], blocks=(4, 8), waves=(1,), paths_per_round=(0, 4, 8),
                                      walks=((3, 8, 0),))
        assert choice["paths_per_round"] == 4
        assert any(k == 8 for k, _ in calls)
    finally:
        autotune.restore(saved)

    def broken_batch(jobs, use_graph=True):
        raise RmError(RM_ERR_INVALID, "bad job")

    monkeypatch.setattr(autotune, "render_batch", broken_batch)
    try:
        with pytest.raises(RmError):
            autotune.tune_render([(None, rm.RmRenderArgs(), None, 0)], blocks=(4,), waves=(1,),
                                 paths_per_round=(0,), walks=((3, 8, 0),))
    finally:
        autotune.restore(saved)


def test_tune_render_decides_paths_per_round_at_the_full_spp(monkeypatch):
    """At the scaled spp 4 paths per round are fastest, at the job's own spp
    8: the choice is made at the full spp (and the trials there really run
    with the job's spp)."""
    calls = []

    def fake_batch(jobs, use_graph=True):
        calls.append((jobs[0][1].paths_per_round, jobs[0][1].spp))

    def fake_time(fn, reps):
        fn()
        k, spp = calls[-1]
        if spp == 64:
            return {4: 2.0, 8: 1.5}.get(k, 3.0)
        return {4: 0.5}.get(k, 1.0)

    monkeypatch.setattr(autotune, "render_batch", fake_batch)
    monkeypatch.setattr(autotune, "_time", fake_time)
    saved = autotune.current()
    try:
        a = rm.RmRenderArgs()
        a.spp = 64
        choice = autotune.tune_render([(None, a, None, 0)], blocks=(4,), waves=(1,), paths_per_round=(0, 4, 8),
                                      walks=((3, 8, 0),), full_paths=(4, 8))
        assert choice["paths_per_round"] == 8
        assert {k for k, spp in calls if spp == 64} == {4, 8}
        assert {spp for _, spp in calls} == {8, 64}  # 64 x 0.125, and the full spp
        assert choice["full_spp_seconds"] == {"4": 2.0, "8": 1.5}
    finally:
        autotune.restore(saved)
