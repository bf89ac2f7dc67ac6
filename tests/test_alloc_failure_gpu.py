"""A render whose work buffers do not fit the device fails with RM_ERR_OOM
and leaves nothing behind: the next render on the same stream runs and gives
the same image as before the failure (a failed hipMalloc stays the runtime's
last error until it is taken, and the library takes it)."""
import numpy as np
import pytest

import raym0nade_amd as rm
from raym0nade_amd import render_check as RC
from raym0nade_amd._lib import RM_ERR_OOM, RmError

pytestmark = pytest.mark.gpu


def args_for(g, side, spp, ppr=0):
    a = rm.RmRenderArgs.from_buffer_copy(g.args)
    a.accuracy = np.float32(a.accuracy * a.width / side)
    a.width = a.height = side
    a.spp = spp
    a.paths_per_round = ppr
    return a


def test_render_after_a_failed_allocation(dev):
    g = rm.generate(0, 1)
    S = rm.Scene(g)
    a = args_for(g, 32, 4)
    gb0, rad0 = RC.gpu_render(S, a)
    # 8192^2 pixels at 16 paths per round: about 1G path slots of ~1 KB each
    big = rm.Frame(8192, 8192)
    with pytest.raises(RmError) as e:
        big.render(S, args_for(g, 8192, 64, ppr=16))
    assert e.value.code == RM_ERR_OOM
    del big
    gb1, rad1 = RC.gpu_render(S, a)
    RC.assert_gbuffer_equal(gb1, gb0)
    assert np.array_equal(rad1.view(np.uint32), rad0.view(np.uint32))
