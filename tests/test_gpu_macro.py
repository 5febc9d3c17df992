"""Micro-macro coupling on the GPU against the oracle (bit-exact traces and rasters)."""
import numpy as np
import pytest

import oracle
from mind_amd import configs, macro as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("window", ["bin", "step"])
@pytest.mark.parametrize("method", ["plain", "batched"])
def test_micro_macro_matches_oracle(gpu, method, window):
    """Windows of one 1 ms bin (ten macro steps) or of one macro step: every macro
    trace and the raster bit for bit, over a run that stops inside a bin and two
    that continue from it."""
    mm, info = configs.micro_macro(n_roi=4, cells=8, nseg=40, input_weight=5e-3)
    print(f"case: micro-macro {method} {window} windows, 4 ROIs x 8 cells, 100 ms", flush=True)
    M.build(mm, method=method, window_steps=0 if window == "bin" else 4)
    M.initialize(mm)
    ref_sim = oracle.OracleMacro(mm).initialize()
    for stop in (100.0, 100.3, 150.0):  # 100.3 ends inside a bin
        got = M.run(mm, stop)
        ref = ref_sim.run(stop)
        for name in ("E", "I", "rate", "coupling", "mean_v"):
            assert np.array_equal(getattr(got, name), getattr(ref, name)), (stop, name)
        assert np.array_equal(got.spike_times, ref.spike_times), stop
        assert np.array_equal(got.spike_sources, ref.spike_sources), stop
    assert (got.spike_sources >= 0).sum() > 0 and (got.spike_sources < 0).sum() > 0
