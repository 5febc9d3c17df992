"""Micro-macro coupling on the GPU against the oracle (bit-exact traces and rasters)."""
import numpy as np
import pytest

import oracle
from mind_amd import configs, macro as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["plain", "batched"])
def test_micro_macro_matches_oracle(gpu, method):
    mm, info = configs.micro_macro(n_roi=4, cells=8, nseg=40, input_weight=5e-3)
    print(f"case: micro-macro {method} 4 ROIs x 8 cells, 100 ms", flush=True)
    M.build(mm, method=method)
    M.initialize(mm)
    got = M.run(mm, 100.0)
    ref = oracle.OracleMacro(mm).initialize().run(100.0)
    for name in ("E", "I", "rate", "coupling", "mean_v"):
        assert np.array_equal(getattr(got, name), getattr(ref, name)), name
    assert np.array_equal(got.spike_times, ref.spike_times)
    assert np.array_equal(got.spike_sources, ref.spike_sources)
    assert (got.spike_sources >= 0).sum() > 0 and (got.spike_sources < 0).sum() > 0
    # a second run continues from the same state
    got2 = M.run(mm, 150.0)
    ref_sim = oracle.OracleMacro(mm).initialize()
    ref_sim.run(100.0)
    ref2 = ref_sim.run(150.0)
    assert np.array_equal(got2.E, ref2.E)
    assert np.array_equal(got2.spike_times, ref2.spike_times)
