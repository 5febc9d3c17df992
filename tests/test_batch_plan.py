"""Host-side consistency of the batched plans (no device): every index the window
kernels derive from a plan's tables stays within what it indexes, for the
autotuner's candidate plans of the 5000-cell benchmark model and for models with
cells beyond each layout's LDS capacity."""
import pytest

from mind_amd import configs
from mind_amd.model import check_batch_plan

# the autotuner's candidates (batched.hip, batched_setup) plus small block sizes
CANDIDATES = ([(b, t, 0) for t in (256, 512) for b in (1024, 1700, 2750)]
              + [(b, t, 5) for t in (256, 512) for b in (1800, 2400, 3000, 3600, 4800, 6000)]
              + [(b, t, 4) for t in (256, 512) for b in (3000, 3600, 4500, 5400, 7200)]
              + [(b, 1024, 5) for b in (4800, 6000)]
              + [(b, 1024, 4) for b in (7200, 9000)]
              + [(300, 64, 0), (600, 64, 5), (900, 128, 4), (2750, 1024, 0), (7200, 1024, 4)])


@pytest.fixture(scope="module")
def bench_model():
    model, _ = configs.tree_network(ncells=5000)
    return model


@pytest.mark.parametrize("plan", CANDIDATES)
def test_benchmark_model_plans_are_consistent(bench_model, plan):
    groups = check_batch_plan(bench_model, *plan)
    assert groups > 0


@pytest.mark.parametrize("plan", [(1, 256, 0), (4000, 256, 4), (2000, 512, 5), (100000, 512, 5)])
def test_oversized_cells_plans_are_consistent(plan):
    model, _ = configs.tree_network(ncells=4, median=6000, sigma=0.8, lo=2000, hi=11500,
                                    indegree=3, synapses=4, rate_hz=50.0, netstim_weight=3e-2)
    assert check_batch_plan(model, *plan) >= 4


@pytest.mark.parametrize("plan", [(300, 128, 0), (600, 64, 4), (500, 64, 5), (3000, 1024, 5)])
def test_many_children_plans_are_consistent(plan):
    model, _ = configs.tree_network(ncells=30, median=120, lo=20, hi=600, indegree=10, fanout=4)
    assert check_batch_plan(model, *plan) > 0


def test_invalid_plans_are_rejected():
    from mind_amd._abi import MindError

    model, _ = configs.single_cell()
    for plan in [(0, 256, 0), (100, 2048, 0), (100, 96, 0), (100, 256, 3)]:
        with pytest.raises(MindError):
            check_batch_plan(model, *plan)
