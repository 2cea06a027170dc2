"""The autotuner (vx3h_autotune): it measures every candidate configuration, returns the
fastest, and the configurations it chooses between give identical results."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads

pytestmark = pytest.mark.gpu


def test_autotune_picks_the_fastest_measured_candidate():
    models = workloads.robots(24, seed=11, grid=10, steps=0)
    cfg, trials = engine.autotune(models, precision=_abi.PREC_F64, warmup_steps=20, trial_steps=200)
    assert len(trials) == 16            # block sizes {0, 128, 256, 512} x steps per launch {250, 500, 1000, 2000}
    assert all(v > 0 for *_, v in trials)
    best = max(trials, key=lambda t: t[3])
    assert (cfg["block_size"], cfg["steps_per_launch"]) == (best[0], best[1])
    assert cfg["collision_rebuild_interval"] == 0    # not tuned unless asked: it changes results


@pytest.mark.parametrize("precision", [_abi.PREC_F64, _abi.PREC_F32])
def test_tuned_configurations_give_identical_results(precision):
    """Shape and launch length decide which threads compute a link or voxel and when a
    launch ends; neither changes any value."""
    models = workloads.robots(6, seed=12, grid=10, steps=0)
    states = []
    for block, spl in ((0, 1000), (256, 250), (512, 333), (128, 2000)):
        if precision == _abi.PREC_F32 and block == 128:
            block = 1024
        b = engine.Batch(precision=precision, path=_abi.PATH_AUTO, block_size=block, steps_per_launch=spl)
        slots = [b.admit(m) for m in models]
        b.step(700)
        states.append([b.read_state(s)["pos"] for s in slots])
        b.close()
    for other in states[1:]:
        for a, c in zip(states[0], other):
            assert np.array_equal(a, c)
