"""Every optional feature at once in one batch beside plain resident simulations:
signals with pacemakers, cilia and a force field, detach of fragile links, a
secondary experiment, a VoxCAD stop condition -- each simulation against the oracle
(FP64 positions within 1e-9 m, same stop step), the plain ones unchanged."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from gpuvoxels_amd.model import MT
from oracle_lib import OSim
from test_gpu_extras import ciliated

pytestmark = pytest.mark.gpu


def everything(seed):
    m = workloads.robot(np.random.default_rng(seed), grid=7, p=0.8, steps=0)
    for k, mt in enumerate(m.materials):
        mt.mat_model, mt.fail_stress = _abi.MDL_LINEAR_FAIL, 800.0
    m.materials[2].is_pacemaker, m.materials[2].pacemaker_period = 1, 0.03
    m.materials[1].remove_after = 0.008
    m.enable_signals = 1
    m.enable_detach = 1
    m.secondary_experiment, m.reinit_after = 1, 0.004
    m.stop_condition_type, m.stop_condition_value = _abi.SC_MAX_TIME_STEPS, 1800.0
    m = ciliated(m, seed=seed)
    m.fitness_tokens = [(MT["mtVAR"], 0.0), (MT["mtEND"], 0.0)]
    return m


def test_all_features_together():
    special = [everything(s) for s in (1, 2)]
    plain = workloads.robots(2, seed=6, grid=8, steps=2500)
    models = plain[:1] + special + plain[1:]
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in models]
    b.step(2500)
    st = b.poll()
    for m, s in zip(models, slots):
        o = OSim(m)
        o.step(2500)
        so = o.status()
        assert st[s].steps == so.steps and st[s].state == so.state
        assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-9
        assert st[s].fitness == pytest.approx(so.fitness, rel=1e-9, abs=1e-12)
    assert st[slots[1]].steps == 1801                       # the VoxCAD condition stopped them
    b.close()
