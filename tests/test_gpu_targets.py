"""VX3 target contacts on the GPU engine against the oracle
(handle_collision_attachment, VX3_VoxelyzeKernel.cu:682-722): hit counts equal, the
MathTree variable hit in the fitness, signals started by target contacts, FP64
positions within 1e-9 m."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine
from gpuvoxels_amd.model import MT
from oracle_lib import OSim
from test_gpu_extras import ciliated
from test_oracle_kat import _target_drop

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("signals", [0, 1])
def test_target_drop_matches_oracle(signals):
    m = _target_drop()
    m.fitness_tokens = [(MT["mtVAR"], 3.0), (MT["mtEND"], 0.0)]
    if signals:
        m.enable_signals = 1
        m = ciliated(m, seed=3, scale=(4.0, 0.0), on_after=(0.0,), field=False)
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s = b.admit(m)
    o = OSim(m)
    for _ in range(6):
        b.step(500)
        o.step(500)
        st, so = b.poll()[s], o.status()
        assert st.collision_count == so.collision_count
        assert st.fitness == so.fitness
        assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-9
    assert so.collision_count > 0
    if signals:
        assert o.local_signal().max() > 0
    b.close()


def test_close_pairs_and_closeness_match_oracle():
    """computeTargetCloseness on the GPU: the counted pairs, the 1/distance sum and the
    stop program numClosePairs > 0 as in the oracle (closed form in test_oracle_kat)."""
    from test_oracle_kat import _close_targets
    m = _close_targets()
    m.stop_tokens = [(MT["mtCONST"], 0.0), (MT["mtVAR"], 7.0), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)]
    m.fitness_tokens = [(MT["mtVAR"], 6.0), (MT["mtEND"], 0.0)]   # fitness = targetCloseness
    o = OSim(m)
    o.step(100)
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s = b.admit(m)
    b.step(100)
    st, so = b.poll()[s], o.status()
    assert st.state == so.state == _abi.SIM_DONE and st.steps == so.steps == 10
    assert st.num_close_pairs == so.num_close_pairs == 1
    assert st.target_closeness == pytest.approx(so.target_closeness, rel=1e-14)
    assert st.fitness == pytest.approx(so.fitness, rel=1e-14)
    b.close()


def test_close_pairs_of_moving_robots_match_oracle():
    """Robots whose actuated material is a target: close pairs and closeness of the
    voxels' positions at the center-of-mass samples, taken once per temperature period
    (0.1 s here, about 11,000 steps of these robots: two samples in 24,000 steps)."""
    from gpuvoxels_amd import workloads
    rs = workloads.robots(3, seed=21, grid=7, steps=0)
    for r in rs:
        r.materials[2].is_target = 1
        r.max_dist_close_pairs = 1.5
        r.temp_period = 0.1
    os_ = [OSim(r) for r in rs]
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    slots = [b.admit(r) for r in rs]
    for o in os_:
        o.step(24000)
    b.step(24000)
    st = b.poll()
    for o, s in zip(os_, slots):
        so, g = o.status(), st[s]
        assert g.num_close_pairs == so.num_close_pairs > 0
        assert g.target_closeness == pytest.approx(so.target_closeness, rel=1e-9)
    b.close()
