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
