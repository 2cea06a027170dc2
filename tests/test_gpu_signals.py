"""VX3 signals on the GPU engine against the oracle (VX3_Voxel.cu:271-348, two-phase
order: every moving voxel updates its own state and posts what it sends, then every
voxel takes its neighbours' posts in link-direction order): local signals equal,
FP64 positions within 1e-9 m."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim
from test_gpu_extras import ciliated
from test_oracle_kat import _chain

pytestmark = pytest.mark.gpu


def test_chain_signal_equals_oracle():
    m = _chain()
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s = b.admit(m)
    o = OSim(m)
    for _ in range(8):
        b.step(500)
        o.step(500)
        rec = b.history_record(s, 0, 1, 0)
        got = np.array([float(e.split(",")[14]) for e in rec.split(">>>", 1)[1].split("<<<>>>")[0].split(";")[:-1]])
        want = np.round(o.local_signal(), 1)
        assert np.allclose(got, want, atol=0.051)          # printed with one decimal
    assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-12
    b.close()


def test_ciliated_robots_with_pacemakers():
    models = []
    for i, m in enumerate(workloads.robots(3, seed=41, grid=8, steps=0)):
        m.enable_signals = 1
        m.materials[2].is_pacemaker, m.materials[2].pacemaker_period = 1, 0.02 * (i + 1)
        models.append(ciliated(m, seed=i, field=False))
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in models]
    b.step(2000)
    for m, s in zip(models, slots):
        o = OSim(m)
        o.step(2000)
        assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-9
        assert o.local_signal().max() > 0
    b.close()
