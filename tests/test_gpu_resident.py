"""GPU parity of the resident (on-chip, multi-step) path."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from gpuvoxels_amd.model import MT
from oracle_lib import OSim

pytestmark = pytest.mark.gpu


def test_resident_f64_cube_matches_oracle():
    m = workloads.cube(6, steps=0)
    o = OSim(m)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT, steps_per_launch=250)
    s = b.admit(m)
    for _ in range(4):
        o.step(250)
        b.step(250)
        g, c = b.read_state(s), o.state()
        assert np.max(np.abs(g["pos"] - c["pos"])) < 1e-12
        assert np.max(np.abs(g["orient"] - c["orient"])) < 1e-10
        assert np.array_equal(g["flags"] & 8, c["flags"] & 8)
    st, ost = b.poll()[s], o.status()
    assert st.steps == ost.steps and st.current_time == ost.current_time
    b.close()


def test_resident_f64_equals_stream_f64_on_robots():
    # both FP64 kernels are built without FMA contraction from the same physics
    # templates and sum the six link slots in the same order: identical results
    rs = workloads.robots(6, seed=11, grid=9, steps=0)
    bs = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    br = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT, steps_per_launch=300)
    ss = [bs.admit(r) for r in rs]
    sr = [br.admit(r) for r in rs]
    bs.step(900)
    br.step(900)
    for a, c in zip(ss, sr):
        ga, gc = bs.read_state(a), br.read_state(c)
        assert np.max(np.abs(ga["pos"] - gc["pos"])) < 1e-14
        assert np.max(np.abs(ga["orient"] - gc["orient"])) < 1e-13
        assert np.array_equal(ga["flags"], gc["flags"])
    pa, pc = bs.poll(), br.poll()
    for a, c in zip(ss, sr):
        assert pa[a].steps == pc[c].steps == 900
    bs.close()
    br.close()


@pytest.mark.parametrize("block", [512, 256])
def test_resident_f64_shapes_equal_stream_on_full_robots(block):
    # 730-908-voxel robots: the (256, 4, 2) and (512, 2, 1) FP64 shapes
    rs = workloads.robots(3, seed=5, steps=0)
    bs = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    br = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT, steps_per_launch=250, block_size=block)
    ss = [bs.admit(r) for r in rs]
    sr = [br.admit(r) for r in rs]
    bs.step(500)
    br.step(500)
    for a, c in zip(ss, sr):
        ga, gc = bs.read_state(a), br.read_state(c)
        assert np.max(np.abs(ga["pos"] - gc["pos"])) < 1e-14
        assert np.array_equal(ga["flags"], gc["flags"])
    bs.close()
    br.close()


@pytest.mark.parametrize("block", [1024, 512, 256])
def test_resident_f32_cube_within_tolerance(block):
    m = workloads.cube(6, steps=0)
    o = OSim(m)
    b = engine.Batch(precision=_abi.PREC_F32, path=_abi.PATH_RESIDENT, steps_per_launch=500, block_size=block)
    s = b.admit(m)
    o.step(1000)
    b.step(1000)
    g, c = b.read_state(s), o.state()
    # FP32 positions within 1e-4 of the 60 mm body scale of the FP64 oracle
    assert np.max(np.abs(g["pos"] - c["pos"])) / 0.06 < 1e-4
    assert np.max(np.abs(g["orient"] - c["orient"])) < 1e-3
    b.close()


def test_resident_chunking_is_seamless_f64():
    m = workloads.cube(5, steps=0)
    b1 = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT, steps_per_launch=1000)
    b2 = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT, steps_per_launch=77)
    s1, s2 = b1.admit(m), b2.admit(m)
    b1.step(700)
    b2.step(700)
    assert np.array_equal(b1.read_state(s1)["pos"], b2.read_state(s2)["pos"])
    b1.close()
    b2.close()


def test_resident_stop_condition_and_budget_match_oracle():
    m = workloads.cube(4, steps=0)
    # stop when t - 0.01 > 0 (reversed-BFS tokens of <mtSUB><mtVAR>t</mtVAR><mtCONST>0.01</mtCONST></mtSUB>)
    m.stop_tokens = [(MT["mtCONST"], 0.01), (MT["mtVAR"], 4.0), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)]
    o = OSim(m)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT)
    s = b.admit(m)
    o.step(5000)
    b.step(5000)
    st, ost = b.poll()[s], o.status()
    assert ost.state == _abi.SIM_DONE and st.state == _abi.SIM_DONE
    assert st.steps == ost.steps
    m2 = workloads.cube(4, steps=123)
    b2 = engine.Batch(precision=_abi.PREC_F32, path=_abi.PATH_RESIDENT)
    s2 = b2.admit(m2)
    b2.step(1000)
    assert b2.poll()[s2].steps == 123
    b.close()
    b2.close()


def test_resident_f32_equals_stream_f32_on_robots():
    # the FP32 kernels are built without contraction too: identical results on either path
    rs = workloads.robots(4, seed=12, grid=9, steps=0)
    bs = engine.Batch(precision=_abi.PREC_F32, path=_abi.PATH_STREAM)
    br = engine.Batch(precision=_abi.PREC_F32, path=_abi.PATH_RESIDENT, steps_per_launch=300)
    ss = [bs.admit(r) for r in rs]
    sr = [br.admit(r) for r in rs]
    bs.step(900)
    br.step(900)
    for a, c in zip(ss, sr):
        ga, gc = bs.read_state(a), br.read_state(c)
        assert np.array_equal(ga["pos"], gc["pos"])
        assert np.array_equal(ga["orient"], gc["orient"])
        assert np.array_equal(ga["flags"], gc["flags"])
    bs.close()
    br.close()
