"""VX3 cilia and force fields on the GPU engine against the oracle
(VX3_VoxelyzeKernel.cu:296-303,846-859; VX3_Voxel.cu:178-184,376): FP64 states within
1e-9 m after 1000 steps, with and without collisions, from a Model and from a VXA."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from gpuvoxels_amd.model import MT
from oracle_lib import OSim
from vxa_cases import CASES

pytestmark = pytest.mark.gpu


def ciliated(m, seed=0, on_after=(0.0, 0.0, 0.005, 0.0), scale=(0.0, 2.5, 8.0, 1.0), field=True):
    rng = np.random.default_rng(seed)
    m.enable_cilia = 1
    m.base_cilia = rng.normal(0, 2e-4, (m.n_voxels, 3))
    m.shift_cilia = rng.normal(0, 2e-4, (m.n_voxels, 3))
    for k, mt in enumerate(m.materials):
        mt.cilia = scale[k % len(scale)]
        mt.cilia_on_after = on_after[k % len(on_after)]
    if field:   # a pull toward x = 0 growing with time, and a constant lift
        m.force_field = [[(MT["mtVAR"], 0.0), (MT["mtVAR"], 4.0), (MT["mtMUL"], 0.0), (MT["mtCONST"], -0.5),
                          (MT["mtMUL"], 0.0), (MT["mtEND"], 0.0)], [],
                         [(MT["mtCONST"], 2e-3), (MT["mtEND"], 0.0)]]
    return m


def compare(models, steps=1000, tol=1e-9, precision=_abi.PREC_F64):
    b = engine.Batch(device=0, precision=precision, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in models]
    b.step(steps)
    worst = 0.0
    for m, s in zip(models, slots):
        o = OSim(m)
        o.step(steps)
        g, r = b.read_state(s), o.state()
        worst = max(worst, float(np.max(np.abs(g["pos"] - r["pos"]))))
        assert np.max(np.abs(g["orient"] - r["orient"])) < 10 * tol
    b.close()
    print("max |GPU - oracle| position: %.3e m" % worst)
    assert worst < tol
    return worst


def test_ciliated_robots_match_oracle():
    models = [ciliated(m, seed=i) for i, m in enumerate(workloads.robots(3, seed=8, grid=8, steps=0))]
    compare(models)


def test_cilia_only_and_field_only():
    a = ciliated(workloads.robots(1, seed=9, grid=8, steps=0)[0], field=False)
    b = workloads.robots(1, seed=10, grid=8, steps=0)[0]
    b.force_field = [[], [(MT["mtVAR"], 2.0), (MT["mtCONST"], 0.3), (MT["mtMUL"], 0.0), (MT["mtEND"], 0.0)], []]
    compare([a, b])


def test_ciliated_drop_with_collisions():
    m = ciliated(workloads.drop(3, gap=1, shift=1, steps=0), scale=(3.0,), on_after=(0.0,))
    compare([m], steps=2000)


def test_mixed_with_plain_and_the_effect_is_real():
    plain = workloads.robots(2, seed=11, grid=8, steps=0)
    cil = ciliated(workloads.robots(1, seed=11, grid=8, steps=0)[0])
    compare(plain + [cil])
    # the same robot without extras ends elsewhere
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s0, s1 = b.admit(plain[0]), b.admit(cil)
    b.step(1000)
    assert np.max(np.abs(b.read_state(s0)["pos"] - b.read_state(s1)["pos"])) > 1e-6
    b.close()


def test_from_vxa_document():
    v = engine.Vxa(CASES["cilia"]())
    assert v.desc.enable_cilia == 1 and v.desc.n_force_field_tokens[0] > 0
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s = b.admit(v)
    b.step(1000)
    o = OSim.from_desc(v.desc)
    o.step(1000)
    assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-9
    b.close()


def test_f32_within_tolerance():
    # a cube, as the FP32 bar is set on (robots' soft contacts amplify FP32 rounding)
    m = ciliated(workloads.cube(4, steps=0), scale=(2.0,), on_after=(0.0,))
    compare([m], tol=1e-4, precision=_abi.PREC_F32)


def test_resident_path_refuses():
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT)
    with pytest.raises(engine.VX3Error) as e:
        b.admit(ciliated(workloads.robots(1, seed=9, grid=8, steps=0)[0]))
    assert e.value.status == _abi.EUNSUPPORTED
    b.close()
