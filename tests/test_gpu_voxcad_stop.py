"""VoxCAD's stop-condition family and equilibrium mode on the GPU engine against the
oracle (VoxCAD/src/VX1/VX_Sim.cpp:862-891, 1015-1160, 1833-1837): the same stop step
and final state, FP64 positions within 1e-9 m.  The engine sums the energies in a
block reduction and the oracle sequentially, so the histories agree to rounding; a
threshold crossing or an equilibrium decision can only differ at a tie."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim
from worker_cases import make_task

pytestmark = pytest.mark.gpu
WORKER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpuvoxels_amd", "vx3_node_worker")
CASES = [(_abi.SC_MAX_TIME_STEPS, 1500.4, 0), (_abi.SC_MAX_SIM_TIME, 0.04, 0), (_abi.SC_TEMP_CYCLES, 0.1, 0),
         (_abi.SC_CONST_MAXENERGY, 1e-3, 0), (_abi.SC_MIN_KE, 1e-3, 0), (_abi.SC_MIN_KE, 1e-3, 1),
         (_abi.SC_MIN_MAXMOVE, 1e-3, 0), (_abi.SC_NONE, 0.0, 1)]


def drop_cube(sc, value, eq, lift=2, n=3):
    m = workloads.cube(n, steps=0)
    m.ijk = m.ijk + np.array([0, 0, lift], np.int32)
    m.vary_temp_enabled = 0
    m.stop_condition_type, m.stop_condition_value, m.equilibrium_mode = sc, value, eq
    m.max_steps = 3000 if sc == _abi.SC_NONE else 0
    return m


def run_oracle(m):
    o = OSim(m)
    o.step(100000)
    return o


def test_each_condition_stops_where_the_oracle_does():
    models = [drop_cube(*c) for c in CASES]
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in models]
    for _ in range(40):
        b.step(500)
        if all(s.state != _abi.SIM_RUNNING for s in b.poll().values()):
            break
    status = b.poll()
    for c, m, s in zip(CASES, models, slots):
        o = run_oracle(m)
        so = o.status()
        print(c, "GPU steps", status[s].steps, "oracle", so.steps)
        assert status[s].state == so.state == _abi.SIM_DONE, c
        assert status[s].steps == so.steps, c
        assert status[s].current_time == pytest.approx(so.current_time, rel=1e-12)
        g, r = b.read_state(s), o.state()
        assert np.max(np.abs(g["pos"] - r["pos"])) < 1e-9, c
        assert np.max(np.abs(g["lin_mom"] - r["lin_mom"])) < 1e-9, c
    b.close()


def test_mixed_with_plain_simulations():
    # VoxCAD-stop simulations run on the stream path next to resident ones; both stay
    # the same as when run alone
    plain = workloads.robots(2, seed=3, grid=8, steps=1200)
    stop = drop_cube(_abi.SC_MIN_KE, 1e-3, 1)
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in plain + [stop]]
    b.step(6000)
    for m, s in zip(plain + [stop], slots):
        o = run_oracle(m)
        st = b.poll()[s]
        assert st.steps == o.status().steps
        assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-9
    b.close()


def test_resident_path_refuses():
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT)
    with pytest.raises(engine.VX3Error) as e:
        b.admit(drop_cube(_abi.SC_MIN_KE, 1e-3, 0))
    assert e.value.status == _abi.EUNSUPPORTED
    b.close()


def test_worker_flag_applies_the_vxa_keys(tmp_path):
    # without --voxcad-stop the worker ignores StopConditionType like the Voxelyze3
    # worker (the MathTree stop t > 0.05 ends the run); with it, the max-steps
    # condition (steps > 1200) ends it first
    vxt, inp, base_text, names = make_task(str(tmp_path), n=1, stop_time=0.05)
    path = os.path.join(inp, names[0])
    text = open(path).read().replace(
        "</VXD>", '<StopConditionType replace="VXA.Simulator.StopCondition.StopConditionType">1</StopConditionType>'
                  '<StopConditionValue replace="VXA.Simulator.StopCondition.StopConditionValue">1200'
                  '</StopConditionValue></VXD>')
    with open(path, "w") as f:
        f.write(text)
    times = []
    for flags in ([], ["--voxcad-stop"]):
        out = tmp_path / ("r%d.xml" % len(flags))
        r = subprocess.run([WORKER, "-i", vxt, "-o", str(out)] + flags, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        node = ET.parse(str(out)).getroot().find("detail").find(names[0].split(".")[0])
        times.append(float(node.findtext("currentTime")))
    v = engine.Vxa(base_text, text)
    assert v.info.stop_condition_type == 1 and v.info.stop_condition_value == 1200
    d = v.desc
    o = OSim.from_desc(d)
    o.step(100000)
    assert times[0] == pytest.approx(o.status().current_time, rel=1e-12)   # MathTree stop only
    d.stop_condition_type, d.stop_condition_value = 1, 1200.0
    o2 = OSim.from_desc(d)
    assert o2.step(100000) == 1201
    assert times[1] == pytest.approx(o2.status().current_time, rel=1e-12)
    assert times[1] < times[0]
