"""The reference's Experiments/xenobot/h.target_only/data1 run (base.vxa + bot_0000.vxd,
kept in tests/golden/target_only.npz): 2,964 voxels of four materials on a 112x112x7
lattice -- a ciliated xenobot with signals among sticky target voxels and fixed
voxels, collisions, attach and detach on, close pairs counted within 3 voxel lengths
(computeTargetCloseness, VX3_VoxelyzeKernel.cu:545-562), stop at t > 10 s.

Loaded by the engine's VXA loader and run on the GPU against the oracle built from
the same description: states, hit counts, close pairs, closeness, links made and pair
lists; then end to end through the worker (task file -> report XML) with the stop
time shortened."""
import os
import re
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine
from oracle_lib import OSim

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
WORKER = os.path.join(os.path.dirname(HERE), "gpuvoxels_amd", "vx3_node_worker")


def documents():
    z = np.load(os.path.join(HERE, "golden", "target_only.npz"))
    return z["vxa"].tobytes().decode(), z["vxd"].tobytes().decode()


def test_target_only_experiment_matches_oracle():
    base, vxd = documents()
    v = engine.Vxa(base, vxd)
    d = v.desc
    assert d.n_voxels == 2964 and d.collisions_enabled and d.enable_attach and d.enable_detach
    assert d.max_dist_close_pairs == 3.0
    o = OSim.from_desc(d)
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s = b.admit(v)
    for _ in range(3):
        b.step(200)
        o.step(200)
        st, so = b.poll()[s], o.status()
        assert st.state == so.state == _abi.SIM_RUNNING and st.steps == so.steps
        assert st.collision_count == so.collision_count
        assert st.num_close_pairs == so.num_close_pairs
        assert st.target_closeness == pytest.approx(so.target_closeness, rel=1e-12)
        assert st.n_links == o.nl   # links made by attach included
        g, r = b.read_state(s, nl=o.nl), o.state()
        assert np.max(np.abs(g["pos"] - r["pos"])) < 1e-9
        assert np.array_equal(g["flags"] & 2, r["flags"] & 2)
        assert np.array_equal(b.collision_pairs(s), o.collision_pairs())
    print("600 steps: hits %d, close pairs %d, closeness %.6g" % (so.collision_count, so.num_close_pairs,
                                                                   so.target_closeness))
    assert so.collision_count > 0 and so.num_close_pairs > 0
    b.close()


def test_target_only_experiment_end_to_end_through_the_worker(tmp_path):
    """The worker on the experiment's documents, the stop condition t > 10 replaced by
    t > 0.1 (about 220 steps) and history recording off: the report's fitness, time
    and voxel count against the oracle run to the same stop."""
    base, vxd = documents()
    base = re.sub(r"<StopConditionFormula>.*?</StopConditionFormula>",
                  "<StopConditionFormula><mtSUB><mtVAR>t</mtVAR><mtCONST>0.1</mtCONST></mtSUB></StopConditionFormula>",
                  base, flags=re.S)
    base = re.sub(r"<RecordStepSize>\d+</RecordStepSize>", "<RecordStepSize>0</RecordStepSize>", base)
    inp = tmp_path / "data1"
    inp.mkdir()
    (inp / "base.vxa").write_text(base)
    (inp / "bot_0000.vxd").write_text(vxd)
    vxt = tmp_path / "task.vxt"
    vxt.write_text('<?xml version="1.0" encoding="utf-8"?>\n<vxa>%s</vxa><input_dir>%s</input_dir><vxd><f>bot_0000.vxd</f></vxd>\n'
                   % (inp / "base.vxa", inp))
    out = tmp_path / "report.xml"
    r = subprocess.run([WORKER, "-i", str(vxt), "-o", str(out)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    node = ET.parse(str(out)).getroot().find("detail").find("bot_0000")
    doc = engine.Vxa(base, vxd)   # kept alive while the oracle is built from its description
    o = OSim.from_desc(doc.desc)
    o.step(100000)   # until t > 0.1
    so = o.status()
    assert so.state == _abi.SIM_DONE
    assert int(node.findtext("num_voxel")) == 2964
    assert float(node.findtext("currentTime")) == pytest.approx(so.current_time, rel=1e-12)
    assert float(node.findtext("fitness_score")) == pytest.approx(so.fitness, rel=1e-6, abs=1e-12)
