"""vx3_node_worker: command-line contracts (CPU) and full
task runs checked against the CPU oracle (GPU)."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine
from oracle_lib import OSim
from worker_cases import make_task

WORKER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpuvoxels_amd", "vx3_node_worker")
pytestmark = pytest.mark.skipif(not os.path.exists(WORKER), reason="worker not built")


def run(args, timeout=300):
    return subprocess.run([WORKER] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)


def test_usage_and_argument_errors(tmp_path):
    assert run([]).returncode == 1
    assert run(["-i", str(tmp_path / "missing.vxt"), "-o", str(tmp_path / "r.xml")]).returncode == 1
    out = tmp_path / "exists.xml"
    out.write_text("x")
    vxt, _, _, _ = make_task(str(tmp_path))
    r = run(["-i", vxt, "-o", str(out)])
    assert r.returncode == 1 and "output file exists" in r.stdout
    assert out.read_text() == "x"


@pytest.mark.skipif(engine.device_count() > 0, reason="checks the no-GPU path")
def test_no_gpu_exit_status(tmp_path):
    vxt, _, _, _ = make_task(str(tmp_path))
    r = run(["-i", vxt, "-o", str(tmp_path / "r.xml")])
    assert r.returncode == 1 and "No GPU found" in r.stdout


@pytest.mark.gpu
def test_worker_task_matches_oracle(tmp_path):
    vxt, inp, base_text, names = make_task(str(tmp_path), n=3, stop_time=0.01, save_positions=True)
    out = tmp_path / "report.xml"
    r = run(["-i", vxt, "-o", str(out)])
    assert r.returncode == 0, r.stdout + r.stderr
    rep = ET.parse(str(out)).getroot()
    assert rep.tag == "report" and rep.findtext("inputdir") == "generation_0001"
    detail = rep.find("detail")
    fits = []
    for name in names:
        node = detail.find(name.split(".")[0])
        v = engine.Vxa(base_text, open(os.path.join(inp, name)).read())
        o = OSim.from_desc(v.desc)
        o.step(100000)    # runs until the stop condition t > 0.01
        st = o.status()
        assert st.state == _abi.SIM_DONE
        assert int(node.findtext("num_voxel")) == v.n_voxels
        assert float(node.findtext("currentTime")) == pytest.approx(st.current_time, rel=1e-12)
        f = float(node.findtext("fitness_score"))
        assert f == pytest.approx(st.fitness, rel=1e-9, abs=1e-15)
        pos = np.array([[float(c) for c in p.split(",")] for p in node.findtext("pos").strip(";").split(";")])
        assert pos.shape == (v.n_voxels, 3)
        assert np.max(np.abs(pos - o.state()["pos"])) < 1e-6   # written with six decimals
        fits.append((f, name))
    best = max(fits)
    assert rep.find("bestfit").findtext("filename") == best[1]
