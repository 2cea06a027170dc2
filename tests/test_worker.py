"""vx3_node_worker and the Voxelyze3 launcher: command-line contracts (CPU) and full
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


LAUNCHER = os.path.join(os.path.dirname(WORKER), "Voxelyze3")


def launch(args, cwd, timeout=300):
    return subprocess.run([LAUNCHER] + args, cwd=str(cwd), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=timeout)


def test_launcher_argument_errors(tmp_path):
    vxt, inp, _, _ = make_task(str(tmp_path))
    assert launch([], tmp_path).returncode == 1
    assert launch(["-i", str(tmp_path / "nope")], tmp_path).returncode == 1
    empty = tmp_path / "empty"
    empty.mkdir()
    r = launch(["-i", str(empty), "-w", WORKER], tmp_path)
    assert r.returncode == 1 and "No base.vxa" in r.stdout
    r = launch(["-i", inp, "-w", str(tmp_path / "no_worker")], tmp_path)
    assert r.returncode == 1 and "worker" in r.stdout
    out = tmp_path / "out.xml"
    out.write_text("keep")
    r = launch(["-i", inp, "-w", WORKER, "-o", str(out), "-l"], tmp_path)
    assert r.returncode == 1 and out.read_text() == "keep"
    assert launch(["-i", inp, "-n", "2", "-l"], tmp_path).returncode == 1


def test_launcher_task_file_and_copy(tmp_path):
    """With a stand-in worker that copies its task file to its output, the report path
    receives the task file the launcher wrote: base.vxa, the input directory and every
    .vxd (any case) in name order."""
    vxt, inp, _, names = make_task(str(tmp_path))
    with open(os.path.join(inp, "extra.VXD"), "w") as f:
        f.write("<VXD/>")
    with open(os.path.join(inp, "ignored.txt"), "w") as f:
        f.write("x")
    fake = tmp_path / "fake_worker"
    fake.write_text('#!/bin/sh\ncp "$2" "$4"\n')
    fake.chmod(0o755)
    out = tmp_path / "report.xml"
    r = launch(["-i", inp, "-w", str(fake), "-o", str(out), "-lf"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    root = ET.fromstring("<t>%s</t>" % out.read_text().split("?>", 1)[1])
    assert root.findtext("vxa") == os.path.join(inp, "base.vxa")
    assert root.findtext("input_dir") == inp
    assert [e.text for e in root.find("vxd")] == sorted(names + ["extra.VXD"])
    assert len(os.listdir(tmp_path / "workspace" / "locally")) == 2      # .vxt and .vxr


@pytest.mark.gpu
def test_launcher_runs_worker(tmp_path):
    vxt, inp, _, names = make_task(str(tmp_path), n=2, stop_time=0.005)
    direct = tmp_path / "direct.xml"
    assert run(["-i", vxt, "-o", str(direct)]).returncode == 0
    out = tmp_path / "report.xml"
    r = launch(["-i", inp, "-w", WORKER, "-o", str(out), "-lf"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    a, b = ET.parse(str(direct)).getroot(), ET.parse(str(out)).getroot()
    for name in names:
        key = name.split(".")[0]
        assert a.find("detail").find(key).findtext("fitness_score") == b.find("detail").find(key).findtext("fitness_score")
