"""VX3_VoxelyzeKernel::doTimeStep(float dt) through the C ABI (vx3h_sim_step), against the
CPU oracle's VX3 driver and the batch path: dt < 0 (DtFrac x the cached recommended
step), dt == 0 (a counted step that changes nothing), dt > 0 (a caller-given step), and
the false return on divergence (VX3_VoxelyzeKernel.cu:237-359)."""
import os
import subprocess

import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim
from test_gpu_stream import two_voxel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_default_dt_matches_batch_and_oracle():
    m = workloads.cube(5, steps=0)
    o = OSim(m)
    o.step(300)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    s = b.admit(m)
    for _ in range(300):
        assert b.do_time_step(s)
    g, st = b.read_state(s), b.poll()[s]
    ref = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    r = ref.admit(m)
    ref.step(300)
    assert np.array_equal(g["pos"], ref.read_state(r)["pos"])
    assert st.steps == 300 and st.current_time == ref.poll()[r].current_time
    assert np.max(np.abs(g["pos"] - o.state()["pos"])) < 1e-12
    assert st.current_time == o.status().current_time
    b.close()
    ref.close()


def test_zero_dt_counts_a_step_only():
    m = workloads.cube(4, steps=0)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    s = b.admit(m)
    for _ in range(20):
        b.do_time_step(s)
    g0, st0 = b.read_state(s), b.poll()[s]
    assert b.do_time_step(s, 0.0)
    g1, st1 = b.read_state(s), b.poll()[s]
    assert st1.steps == st0.steps + 1
    assert st1.current_time == st0.current_time
    for k in ("pos", "orient", "lin_mom", "ang_mom"):
        assert np.array_equal(g0[k], g1[k])
    b.close()


def test_given_dt_matches_oracle_fixed_step():
    """doTimeStep(dt) with dt = 0.7 x the recommended step every call equals the oracle
    run with that fixed step: time, sampling period (int(TempPeriod/dt)) and damping's
    previous step all follow the given dt."""
    m = workloads.cube(5, steps=0)
    probe = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    p = probe.admit(m)
    probe.step(1)
    dt = float(np.float32(0.7 * probe.poll()[p].recommended_dt))
    probe.close()
    mf = workloads.cube(5, steps=0)
    mf.dt_fixed = dt
    o = OSim(mf)
    o.step(2500)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    s = b.admit(m)
    for _ in range(2500):
        assert b.do_time_step(s, dt)
    st, ost = b.poll()[s], o.status()
    assert st.steps == ost.steps == 2500
    assert abs(st.current_time - ost.current_time) < 1e-15
    assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-12
    assert st.recent_angle == pytest.approx(ost.recent_angle, rel=1e-9, abs=1e-12)
    b.close()


def test_divergence_returns_false_at_the_oracle_step():
    m = two_voxel((1e3, 0, 0))
    o = OSim(m)
    o.step(200)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    s = b.admit(m)
    n = 0
    while b.do_time_step(s):
        n += 1
        assert n < 200
    st = b.poll()[s]
    assert st.state == o.status().state == _abi.SIM_DIVERGED
    assert st.steps == o.status().steps
    assert not b.do_time_step(s)              # stays diverged, no further step
    assert b.poll()[s].steps == st.steps
    b.close()


def test_single_steps_of_a_resident_simulation_in_a_batch():
    """A simulation admitted to the resident path, stepped alone by doTimeStep between
    batch steps, follows the run that only ever takes batch steps (both FP64 kernels
    compute identically); its neighbour in the batch does not move during those calls."""
    rs = workloads.robots(2, seed=9, grid=8, steps=0)
    a = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_AUTO, steps_per_launch=100)
    sa = [a.admit(r) for r in rs]
    c = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_AUTO, steps_per_launch=100)
    sc = [c.admit(r) for r in rs]
    a.step(100)
    other = a.read_state(sa[1])["pos"]
    for _ in range(7):
        assert a.do_time_step(sa[0])
    assert np.array_equal(a.read_state(sa[1])["pos"], other)
    a.step(93)          # robot 0 reaches 200 steps ...
    c.step(200)
    for _ in range(7):  # ... robot 1 catches up
        assert a.do_time_step(sa[1])
    for x, y in zip(sa, sc):
        assert a.poll()[x].steps == c.poll()[y].steps == 200
        assert np.max(np.abs(a.read_state(x)["pos"] - c.read_state(y)["pos"])) < 1e-14
    a.close()
    c.close()


def test_cpp_facade_program(tmp_path):
    """The C++ classes with the reference's names, compiled against the engine: a
    VX3_VoxelyzeKernel stepped until StopConditionMet, as the reference's device
    thread does (VX3_SimulationManager.cu:62-69)."""
    src = tmp_path / "facade.cpp"
    src.write_text(CPP)
    exe = tmp_path / "facade"
    lib = os.path.join(ROOT, "gpuvoxels_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), str(src), "-I" + ROOT, "-L" + lib, "-lvx3h",
                    "-Wl,-rpath," + lib], check=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    steps, t = r.stdout.split()
    assert int(steps) == 123 and float(t) > 0


CPP = r"""
#include <cstdio>
#include <vector>
#include "gpuvoxels_amd/csrc/vx3_kernel.h"
#include "gpuvoxels_amd/csrc/vx3_manager.h"

int main() {
    vx3h_material_desc mat = {};
    mat.elastic_mod = 1e6; mat.density = 1000; mat.poissons_ratio = 0; mat.u_static = 1; mat.u_dynamic = 0.5;
    mat.zeta_internal = 1; mat.zeta_global = 0.01; mat.zeta_collision = 0.8; mat.is_measured = 1;
    std::vector<int> ijk, m;
    for (int z = 0; z < 3; z++) for (int y = 0; y < 3; y++) for (int x = 0; x < 3; x++) {
        ijk.push_back(x); ijk.push_back(y); ijk.push_back(z); m.push_back(0);
    }
    vx3h_sim_desc d = {};
    d.voxel_size = 0.01; d.n_materials = 1; d.materials = &mat; d.n_voxels = 27; d.ijk = ijk.data();
    d.material = m.data(); d.gravity_mult = 1; d.floor_enabled = 1; d.dt_frac = 0.9; d.max_steps = 123;
    VX3_VoxelyzeKernel k(&d);
    while (!k.StopConditionMet())
        if (!k.doTimeStep()) return 2;
    std::printf("%d %g\n", k.status().steps, k.status().current_time);
    VX3_SimulationManager* unused = nullptr;   // the manager under the reference's name
    (void)unused;
    return 0;
}
"""
