"""The collision-heavy multi-body configuration at full size: 2,048 radius-4 balls
(526,336 voxels, 1,277,952 links) in one simulation, random non-overlapping
positions in a 400x400-voxel arena, per-ball initial velocities N(0, 0.05 m/s),
gravity and floor, the broadphase rebuilt every 8 steps.  Checked against the CPU
oracle (the reference's CPU pair rule: watch radius, link-hop exclusion, pairs in
ascending order) for a bounded number of steps spanning three rebuilds."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim

pytestmark = pytest.mark.gpu

ARENA = 4.0   # m: the 400-voxel arena, the position scale of the relative error


def step_to(b, slot, n, calls):
    """Steps until the simulation has done n steps (a simulation whose collision table
    had to grow stands still until a poll enlarges it, then carries on); calls[0]
    counts the vx3h_batch_step calls this took."""
    for _ in range(20):
        st = b.poll()[slot]
        if st.steps >= n or st.state != _abi.SIM_RUNNING:
            return st
        b.step(n - st.steps)
        calls[0] += 1
    raise AssertionError("simulation does not advance")


def run_chunks(model, chunks, table=0, precision=_abi.PREC_F64, calls=None):
    o = OSim(model)
    o.set_collision_interval(8)
    b = engine.Batch(precision=precision, path=_abi.PATH_AUTO, collision_rebuild_interval=8,
                     collision_table_entries=table)
    s = b.admit(model)
    out, done = [], 0
    calls = calls if calls is not None else [0]
    for n in chunks:
        done += n
        o.step(n)
        gst = step_to(b, s, done, calls)
        out.append((o.collision_pairs(), b.collision_pairs(s), o.state(), b.read_state(s), o.status(), gst))
    b.close()
    return out


def check(res, model, tol):
    for op, gp, os_, gs, ost, gst in res:
        assert gst.state == ost.state == _abi.SIM_RUNNING and gst.steps == ost.steps
        assert np.array_equal(op, gp)
        err = np.max(np.abs(gs["pos"] - os_["pos"])) / ARENA
        print("steps %d pairs %d max |dpos|/arena %.3e" % (gst.steps, len(gp), err))
        assert err < tol
        assert np.max(np.abs(gs["orient"] - os_["orient"])) < 1e-9
        assert np.array_equal(gs["flags"] & 8, os_["flags"] & 8)


def test_c3_config_full_size_f64():
    m = workloads.arena_balls(steps=0)
    assert m.n_voxels == 526336
    # rebuilds at step indices 0, 8 and 16
    res = run_chunks(m, [1, 8, 8])
    check(res, m, 1e-9)
    # the balls start >= 2.5 voxels apart: nothing is watched until they close in
    assert len(res[-1][1]) == len(res[-1][0])


def test_c3_dense_full_size_contacts_f64():
    """The same 2,048 balls packed with no gap between surfaces (centres 9 voxels
    apart): neighbouring balls are watched and touching from the first step."""
    m = workloads.arena_balls(gap=0.0, steps=0, seed=1)
    res = run_chunks(m, [1, 8, 8])
    check(res, m, 1e-9)
    assert min(len(r[1]) for r in res) > 1000


def test_collision_table_grows_instead_of_failing():
    """A neighbour table of 16 entries is too small for touching balls:
    the simulation stands still until a poll enlarges the table, then runs exactly as
    with a large table from the start."""
    m = workloads.arena_balls(n_balls=64, arena=80, gap=0.0, steps=0, seed=2)
    calls = [0]
    small = run_chunks(m, [1, 8, 8], table=16, calls=calls)
    check(small, m, 1e-11)
    assert calls[0] > 3, "the table never had to grow"
    big = run_chunks(m, [17], table=0)
    assert np.array_equal(small[-1][3]["pos"], big[-1][3]["pos"])


def test_c3_f32_within_tolerance():
    """The configuration's stated precision: FP32 state over 17 steps (three rebuilds)
    against the FP64 oracle, positions relative to the arena scale."""
    m = workloads.arena_balls(steps=0)
    res = run_chunks(m, [17], precision=_abi.PREC_F32)
    op, gp, os_, gs, ost, gst = res[-1]
    assert gst.steps == 17
    assert np.array_equal(op, gp)
    assert np.max(np.abs(gs["pos"] - os_["pos"])) / ARENA < 1e-6
