"""GPU collisions (spatial-hash neighbour lists) against the CPU oracle: the watched
pair lists match exactly and FP64 trajectories stay at rounding level."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim

pytestmark = pytest.mark.gpu


def run_both(m, steps, checkpoints, precision=_abi.PREC_F64, interval=0):
    o = OSim(m)
    b = engine.Batch(precision=precision, collision_rebuild_interval=interval)
    s = b.admit(m)
    out = []
    per = steps // checkpoints
    for _ in range(checkpoints):
        o.step(per)
        b.step(per)
        out.append((o.collision_pairs(), b.collision_pairs(s), o.state(), b.read_state(s), o.status(), b.poll()[s]))
    b.close()
    return out


def test_drop_pairs_and_trajectory_f64():
    m = workloads.drop(3, gap=1, shift=1, steps=0)
    res = run_both(m, 4200, 6)
    contact = False
    for op, gp, os_, gs, ost, gst in res:
        assert gst.state == ost.state == _abi.SIM_RUNNING
        assert np.array_equal(op, gp)
        assert np.max(np.abs(os_["pos"] - gs["pos"])) < 1e-10
        # some cross-body pair closer than the contact envelope (1.25 voxel)
        top = np.arange(m.n_voxels) >= 27
        d = np.linalg.norm(gs["pos"][gp[:, 0]] - gs["pos"][gp[:, 1]], axis=1)
        cross = top[gp[:, 0]] != top[gp[:, 1]]
        if np.any(d[cross] < 1.25 * m.voxel_size):
            contact = True
    assert contact, "the bodies never touched"


def test_slab_collisions_f64():
    m = workloads.slab(8, 6, 4, collisions=True)
    for op, gp, os_, gs, ost, gst in run_both(m, 400, 2):
        assert np.array_equal(op, gp)
        assert np.max(np.abs(os_["pos"] - gs["pos"])) < 1e-9   # actuation sin() differs in the last ulp


def test_rebuild_interval_keeps_physics():
    # rebuilding more often than the staleness rule only changes which non-touching
    # pairs are watched: the forces, and so the trajectory, are identical
    m = workloads.drop(3, gap=1, shift=1, steps=0)
    a = run_both(m, 3600, 1, interval=0)[-1]
    c = run_both(m, 3600, 1, interval=8)[-1]
    assert np.array_equal(a[3]["pos"], c[3]["pos"])


def test_drop_f32_within_tolerance():
    m = workloads.drop(3, gap=1, shift=1, steps=0)
    op, gp, os_, gs, ost, gst = run_both(m, 1000, 1, precision=_abi.PREC_F32)[-1]
    assert np.max(np.abs(os_["pos"] - gs["pos"])) / (6 * m.voxel_size) < 1e-4


def test_many_balls_pairs_and_trajectory_f64():
    """18 balls of 257 voxels in one simulation (two offset layers): the upper layer
    falls into the lower one; pair lists equal the oracle's at every checkpoint and
    some cross-ball pair comes within the contact envelope."""
    m = workloads.balls(3, 2, 4, 1.5, steps=0)
    ball = np.zeros(m.n_voxels, np.int64)
    # ball id of each voxel: nearest of the generating centres, recovered by clustering the lattice cells
    from scipy.ndimage import label
    grid = np.zeros(m.ijk.max(axis=0) + 1, bool)
    grid[tuple(m.ijk.T)] = True
    lab, n = label(grid)
    assert n == 18
    ball[:] = lab[tuple(m.ijk.T)]
    contact = False
    for op, gp, os_, gs, ost, gst in run_both(m, 4800, 6):
        assert gst.state == ost.state == _abi.SIM_RUNNING
        assert np.array_equal(op, gp)
        assert np.max(np.abs(os_["pos"] - gs["pos"])) < 1e-9
        if len(gp):
            d = np.linalg.norm(gs["pos"][gp[:, 0]] - gs["pos"][gp[:, 1]], axis=1)
            cross = ball[gp[:, 0]] != ball[gp[:, 1]]
            contact |= bool(np.any(d[cross] < 1.25 * m.voxel_size))
    assert contact, "no two balls touched"
