"""Collisions on the resident path (one workgroup per simulation, contact lists in LDS,
rebuilt by a counting sort into LDS cell buckets when stale): pair lists and FP64 states
equal to the per-step kernels' and to the CPU oracle's; a simulation whose lists outgrow
the LDS table moves to the per-step kernels and carries on unchanged."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim

pytestmark = pytest.mark.gpu


def run(models, path, steps, chunks, precision=_abi.PREC_F64, table=0, poll_between=False):
    b = engine.Batch(precision=precision, path=path, steps_per_launch=250, collision_table_entries=table)
    slots = [b.admit(m) for m in models]
    out = []
    per = steps // chunks
    for c in range(chunks):
        if poll_between:   # a simulation that stood still for a table catches up after the poll
            for _ in range(10):
                st = b.poll()
                lag = [(c + 1) * per - st[s].steps for s in slots]
                if max(lag) <= 0:
                    break
                b.step(max(lag))
        else:
            b.step(per)
        out.append([(b.collision_pairs(s), b.read_state(s), b.poll()[s]) for s in slots])
    b.close()
    return out


def test_balls_resident_equal_stream_and_oracle_f64():
    """8 radius-2 balls half a voxel apart (264 voxels, one resident simulation), the
    upper layer falling into the lower one's gaps."""
    m = workloads.balls(n_side=2, layers=2, radius=2, gap=0.5, steps=0)
    res = run([m], _abi.PATH_RESIDENT, 4000, 4)
    stm = run([m], _abi.PATH_STREAM, 4000, 4)
    o = OSim(m)
    touched = False
    for (rp, rs, rst), (sp, ss, sst) in zip((r[0] for r in res), (s[0] for s in stm)):
        o.step(1000)
        assert rst.steps == sst.steps == o.status().steps
        assert np.array_equal(rp, sp) and np.array_equal(rp, o.collision_pairs())
        assert np.array_equal(rs["pos"], ss["pos"]) and np.array_equal(rs["orient"], ss["orient"])
        assert np.max(np.abs(rs["pos"] - o.state()["pos"])) < 1e-9
        d = np.linalg.norm(rs["pos"][rp[:, 0]] - rs["pos"][rp[:, 1]], axis=1)
        touched |= bool(np.any(d < 1.25 * m.voxel_size))
    assert touched, "the balls never touched"


def test_robots_with_collisions_resident_f64():
    """The 512-robot workload's robots with self-collision on (the Voxelyze3 default):
    about 3,000 watched pairs each (6,000 partner entries, within the resident table),
    resident, equal to the per-step kernels and within 1e-9 m of the oracle."""
    rs = workloads.robots(6, seed=0, steps=0)
    for r in rs:
        r.collisions_enabled = 1
    res = run(rs, _abi.PATH_RESIDENT, 2000, 2)
    stm = run(rs, _abi.PATH_STREAM, 2000, 2)
    for r, (rp, rst_, rs_), (sp, sst_, ss_), in zip(rs, res[-1], stm[-1]):
        assert rs_.steps == ss_.steps == 2000   # no simulation stood still for a larger table
        assert np.array_equal(rp, sp)
        assert np.array_equal(rst_["pos"], sst_["pos"])
    os_ = [OSim(r) for r in rs]
    for o, r, (rp, rst_, st) in zip(os_, rs, res[-1]):
        o.step(2000)
        assert st.steps == 2000
        assert np.array_equal(rp, o.collision_pairs())
        assert np.max(np.abs(rst_["pos"] - o.state()["pos"])) < 1e-9


def test_resident_collisions_f32_equal_stream():
    m = workloads.drop(3, gap=1, shift=1, steps=0)
    res = run([m], _abi.PATH_RESIDENT, 3000, 3, precision=_abi.PREC_F32)
    stm = run([m], _abi.PATH_STREAM, 3000, 3, precision=_abi.PREC_F32)
    for (rp, rs, _), (sp, ss, _) in zip((r[0] for r in res), (s[0] for s in stm)):
        assert np.array_equal(rp, sp)
        assert np.array_equal(rs["pos"], ss["pos"])


def test_resident_lists_beyond_the_table_move_to_the_step_kernels():
    """A 16-entry collision table is too small once the balls touch: the resident kernel
    stops the simulation in place, the next poll moves it to the per-step kernels (which
    grow the table), and the run ends as the run with a large table from the start."""
    m = workloads.balls(n_side=2, layers=2, radius=3, steps=0)
    small = run([m], _abi.PATH_AUTO, 4000, 4, table=16, poll_between=True)
    big = run([m], _abi.PATH_AUTO, 4000, 4)
    (sp, ss, sst), (bp, bs, bst) = small[-1][0], big[-1][0]
    assert sst.steps == bst.steps == 4000
    assert np.array_equal(sp, bp)
    assert np.array_equal(ss["pos"], bs["pos"])
