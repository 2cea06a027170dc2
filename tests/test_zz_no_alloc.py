"""vx3h_batch_step_noalloc through the C ABI against vx3h_batch_step: the same states
(the same kernels, with the step's end loads in the library's static device array
instead of the batch's buffer), on two calls in a row, the second on a smaller batch."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads

pytestmark = pytest.mark.gpu


def states(models, precision, noalloc, chunks):
    b = engine.Batch(precision=precision, path=_abi.PATH_AUTO, steps_per_launch=100)
    slots = [b.admit(m) for m in models]
    for n in chunks:
        if noalloc:
            b.step_noalloc(n)
        else:
            b.step(n)
    st = b.poll()
    out = [(b.read_state(s), st[s]) for s in slots]
    b.close()
    return out


@pytest.mark.parametrize("precision", [_abi.PREC_F64, _abi.PREC_F32])
def test_noalloc_step_equals_batch_step(precision):
    big = workloads.robots(6, seed=17, steps=0) + [workloads.drop(3, steps=0)]
    small = [workloads.cube(4, steps=0)]
    for models, chunks in ((big, [1, 150, 149]), (small, [60, 40])):   # second call: a smaller batch
        ref = states(models, precision, False, chunks)
        got = states(models, precision, True, chunks)
        for (g, gs), (r, rs) in zip(got, ref):
            assert gs.steps == rs.steps == sum(chunks)
            assert gs.current_time == rs.current_time
            for k in ("pos", "orient", "lin_mom", "ang_mom", "flags"):
                assert np.array_equal(g[k], r[k]), k


def test_noalloc_refuses_batches_beyond_its_scratch():
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    m = workloads.slab(80, 80, 80, steps=0, collisions=False)   # 512,000 voxels: 147 MB of FP64 end loads
    b.admit(m)
    with pytest.raises(engine.VX3Error) as e:
        b.step_noalloc(1)
    assert e.value.status == _abi.EINVAL
    b.close()


@pytest.mark.parametrize("first_noalloc", [False, True])
def test_resume_after_link_pass_across_entry_points(first_noalloc):
    """Touching balls with a 16-entry neighbour table: the first step call stands the
    simulation still after the link pass of its first step until a poll enlarges the
    table; the next call, through the other entry point, resumes it (its end loads
    carried between the batch's buffer and the static array).  It ends where
    vx3h_batch_step with a large table from the start ends."""
    m = workloads.arena_balls(n_balls=64, arena=80, gap=0.0, steps=0, seed=2)

    def run(alternate, table):
        b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_AUTO, collision_table_entries=table,
                         collision_rebuild_interval=8)
        s = b.admit(m)
        calls, after_first = 0, None
        while True:
            st = b.poll()[s]
            if calls == 1:
                after_first = st.steps
            if st.steps >= 17:
                break
            if alternate and (calls % 2 == 0) == first_noalloc:
                b.step_noalloc(17 - st.steps)
            else:
                b.step(17 - st.steps)
            calls += 1
            assert calls < 200, "simulation does not advance"
        out = b.read_state(s), b.collision_pairs(s), calls, after_first
        b.close()
        return out

    (ga, pa, ca, fa), (gc, pc, _, _) = run(True, 16), run(False, 0)
    assert fa == 0 and ca >= 2, "the first call did not stop for the table"
    assert np.array_equal(pa, pc)
    for k in ("pos", "orient", "lin_mom", "ang_mom"):
        assert np.array_equal(ga[k], gc[k]), k
