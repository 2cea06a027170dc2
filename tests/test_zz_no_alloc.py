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
