"""Serving: continuous batching gives every request the result it gets alone."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads

pytestmark = pytest.mark.gpu


def solo(model, steps, precision):
    b = engine.Batch(precision=precision)
    s = b.admit(model)
    b.step(steps)
    st = b.poll()[s]
    b.close()
    return st


@pytest.mark.parametrize("precision", [_abi.PREC_F64])
def test_results_same_alone_and_in_mix(precision):
    rng = np.random.default_rng(5)
    grids = [6, 9, 12, 7, 14, 8, 10, 16]
    models = []
    for i, g in enumerate(grids):
        m = workloads.robot(rng, grid=g, steps=0, name="req%d" % i)
        models.append(m)
    horizons = [int(h) for h in rng.integers(300, 1500, size=len(models))]
    # a small budget forces queueing; a quantum that does not divide the horizons forces
    # retirements mid-quantum and admissions between rounds
    sv = engine.Server(precision=precision, quantum_steps=250, max_sims=3)
    tickets = {}
    for m, h in zip(models[:4], horizons[:4]):
        tickets[sv.submit(m, h)] = (m, h)
    results = {}
    rounds = 0
    while True:
        active = sv.round()
        rounds += 1
        if rounds == 2:                       # late arrivals
            for m, h in zip(models[4:], horizons[4:]):
                tickets[sv.submit(m, h)] = (m, h)
        for r in sv.results():
            results[r.ticket] = r
        if active == 0 and rounds > 2:
            break
        assert rounds < 200
    assert len(results) == len(models)
    for t, r in results.items():
        m, h = tickets[t]
        assert r.status.state == _abi.SIM_DONE and r.status.steps == h
        ref = solo(m, h, precision)
        assert r.status.current_time == ref.current_time
        assert r.status.fitness == ref.fitness
        assert list(r.status.current_com) == list(ref.current_com)
        assert r.t_submit <= r.t_admit <= r.t_done
    sv.close()
