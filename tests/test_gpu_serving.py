"""Serving: continuous batching gives every request the result it gets alone."""
import numpy as np
import pytest

from concurrent.futures import ThreadPoolExecutor

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim

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


def full_size_requests():
    """Requests spanning the serving stream's range: robots of ~100 to ~50,000 voxels
    (grids 7 to 47), horizons of 2,000 steps and more."""
    rng = np.random.default_rng(7)
    out = []
    for i, g in enumerate([7, 11, 19, 30, 40, 47]):
        m = workloads.robot(rng, grid=g, steps=0, name="full%d" % i)
        out.append((m, 2000 + 97 * i))
    return out


def serve(reqs, precision, quantum=700, max_sims=4):
    sv = engine.Server(precision=precision, quantum_steps=quantum, max_sims=max_sims)
    tickets = {sv.submit(m, h): i for i, (m, h) in enumerate(reqs)}
    results, rounds = {}, 0
    while True:
        active = sv.round()
        rounds += 1
        for r in sv.results():
            results[tickets[r.ticket]] = r
        if active == 0:
            break
        assert rounds < 500
    sv.close()
    return results


@pytest.mark.parametrize("precision", [_abi.PREC_F64, _abi.PREC_F32])
def test_full_size_requests_same_alone_and_in_mix(precision):
    reqs = full_size_requests()
    assert max(m.n_voxels for m, _ in reqs) > 45000 and min(m.n_voxels for m, _ in reqs) < 250
    results = serve(reqs, precision)
    assert len(results) == len(reqs)
    for i, (m, h) in enumerate(reqs):
        r = results[i]
        assert r.status.state == _abi.SIM_DONE and r.status.steps == h
        ref = solo(m, h, precision)
        assert r.status.current_time == ref.current_time
        assert r.status.fitness == ref.fitness
        assert list(r.status.current_com) == list(ref.current_com)


def test_full_size_requests_match_oracle_f64():
    reqs = full_size_requests()
    picked = [reqs[0], reqs[2], reqs[5]]   # ~100, ~3,000 and ~50,000 voxels
    results = serve(picked, _abi.PREC_F64)

    def oracle(mh):
        o = OSim(mh[0])
        o.step(mh[1])
        return o.status()
    with ThreadPoolExecutor(max_workers=3) as ex:
        ost = list(ex.map(oracle, picked))
    for i, ((m, h), o) in enumerate(zip(picked, ost)):
        g = results[i].status
        assert g.steps == o.steps == h
        err = abs(g.fitness - o.fitness) / abs(o.fitness)
        print("%d voxels, %d steps: relative fitness error %.3e" % (m.n_voxels, h, err))
        assert err < 1e-6
        assert np.max(np.abs(np.array(g.current_com) - np.array(o.current_com))) < 1e-9


def test_submissions_from_another_thread_while_rounds_run():
    """Requests submitted from a second thread while the driving thread runs rounds (the
    server prepares them on its own threads meanwhile) get the results they get alone."""
    import threading
    import time
    rng = np.random.default_rng(11)
    reqs = [(workloads.robot(rng, grid=int(g), steps=0, name="thr%d" % i), int(h))
            for i, (g, h) in enumerate(zip(rng.integers(5, 13, size=12), rng.integers(200, 900, size=12)))]
    sv = engine.Server(precision=_abi.PREC_F64, quantum_steps=150, max_sims=5)
    tickets, lock = {}, threading.Lock()

    def submitter():
        for i, (m, h) in enumerate(reqs):
            t = sv.submit(m, h)
            with lock:
                tickets[t] = i
            time.sleep(0.005)

    th = threading.Thread(target=submitter)
    th.start()
    results, t0 = {}, time.time()
    while True:
        active = sv.round()
        for r in sv.results():
            results[r.ticket] = r
        if not th.is_alive() and active == 0 and len(results) == len(reqs):
            break
        if active == 0:
            time.sleep(0.002)   # nothing queued yet: wait for the submitter
        assert time.time() - t0 < 120
    th.join()
    sv.close()
    assert sorted(results) == sorted(tickets)
    for t, r in results.items():
        m, h = reqs[tickets[t]]
        assert r.status.state == _abi.SIM_DONE and r.status.steps == h
        ref = solo(m, h, _abi.PREC_F64)
        assert r.status.current_time == ref.current_time
        assert r.status.fitness == ref.fitness
        assert list(r.status.current_com) == list(ref.current_com)
