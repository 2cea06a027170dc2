"""The tiled resident kernel on the GPU: simulations without collisions too large for one
workgroup run as tiles, one workgroup each, exchanging boundary poses through a
barrier per step.  Every tile computes its links and voxels with the same code as the
one-workgroup kernel and the per-step kernels, so positions, orientations and flags
equal the per-step kernels' (FP64 and FP32); only the centre of mass is summed in
another order (tiles, then their waves).  Against the oracle at 1e-9 m; divergence at
the oracle's step; stop conditions; several launch groups; a mixed batch."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from gpuvoxels_amd.model import MT, Ext
from oracle_lib import OSim

pytestmark = pytest.mark.gpu


def robot(seed, grid, steps=0):
    return workloads.robot(np.random.default_rng(seed), grid=grid, steps=steps, name="r%d_%d" % (grid, seed))


def run(models, path, steps, precision=_abi.PREC_F64, spl=500):
    b = engine.Batch(precision=precision, path=path, steps_per_launch=spl)
    slots = [b.admit(m) for m in models]
    b.step(steps)
    return b, slots


def assert_same(bt, st, bs, ss, pos_tol, orient_tol):
    pt, ps = bt.poll(), bs.poll()
    for a, c in zip(st, ss):
        ga, gc = bt.read_state(a), bs.read_state(c)
        assert np.max(np.abs(ga["pos"] - gc["pos"])) <= pos_tol
        assert np.max(np.abs(ga["orient"] - gc["orient"])) <= orient_tol
        assert np.array_equal(ga["flags"], gc["flags"])
        assert pt[a].steps == ps[c].steps and pt[a].state == ps[c].state
        assert pt[a].current_time == ps[c].current_time
        assert pt[a].fitness == pytest.approx(ps[c].fitness, rel=1e-9, abs=1e-15)


def test_tiled_f64_equals_stream():
    models = [robot(1, 16), robot(2, 22), robot(3, 30)]
    bt, st = run(models, _abi.PATH_AUTO, 1500)
    tiles = [bt.workgroups(s) for s in st]
    print("tiles", tiles)
    assert all(n > 1 for n in tiles)
    bs, ss = run(models, _abi.PATH_STREAM, 1500)
    assert all(bs.workgroups(s) == 0 for s in ss)
    assert_same(bt, st, bs, ss, 1e-14, 1e-13)
    bt.close()
    bs.close()


def test_more_tiles_than_a_wave_has_lanes():
    """A simulation of more than 64 tiles: a waiting tile's first wave checks several
    arrival flags per lane."""
    m = robot(6, 56)
    bt, st = run([m], _abi.PATH_AUTO, 400)
    print("voxels", m.n_voxels, "tiles", bt.workgroups(st[0]))
    assert bt.workgroups(st[0]) > 64
    bs, ss = run([m], _abi.PATH_STREAM, 400)
    assert_same(bt, st, bs, ss, 1e-14, 1e-13)
    bt.close()
    bs.close()


def test_tiled_f32_equals_stream():
    models = [robot(4, 14), robot(5, 20)]
    bt, st = run(models, _abi.PATH_AUTO, 1000, precision=_abi.PREC_F32)
    assert all(bt.workgroups(s) > 1 for s in st)
    bs, ss = run(models, _abi.PATH_STREAM, 1000, precision=_abi.PREC_F32)
    assert_same(bt, st, bs, ss, 1e-6, 1e-5)
    bt.close()
    bs.close()


def test_tiled_matches_oracle():
    m = robot(6, 14)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT, steps_per_launch=400)
    s = b.admit(m)
    assert b.workgroups(s) > 1
    o = OSim(m)
    for _ in range(3):
        b.step(400)
        o.step(400)
        g, r = b.read_state(s), o.state()
        assert np.max(np.abs(g["pos"] - r["pos"])) < 1e-9
        assert np.array_equal(g["flags"] & 8, r["flags"] & 8)
    st, so = b.poll()[s], o.status()
    assert st.steps == so.steps == 1200
    assert st.current_time == pytest.approx(so.current_time, rel=1e-12)
    assert st.fitness == pytest.approx(so.fitness, rel=1e-6, abs=1e-12)
    b.close()


def test_tiled_divergence_stops_at_the_step_the_stream_path_does():
    """A huge force on one voxel tears its links apart: the simulation diverges in one
    tile; the repair launch runs it again to that step's link pass.  A second tiled
    simulation in the same launch is untouched."""
    bad = robot(7, 18)
    bad.externals = [Ext(len(bad.ijk) // 2, force=(1e7, 0, 0))]
    good = robot(8, 18)
    bt, st = run([bad, good], _abi.PATH_AUTO, 900, spl=300)
    assert all(bt.workgroups(s) > 1 for s in st)
    bs, ss = run([bad, good], _abi.PATH_STREAM, 900)
    pt, ps = bt.poll(), bs.poll()
    print("diverged at", ps[ss[0]].steps)
    assert ps[ss[0]].state == _abi.SIM_DIVERGED and ps[ss[0]].steps < 900
    assert pt[st[1]].state == _abi.SIM_RUNNING and pt[st[1]].steps == 900
    assert_same(bt, st, bs, ss, 1e-14, 1e-13)
    bt.close()
    bs.close()


def test_tiled_stop_condition_and_step_budget():
    # stop when t - 0.01 > 0, and a step budget of 610: both inside a launch
    a = robot(9, 16, steps=0)
    a.stop_tokens = [(MT["mtCONST"], 0.01), (MT["mtVAR"], 4.0), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)]
    c = robot(10, 16, steps=610)
    bt, st = run([a, c], _abi.PATH_AUTO, 2000, spl=400)
    assert all(bt.workgroups(s) > 1 for s in st)
    bs, ss = run([a, c], _abi.PATH_STREAM, 2000)
    pt = bt.poll()
    assert pt[st[0]].state == _abi.SIM_DONE and 0 < pt[st[0]].steps < 2000
    assert pt[st[1]].state == _abi.SIM_DONE and pt[st[1]].steps == 610
    assert_same(bt, st, bs, ss, 1e-14, 1e-13)
    bt.close()
    bs.close()


def test_several_launch_groups_and_a_mixed_batch():
    """Five 50k-voxel robots need more tiles than one cooperative launch holds; beside
    them a one-workgroup robot and a slab with collisions on the per-step kernels."""
    big = [robot(20 + i, 47) for i in range(5)]
    small = robot(30, 9)
    slab = workloads.slab(24, 24, 4, steps=0)
    models = big + [small, slab]
    bt, st = run(models, _abi.PATH_AUTO, 300, spl=150)
    wg = [bt.workgroups(s) for s in st]
    print("workgroups", wg)
    assert sum(wg[:5]) > 256 and wg[5] == 1 and wg[6] == 0
    bs, ss = run(models, _abi.PATH_STREAM, 300)
    assert_same(bt, st, bs, ss, 1e-14, 1e-13)
    bt.close()
    bs.close()


def test_one_workgroup_sims_inside_tiled_launches_equal_their_own_kernel():
    """While tiled simulations run, one-workgroup simulations run inside the same
    cooperative launches as one-tile entries: the same results, bit for bit, as on their
    own kernel."""
    small = [robot(40 + i, 9) for i in range(3)]
    small[1].stop_tokens = [(MT["mtCONST"], 0.01), (MT["mtVAR"], 4.0), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)]
    big = robot(50, 22)
    bm, sm = run(small + [big], _abi.PATH_AUTO, 1200, spl=400)
    assert [bm.workgroups(s) for s in sm[:3]] == [1, 1, 1] and bm.workgroups(sm[3]) > 1
    ba, sa = run(small, _abi.PATH_AUTO, 1200, spl=400)
    pm, pa = bm.poll(), ba.poll()
    for a, c in zip(sm[:3], sa):
        ga, gc = bm.read_state(a), ba.read_state(c)
        assert np.array_equal(ga["pos"], gc["pos"]) and np.array_equal(ga["orient"], gc["orient"])
        assert np.array_equal(ga["flags"], gc["flags"])
        assert pm[a].steps == pa[c].steps and pm[a].state == pa[c].state
        assert pm[a].fitness == pa[c].fitness and pm[a].current_time == pa[c].current_time
    assert pm[sm[1]].state == _abi.SIM_DONE and pm[sm[1]].steps < 1200
    bm.close()
    ba.close()


def test_single_steps_and_the_no_allocation_step_on_tiled_simulations():
    """doTimeStep (vx3h_sim_step) on a tiled simulation between tiled launches, and the
    no-allocation entry point (every simulation on the per-step kernels) on a batch with
    a tiled simulation: the same states as the per-step kernels throughout."""
    m = robot(60, 16)
    bt = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_AUTO, steps_per_launch=200)
    st = bt.admit(m)
    assert bt.workgroups(st) > 1
    bs = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    ss = bs.admit(m)
    bt.step(300)
    for _ in range(5):
        assert bt.do_time_step(st)
    bt.step(200)
    bt.step_noalloc(50)
    bt.step(250)
    bs.step(805)
    assert_same(bt, [st], bs, [ss], 1e-14, 1e-13)
    assert bt.poll()[st].steps == 805
    bt.close()
    bs.close()
