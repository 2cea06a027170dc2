"""VX3 secondary experiment on the GPU engine against the oracle
(VX3_VoxelyzeKernel.cu:336-350, 365-397): timed removal of a material's voxels and
links, and the reset of the initial center of mass and positions.  FP64 positions
within 1e-9 m, flags equal, status totals and fitness equal to 1e-9 relative."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from gpuvoxels_amd.model import MT
from oracle_lib import OSim
from test_oracle_kat import _layered_cube

pytestmark = pytest.mark.gpu
REMOVED, INTERIOR = 16, 2


def check(models, steps):
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in models]
    b.step(steps)
    st = b.poll()
    for m, s in zip(models, slots):
        o = OSim(m)
        o.step(steps)
        g, r, so = b.read_state(s), o.state(), o.status()
        assert np.max(np.abs(g["pos"] - r["pos"])) < 1e-9
        assert np.array_equal(g["flags"] & (REMOVED | INTERIOR), r["flags"] & (REMOVED | INTERIOR))
        assert np.max(np.abs(np.array(st[s].initial_com[:]) - np.array(so.initial_com[:]))) < 1e-12
        assert st[s].total_distance_of_all_voxels == pytest.approx(so.total_distance_of_all_voxels, rel=1e-9)
        assert st[s].fitness == pytest.approx(so.fitness, rel=1e-9, abs=1e-15)
    return b, slots


def test_layered_cube_removal_and_reset():
    m = _layered_cube(60 * 4.5e-5, reinit_after=130 * 4.5e-5)
    m.fitness_tokens = [(MT["mtVAR"], 2.0), (MT["mtEND"], 0.0)]
    b, (s,) = check([m], 600)
    g = b.read_state(s)
    assert (g["flags"][18:] & REMOVED).all() and not (g["flags"][:18] & REMOVED).any()
    assert not np.array_equal(g["init_pos"], m.ijk * m.voxel_size)     # saved at the reset
    rec = b.history_record(s, 599, 1, 1)
    assert rec.split("<<<>>>")[0].count(";") == 18                       # removed voxels are not drawn
    assert rec.split("|[[[599]]]")[1].count(";") == 54 - 12 - 9           # minus the top layer's 12 links and its 9 down
    b.close()


def test_robots_losing_a_material():
    models = []
    for i, m in enumerate(workloads.robots(3, seed=31, grid=8, steps=0)):
        m.secondary_experiment = 1
        m.reinit_after = 0.002 * (i + 1)
        m.materials[i % 4].remove_after = 0.004
        models.append(m)
    b, slots = check(models, 1500)
    assert any((b.read_state(s)["flags"] & REMOVED).any() for s in slots)
    b.close()


def test_removal_with_collisions_like_the_oracle():
    """Voxel removal with collisions: the removed layer's voxels lose their links and the
    nearby lists are rebuilt from the links left."""
    m = _layered_cube(0.003)
    m.collisions_enabled = 1
    b, slots = check([m], 1500)
    assert (b.read_state(slots[0])["flags"] & REMOVED).any()
    b.close()
