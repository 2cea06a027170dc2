"""VX3 attach on the GPU engine against the oracle (handle_collision_attachment,
VX3_VoxelyzeKernel.cu:724-826; VX3_Link.cu:30-55, 206-213): links made between touching
sticky voxels, in the same order, on the same faces, with the same guard.  FP64
positions within 1e-9 m, link counts and interior flags equal, link strains within 1e-6."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from gpuvoxels_amd.model import MT
from oracle_lib import OSim
from test_oracle_kat import _sticky_bars

pytestmark = pytest.mark.gpu
INTERIOR = 2


def check(models, steps):
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    slots = [b.admit(m) for m in models]
    b.step(steps)
    st = b.poll()
    for m, s in zip(models, slots):
        o = OSim(m)
        nl0 = o.nl
        o.step(steps)
        g, r = b.read_state(s, nl=o.nl), o.state()
        assert st[s].n_links == o.nl
        assert np.max(np.abs(g["pos"] - r["pos"])) < 1e-9
        assert np.array_equal(g["flags"] & INTERIOR, r["flags"] & INTERIOR)
        if o.nl:
            assert np.max(np.abs(g["strain"] - o.link_state()["strain"])) < 1e-6
        if m.enable_attach and len(set(m.mat.tolist())) == 1:
            assert o.nl > nl0   # one sticky material: the case attached something
        else:
            assert o.nl == nl0
    return b, slots


def test_bars_attach_like_the_oracle():
    cond = [[(MT["mtVAR"], 2.0), (MT["mtCONST"], 0.0115), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)], [], [], [], []]
    models = [_sticky_bars(7), _sticky_bars(7, conditions=cond), _sticky_bars(1, guard=0),
              _sticky_bars(5, sticky=(1, 2))] + workloads.robots(1, seed=5, grid=6, steps=0)
    b, slots = check(models, 3000)
    rec = b.history_record(slots[0], 2999, 1, 1)
    assert rec.split("|[[[2999]]]")[1].count(";") == 12 + 3     # the attached links are drawn
    b.close()


def test_sticky_robots_fall_onto_a_sticky_slab():
    # a 4x4 slab on the floor and a robot of the same sticky material dropped onto it
    models = []
    for seed in range(2):
        r = workloads.robot(np.random.default_rng(40 + seed), grid=4, p=0.7, steps=0)
        x, y = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
        slab = np.stack([x.ravel(), y.ravel(), 0 * x.ravel()], 1)
        top = r.ijk + np.array([0, 0, 2])
        m = _sticky_bars(1)
        m.ijk = np.concatenate([slab, top]).astype(np.int32)
        m.mat = np.zeros(len(m.ijk), np.int32)
        m.phase = None
        models.append(m)
    check(models, 4000)


def test_refusals():
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT)
    with pytest.raises(engine.VX3Error) as e:
        b.admit(_sticky_bars(3))
    assert e.value.status == _abi.EUNSUPPORTED
    b.close()


def test_attach_with_collisions_like_the_oracle():
    """Sticky bars with collisions: contacts push the touching voxels while links are
    made; after each attachment the nearby lists (and so the contact pairs) follow the
    new links."""
    models = []
    for v in (1, 3):
        m = _sticky_bars(v)
        m.collisions_enabled = 1
        models.append(m)
    check(models, 4000)
