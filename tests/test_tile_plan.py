"""The tile planner of the tiled resident kernel (gpuvoxels_amd/csrc/vx3h_tile.cpp), on
the CPU: every voxel owned by one tile, every tile within its LDS slots (owned + halo)
and link slots per axis, every link with an owned end computed by that tile, its end
slots pointing at the right voxels, written back by exactly one tile, and the publish
bits set on exactly the voxels another tile reads (tests/tile_plan_check.cpp)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gpuvoxels_amd import engine, workloads
from gpuvoxels_amd.model import Material, Model

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "gpuvoxels_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = str(tmp_path_factory.mktemp("tile") / "tile_plan_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "tile_plan_check.cpp"),
                    os.path.join(CSRC, "vx3h_tile.cpp")], check=True)
    return exe


def lattice(ijk):
    m = Model(voxel_size=0.01)
    m.materials = [Material()]
    m.ijk = np.asarray(ijk, np.int32)
    m.mat = np.zeros(len(m.ijk), np.int32)
    return m


def plan(checker, m, vmax=1024, lax=512, max_tiles=256):
    vneg, vpos, axis, _ = engine.topology(m)
    lines = ["%d %d %d %d %d" % (len(m.ijk), len(vneg), vmax, lax, max_tiles)]
    lines += ["%d %d %d" % tuple(r) for r in m.ijk]
    lines += ["%d %d %d" % (a, b, c) for a, b, c in zip(vneg, vpos, axis)]
    r = subprocess.run([checker], input="\n".join(lines) + "\n", stdout=subprocess.PIPE, text=True, check=True, timeout=120)
    return r.stdout.split()


@pytest.mark.parametrize("grid,seed", [(14, 1), (22, 2), (30, 3), (47, 4)])
def test_robot_plans_hold_their_invariants(checker, grid, seed):
    m = workloads.robot(np.random.default_rng(seed), grid=grid, steps=0)
    out = plan(checker, m)
    assert out[0] == "ok", out
    tiles, owned = int(out[1]), int(out[2])
    assert tiles >= 2
    assert owned / tiles >= 0.4 * 1024   # blocks are filled, not halved down to crumbs


def test_dense_and_flat_bodies(checker):
    z, y, x = np.meshgrid(np.arange(20), np.arange(20), np.arange(20), indexing="ij")
    cube = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    out = plan(checker, lattice(cube))   # 8000 voxels, 3 links per voxel: the per-axis link limit binds
    assert out[0] == "ok", out
    z, y, x = np.meshgrid(np.arange(3), np.arange(80), np.arange(80), indexing="ij")
    out = plan(checker, lattice(np.stack([x.ravel(), y.ravel(), z.ravel()], 1)))
    assert out[0] == "ok", out
    line = np.stack([np.arange(5000), np.zeros(5000, int), np.zeros(5000, int)], 1)
    out = plan(checker, lattice(line))
    assert out[0] == "ok" and int(out[3]) == 2 * (int(out[1]) - 1), out   # a chain: two halo voxels per cut


def test_too_many_tiles_is_refused(checker):
    z, y, x = np.meshgrid(np.arange(20), np.arange(20), np.arange(20), indexing="ij")
    out = plan(checker, lattice(np.stack([x.ravel(), y.ravel(), z.ravel()], 1)), max_tiles=4)
    assert out == ["none"]
