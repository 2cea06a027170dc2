"""The 64x64x32 bilinear slab with self-collision at full size (131,072 voxels,
385,024 links in one simulation, the stream path) against the CPU oracle."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim

pytestmark = pytest.mark.gpu


def test_full_size_slab_f64():
    m = workloads.slab(64, 64, 32, steps=0)
    o = OSim(m)
    b = engine.Batch(precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s = b.admit(m)
    o.step(100)
    b.step(100)
    ost, gst = o.status(), b.poll()[s]
    assert gst.state == ost.state == _abi.SIM_RUNNING and gst.steps == ost.steps == 100
    assert np.array_equal(o.collision_pairs(), b.collision_pairs(s))
    g, c = b.read_state(s), o.state()
    assert np.max(np.abs(g["pos"] - c["pos"])) / 0.64 < 1e-9      # relative to the slab's 0.64 m span
    # rounding-level: the device's transcendental functions differ from glibc's in the
    # last ulp, which grows to ~1e-9 in the orientation of a few of the 131k voxels
    assert np.max(np.abs(g["orient"] - c["orient"])) < 1e-8
    assert np.array_equal(g["flags"] & 8, c["flags"] & 8)           # floor static-friction states
    assert abs(gst.current_com[2] - ost.current_com[2]) < 1e-12
