"""Edge cases of the resident (on-chip, multi-step) FP64 kernel against the oracle:
no links, externals with force, moment and partial rotation constraints, a
stress-strain-curve material, bilinear links that fail, material-fixed voxels, no
Poisson effects, bodies at the kernel shapes' capacities, and a body with more
distinct phase offsets than the kernel evaluates (which falls to the stream path)."""
import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from gpuvoxels_amd.model import Ext, Material, Model
from oracle_lib import OSim

pytestmark = pytest.mark.gpu
ALL = 63


def box(nx, ny, nz, **kw):
    m = workloads.cube(1, steps=0)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m.ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.int32)
    m.mat = np.zeros(len(m.ijk), np.int32)
    m.phase = np.zeros(len(m.ijk))
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def single():
    m = box(1, 1, 1)
    m.ijk[:, 2] += 2
    return m


def bar_with_externals():
    m = Model(voxel_size=0.01)
    m.materials = [Material(elastic_mod=1e6, density=1e3, zeta_internal=1.0, zeta_global=0.1)]
    m.ijk = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.int32)
    m.mat = np.zeros(3, np.int32)
    m.externals = [Ext(0, dof_fixed=ALL), Ext(2, dof_fixed=8, force=(0.02, 0.0, -0.01), moment=(0.0, 1e-4, 2e-4))]
    return m


def data_material():
    m = box(4, 4, 3)
    m.materials = [workloads.base_material(mat_model=_abi.MDL_DATA, strain_data=[0.0, 0.01, 0.05, 0.2],
                                           stress_data=[0.0, 1e4, 3e4, 4e4])]
    m.temp_amplitude, m.temp_period = 10.0, 0.01
    m.phase = (m.ijk[:, 0] % 2) * 0.5                 # neighbours in antiphase: large strains
    return m


def failing_bilinear():
    m = box(4, 3, 3)
    m.materials = [workloads.base_material(mat_model=_abi.MDL_BILINEAR, plastic_mod=2e5, yield_stress=1e3,
                                           fail_stress=3e3)]
    m.temp_amplitude, m.temp_period = 10.0, 0.01
    m.phase = (m.ijk[:, 0] % 2) * 0.5
    return m


def fixed_base():
    m = box(4, 4, 3)
    m.materials = [workloads.base_material(fixed=1, cte=0.0), workloads.base_material()]
    m.mat = (m.ijk[:, 2] > 0).astype(np.int32)
    return m


def no_poisson():
    m = workloads.robots(1, seed=3, grid=8, steps=0)[0]
    m.volume_effects = 0
    return m


CASES = {"single": single, "externals": bar_with_externals, "data": data_material, "bilinear_fail": failing_bilinear,
         "fixed_base": fixed_base, "no_poisson": no_poisson,
         "cap_256": lambda: box(8, 8, 4), "cap_512": lambda: box(8, 8, 8)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_resident_edge_case_matches_oracle(name):
    m = CASES[name]()
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT, steps_per_launch=300)
    s = b.admit(m)
    o = OSim(m)
    for _ in range(3):
        b.step(300)
        o.step(300)
        g, r = b.read_state(s), o.state()
        print(name, "max |dpos| %.3e" % np.max(np.abs(g["pos"] - r["pos"])))
        assert np.max(np.abs(g["pos"] - r["pos"])) < 1e-9
        assert np.max(np.abs(g["orient"] - r["orient"])) < 1e-8
        assert np.array_equal(g["flags"] & 8, r["flags"] & 8)
    assert b.poll()[s].steps == o.status().steps == 900
    b.close()


def test_many_phases_fall_to_the_stream_path():
    m = box(10, 10, 1)
    m.phase = np.linspace(0, 0.99, 100)          # 100 distinct offsets
    m.vary_temp_enabled = 1
    with pytest.raises(engine.VX3Error):
        engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_RESIDENT).admit(m)
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO)
    s = b.admit(m)
    b.step(600)
    o = OSim(m)
    o.step(600)
    assert np.max(np.abs(b.read_state(s)["pos"] - o.state()["pos"])) < 1e-9
    b.close()
