"""The C ABI (include/vx3h.h) without a GPU: libvx3h.so loads and exports every
function the header declares, the ctypes mirrors in gpuvoxels_amd/_abi.py have the
C layout (sizes and field offsets from a program compiled against the header), and
the host-only entry points behave."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vx3h.h")

STRUCTS = {
    "vx3h_material_desc": _abi.MaterialDesc, "vx3h_external": _abi.External, "vx3h_token": _abi.Token,
    "vx3h_sim_desc": _abi.SimDesc, "vx3h_batch_cfg": _abi.BatchCfg, "vx3h_sim_status": _abi.SimStatus,
    "vx3h_state_view": _abi.StateView, "vx3h_topology_view": _abi.TopologyView,
    "vx3h_report_entry": _abi.ReportEntry, "vx3h_vxa_info": _abi.VxaInfo, "vx3h_server_cfg": _abi.ServerCfg,
    "vx3h_server_result": _abi.ServerResult, "vx3h_tune_trial": _abi.TuneTrial,
}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vx3h_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(engine.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 28
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_ctypes_layout_matches_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True,
                                                 check=True).stdout.split("\n") if l)
    for cname, py in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got["%s.%s" % (cname, fname)]) == getattr(py, fname).offset, (cname, fname)


def test_constants_match_header():
    text = open(HEADER).read()
    for name, val in [("VX3H_SIM_FAILED", _abi.SIM_FAILED), ("VX3H_MDL_BILINEAR", _abi.MDL_BILINEAR),
                      ("VX3H_MDL_DATA", _abi.MDL_DATA), ("VX3H_MAX_CURVE", 16), ("VX3H_MAX_TOKENS", 256),
                      ("VX3H_SC_MIN_KE", _abi.SC_MIN_KE), ("VX3H_SC_MIN_MAXMOVE", _abi.SC_MIN_MAXMOVE)]:
        m = re.search(r"\b%s\s*=?\s*(\d+)" % name, text)
        assert m and int(m.group(1)) == val, name


def topology(model):
    lib = engine.lib()
    d, keep = model.desc()
    tv = _abi.TopologyView()
    assert lib.vx3h_build_topology(C.byref(d), C.byref(tv)) == 0
    nv, nl = tv.n_voxels, tv.n_links
    arrs = [np.zeros(nl, np.int32) for _ in range(3)] + [np.zeros(6 * nv, np.int32)]
    tv.link_vneg, tv.link_vpos, tv.link_axis, tv.voxel_links = [a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs]
    assert lib.vx3h_build_topology(C.byref(d), C.byref(tv)) == 0
    return nv, nl, arrs


@pytest.mark.parametrize("nx,ny,nz", [(6, 6, 6), (64, 64, 32), (1, 1, 1), (5, 1, 1)])
def test_box_topology_closed_form(nx, ny, nz):
    """A full nx*ny*nz box has (nx-1)ny nz + nx(ny-1)nz + nx ny(nz-1) links; each link
    joins lattice neighbours along its axis, and each voxel's six slots name the
    links on its faces."""
    m = workloads.slab(nx, ny, nz, steps=0)
    nv, nl, (vneg, vpos, axis, vl) = topology(m)
    assert nv == nx * ny * nz
    assert nl == (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    ijk = m.ijk
    d = ijk[vpos] - ijk[vneg]
    assert np.array_equal(d, np.eye(3, dtype=np.int32)[axis])
    vl = vl.reshape(-1, 6)
    for k in range(3):
        has_pos = vl[:, 2 * k] >= 0
        assert np.array_equal(has_pos, ijk[:, k] < (nx, ny, nz)[k] - 1)
        assert np.array_equal(vneg[vl[has_pos, 2 * k]], np.nonzero(has_pos)[0])
        has_neg = vl[:, 2 * k + 1] >= 0
        assert np.array_equal(vpos[vl[has_neg, 2 * k + 1]], np.nonzero(has_neg)[0])


def test_topology_matches_oracle_order():
    """Link numbering follows the reference's creation order (CVoxelyze::setVoxel adds
    a voxel's links X_POS..Z_NEG as it is added, Voxelyze3/src/old/Voxelyze.cpp:313-341)."""
    from oracle_lib import OSim
    for m in workloads.robots(4, seed=3, grid=8, steps=0):
        nv, nl, (vneg, vpos, axis, vl) = topology(m)
        ov, op, oa, ovl = OSim(m).topology()
        assert np.array_equal(vneg, ov) and np.array_equal(vpos, op) and np.array_equal(axis, oa)
        assert np.array_equal(vl.reshape(-1, 6), ovl)


def test_largest_component_closed_form():
    body = np.zeros((3, 3, 3), np.int32)
    body[0, 0, 0] = 1          # singleton
    body[2, :, :] = 1          # a full 3x3 layer: the largest
    body[1, 1, 1] = 1          # joins the layer through a face
    out = workloads.largest_component(body)
    want = np.zeros_like(body)
    want[2] = 1
    want[1, 1, 1] = 1
    assert np.array_equal(out != 0, want != 0)


def test_footprint_and_errors():
    lib = engine.lib()
    m = workloads.cube(6, steps=0)
    d, keep = m.desc()
    b32, b64 = C.c_size_t(), C.c_size_t()
    assert lib.vx3h_sim_footprint(C.byref(d), _abi.PREC_F32, C.byref(b32)) == 0
    assert lib.vx3h_sim_footprint(C.byref(d), _abi.PREC_F64, C.byref(b64)) == 0
    assert 0 < b32.value < b64.value
    bad = workloads.cube(3, steps=0)
    bad.mat[4] = 7
    d, keep = bad.desc()
    assert lib.vx3h_sim_footprint(C.byref(d), _abi.PREC_F32, C.byref(b32)) == _abi.EINVAL
    assert b"material index out of range" in lib.vx3h_last_error()
    bad = workloads.cube(3, steps=0)
    bad.ijk[5] = bad.ijk[2]
    d, keep = bad.desc()
    assert lib.vx3h_sim_footprint(C.byref(d), _abi.PREC_F32, C.byref(b32)) == _abi.EINVAL
    assert b"duplicate voxel position" in lib.vx3h_last_error()


def test_cpp_facade_compiles_and_links(tmp_path):
    """VX3_VoxelyzeKernel (doTimeStep/StopConditionMet) and VX3_SimulationManager, the
    reference's class names over the C ABI, compile and link against libvx3h.so."""
    import subprocess
    from test_gpu_facade import CPP, ROOT as R
    src = tmp_path / "facade.cpp"
    src.write_text(CPP)
    lib = os.path.join(R, "gpuvoxels_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(tmp_path / "facade"), str(src), "-I" + R, "-L" + lib, "-lvx3h",
                    "-Wl,-rpath," + lib], check=True)
