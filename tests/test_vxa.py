"""VXA/VXD loader and report writer (host side of the drop-in boundary).

The loader (vx3h_vxa_parse) is pinned against what the reference's own loader and
CPU engine make of the same documents: the golden vectors in tests/golden (written by
tests/golden/make_vxa_golden.py from oracle/_ref) and, where oracle/_ref is built, the
live reference.  Pinned per document: lattice indices, material indices, phase
offsets, per-voxel externals from boundary-condition regions, environment, derived
material constants, and the trajectory after 200 steps (bit-exact, FP64).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, vxa, workloads
from gpuvoxels_amd.model import FITNESS_XY_DISTANCE, MT
from oracle_lib import OSim, load, mathtree_eval, ref_available
from vxa_cases import CASES

HERE = os.path.dirname(os.path.abspath(__file__))
REF_TREE = "/root/reference"


def loader_vectors(text, steps=200):
    v = engine.Vxa(text)
    d = v.desc
    nv = d.n_voxels
    ijk = np.ctypeslib.as_array(d.ijk, (3 * nv,)).copy()
    mat = np.ctypeslib.as_array(d.material, (nv,)).copy()
    ph = np.ctypeslib.as_array(d.phase, (nv,)).copy()
    dof, tr, ro = np.zeros(nv, np.int32), np.zeros(3 * nv), np.zeros(3 * nv)
    fo, mo = np.zeros(3 * nv, np.float32), np.zeros(3 * nv, np.float32)
    for i in range(d.n_externals):
        e = d.externals[i]
        dof[e.voxel] = e.dof_fixed
        for k in range(3):
            tr[3 * e.voxel + k], ro[3 * e.voxel + k] = e.translate[k], e.rotate[k]
            fo[3 * e.voxel + k], mo[3 * e.voxel + k] = e.force[k], e.moment[k]
    color = np.array([list(d.materials[i].color) for i in range(d.n_materials)], np.int32).reshape(-1, 4)
    mcil = np.array([[d.materials[i].cilia, d.materials[i].cilia_on_after] for i in range(d.n_materials)]).reshape(-1, 2)
    cb = np.ctypeslib.as_array(d.base_cilia_force, (3 * nv,)).copy() if d.base_cilia_force else np.zeros(3 * nv)
    cs = np.ctypeslib.as_array(d.shift_cilia_force, (3 * nv,)).copy() if d.shift_cilia_force else np.zeros(3 * nv)
    mid = np.array([d.materials[i].matid for i in range(d.n_materials)], np.int32)
    # the trajectory pins the loader's physics fields against the reference's CPU engine,
    # which has no VX3 cilia or force field: those run separately (test_gpu_extras.py)
    saved = d.enable_cilia, list(d.n_force_field_tokens)
    d.enable_cilia = 0
    for k in range(3):
        d.n_force_field_tokens[k] = 0
    o = OSim.from_desc(d)
    d.enable_cilia = saved[0]
    for k in range(3):
        d.n_force_field_tokens[k] = saved[1][k]
    vmc = o.voxel_material_consts()
    lidx, lmc = o.link_material_consts()
    o.step(steps)
    st = o.state()
    return v, dict(nv=nv, nl=o.nl, ijk=ijk, mat=mat, phase=ph, ext_dof=dof, ext_trans=tr, ext_rot=ro, ext_force=fo,
                   ext_moment=mo, env_grav=np.float32(d.gravity_mult), env_floor=int(d.floor_enabled),
                   env_dtfrac=d.dt_frac, env_temp_amp=d.temp_amplitude, env_temp_period=d.temp_period, vmat_consts=vmc,
                   lmat_index=lidx, lmat_consts=lmc, mat_color=color, mat_id=mid, cilia_base=cb, cilia_shift=cs,
                   mat_cilia=mcil,
                   stop=np.array([v.info.stop_condition_type, v.info.stop_condition_value, v.info.equilibrium_mode],
                                 np.float64), state=st)


def check_against(ref, got):
    assert int(ref["nv"]) == got["nv"] and int(ref["nl"]) == got["nl"]
    for k in ("ijk", "mat", "phase", "ext_dof", "ext_trans", "ext_rot", "ext_force", "ext_moment", "vmat_consts",
              "lmat_index", "lmat_consts", "mat_color", "mat_id", "stop", "cilia_base", "cilia_shift", "mat_cilia"):
        assert np.array_equal(np.asarray(ref[k]), np.asarray(got[k])), k
    assert np.float32(ref["env_grav"]) == got["env_grav"]
    assert int(ref["env_floor"]) == got["env_floor"]
    assert float(ref["env_dtfrac"]) == got["env_dtfrac"]
    assert float(ref["env_temp_amp"]) == got["env_temp_amp"]
    assert float(ref["env_temp_period"]) == got["env_temp_period"]
    s = ref["sample"]
    assert np.array_equal(ref["pos"], got["state"]["pos"][s])
    assert np.array_equal(ref["orient"], got["state"]["orient"][s])
    assert float(ref["pos_sum"]) == float(got["state"]["pos"].sum())
    assert float(ref["orient_sum"]) == float(got["state"]["orient"].sum())


@pytest.mark.parametrize("name", sorted(CASES))
def test_loader_matches_reference_golden(name):
    ref = np.load(os.path.join(HERE, "golden", "vxa_%s.npz" % name), allow_pickle=False)
    _, got = loader_vectors(CASES[name]())
    check_against(ref, got)


@pytest.mark.ref
@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", ["cube_bcs", "robot"])
def test_loader_matches_reference_live(name):
    import golden.make_vxa_golden as g
    text = CASES[name]()
    ref = g.reference_vectors(text)
    _, got = loader_vectors(text)
    check_against(ref, got)


def test_boundary_conditions_resolved():
    # closed form for the cube_bcs document: bottom layer (z = 0) fully fixed; the
    # overlapping force regions give the add-twice rule of Import on shared voxels
    v, got = loader_vectors(CASES["cube_bcs"](), steps=0)
    ijk = got["ijk"].reshape(-1, 3)
    bottom = ijk[:, 2] == 0
    assert np.all(got["ext_dof"][bottom] & 63 == 63)
    assert not np.any(got["ext_dof"][~bottom] & 62)
    # the rod (cylinder along x through the centre) only fixes x, with the prescribed displacement
    rod = (got["ext_dof"] == 1)
    assert rod.sum() > 0 and np.all(got["ext_trans"].reshape(-1, 3)[rod, 0] == 0.0005)


def test_round_trip_of_model_fields():
    for m in (workloads.robots(1, seed=4, steps=0)[0], workloads.cube(4, steps=0)):
        v = engine.Vxa(vxa.model_to_vxa(m))
        d, (dm, keep) = v.desc, m.desc()
        assert d.n_voxels == m.n_voxels
        assert np.array_equal(np.ctypeslib.as_array(d.ijk, (3 * d.n_voxels,)), m.ijk.ravel())
        assert np.array_equal(np.ctypeslib.as_array(d.material, (d.n_voxels,)), m.mat)
        for f in ("voxel_size", "gravity_mult", "floor_enabled", "temp_enabled", "vary_temp_enabled", "temp_base",
                  "temp_amplitude", "temp_period", "enable_expansion", "dt_frac", "volume_effects", "collisions_enabled",
                  "bounding_radius", "watch_distance", "n_fitness_tokens", "save_position_of_all_voxels"):
            assert getattr(d, f) == getattr(dm, f), f
        assert d.max_steps == 1000000   # the reference driver's step cap
        for i in range(d.n_materials):
            a, b = d.materials[i], dm.materials[i]
            for f in ("mat_model", "elastic_mod", "density", "poissons_ratio", "cte", "u_static", "u_dynamic",
                      "is_measured", "fixed", "matid"):
                assert getattr(a, f) == getattr(b, f), f
            for f in ("zeta_internal", "zeta_global", "zeta_collision"):   # the reader's float damping values
                assert getattr(a, f) == float(np.float32(getattr(b, f))), f


def test_vxd_replaces_paths():
    m = workloads.cube(3, steps=0)
    base = vxa.model_to_vxa(m)
    patch = vxa.vxd({
        "VXA.Environment.Thermal": "<TempEnabled>1</TempEnabled><TempAmplitude>3.5</TempAmplitude><TempBase>20</TempBase>"
                                   "<VaryTempEnabled>0</VaryTempEnabled><TempPeriod>0.25</TempPeriod>",
        "VXA.Simulator.StopCondition": "<StopConditionFormula><mtSUB><mtVAR>t</mtVAR><mtCONST>0.5</mtCONST></mtSUB>"
                                       "</StopConditionFormula>",
        "VXA.Simulator.EnableExpansion": "1",
        "VXA.Simulator.AttachDetach.boundingRadius": "0.6",
    })
    d = engine.Vxa(base, patch).desc
    assert d.temp_amplitude == 3.5 and d.temp_base == 20 and d.temp_period == 0.25 and d.vary_temp_enabled == 0
    assert d.enable_expansion == 1 and d.bounding_radius == 0.6
    toks = [(d.stop_tokens[i].op, d.stop_tokens[i].value) for i in range(d.n_stop_tokens)]
    assert toks == [(MT["mtCONST"], 0.5), (MT["mtVAR"], 4.0), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)]
    # a VXD child without replace= and a path outside the document are ignored
    d2 = engine.Vxa(base, '<VXD><Thermal><TempBase>99</TempBase></Thermal>'
                          '<X replace="Other.Simulator">1</X></VXD>').desc
    assert d2.temp_base == m.temp_base


def test_record_history_keys():
    # VX3_SimulationManager.cu:326-367 reads RecordHistory.RecordStepSize [0],
    # RecordVoxel [1] and RecordLink [0]
    base = vxa.model_to_vxa(workloads.cube(3, steps=0))
    info = engine.Vxa(base).info
    assert (info.record_step_size, info.record_voxel, info.record_link) == (0, 1, 0)
    info = engine.Vxa(base, vxa.vxd({"VXA.Simulator.RecordHistory": "<RecordStepSize>100</RecordStepSize>"
                                                                    "<RecordLink>1</RecordLink>"})).info
    assert (info.record_step_size, info.record_voxel, info.record_link) == (100, 1, 1)
    info = engine.Vxa(base, vxa.vxd({"VXA.Simulator.RecordHistory": "<RecordStepSize>7</RecordStepSize>"
                                                                    "<RecordVoxel>0</RecordVoxel>"})).info
    assert (info.record_step_size, info.record_voxel, info.record_link) == (7, 0, 0)
    # the leaf form the experiment workflow writes (workflow.py:90-97) into a base without RecordHistory
    assert "RecordHistory" not in base
    info = engine.Vxa(base, '<VXD><RecordStepSize replace="VXA.Simulator.RecordHistory.RecordStepSize">100'
                            '</RecordStepSize></VXD>').info
    assert (info.record_step_size, info.record_voxel, info.record_link) == (100, 1, 0)


def _with_fitness(base, xml):
    return re.sub(r"<FitnessFunction>.*?</FitnessFunction>", "<FitnessFunction>%s</FitnessFunction>" % xml, base, flags=re.S)


def _fit_tokens(d):
    return [(d.fitness_tokens[i].op, d.fitness_tokens[i].value) for i in range(d.n_fitness_tokens)]


def test_mathtree_reversed_bfs():
    xml = "<mtSQRT><mtADD><mtMUL><mtVAR>x</mtVAR><mtVAR>x</mtVAR></mtMUL><mtMUL><mtVAR>y</mtVAR><mtVAR>y</mtVAR></mtMUL></mtADD></mtSQRT>"
    base = vxa.model_to_vxa(workloads.cube(2, steps=0))
    assert _fit_tokens(engine.Vxa(_with_fitness(base, xml)).desc) == FITNESS_XY_DISTANCE
    # replaced through a VXD, the element carries replace="...": as a property tree that is a
    # first child <xmlattr>/<replace>, the last breadth-first node, taken first and not a
    # token, so the program ends before the square root (VX3_SimulationManager.cu:165-273)
    d = engine.Vxa(base, vxa.vxd({"VXA.Simulator.FitnessFunction": xml})).desc
    assert _fit_tokens(d) == FITNESS_XY_DISTANCE[:7] + [(MT["mtEND"], 0.0)]
    # an unknown variable name leaves an mtVAR 0 (x) and ends the program there
    bad = xml.replace("<mtVAR>x</mtVAR><mtVAR>x</mtVAR>", "<mtVAR>x</mtVAR><mtVAR>w</mtVAR>")
    assert _fit_tokens(engine.Vxa(_with_fitness(base, bad)).desc) == [(MT["mtVAR"], 1.0), (MT["mtVAR"], 1.0),
                                                                     (MT["mtVAR"], 0.0), (MT["mtEND"], 0.0)]
    # a deeper program: (3 + pi) * sin(t) / 2 - |angle|, checked against the closed form
    xml2 = ("<mtSUB><mtDIV><mtMUL><mtADD><mtCONST>3</mtCONST><mtPI/></mtADD><mtSIN><mtVAR>t</mtVAR></mtSIN></mtMUL>"
            "<mtCONST>2</mtCONST></mtDIV><mtABS><mtVAR>angle</mtVAR></mtABS></mtSUB>")
    toks = _fit_tokens(engine.Vxa(_with_fitness(base, xml2)).desc)
    val = mathtree_eval(toks, t=0.7, angle=-0.3)
    assert val == pytest.approx((3 + np.pi) * np.sin(0.7) / 2 - 0.3, rel=1e-14)
    # the writer's inverse reproduces the element tree
    assert vxa.mathtree_xml(toks).replace("<mtCONST>3.0</mtCONST>", "<mtCONST>3</mtCONST>").replace(
        "<mtCONST>2.0</mtCONST>", "<mtCONST>2</mtCONST>") == xml2


def test_loader_errors():
    base = vxa.model_to_vxa(workloads.cube(2, steps=0))
    with pytest.raises(engine.VX3Error) as e:
        engine.Vxa("<VXA><Simulator></VXA>")
    assert e.value.status == _abi.EPARSE
    zl = base.replace('Compression="ASCII_READABLE"', 'Compression="ZLIB"')
    with pytest.raises(engine.VX3Error) as e:
        engine.Vxa(zl)
    assert e.value.status == _abi.EUNSUPPORTED
    short = re.sub(r"<Layer><!\[CDATA\[([0-9]*)\]\]></Layer>", lambda mm: "<Layer><![CDATA[%s]]></Layer>" % mm.group(1)[:-1],
                   base, count=1)
    with pytest.raises(engine.VX3Error) as e:
        engine.Vxa(short)
    assert e.value.status == _abi.EPARSE


def _nonzero_cells(text):
    st = re.search(r"<Structure[^>]*>(.*?)</Structure>", text, re.S).group(1)
    comp = re.search(r'<Structure\s+Compression="([^"]*)"', text)
    data = re.search(r"<Data>(.*?)</Data>", st, re.S).group(1)
    layers = re.findall(r"<Layer>\s*(?:<!\[CDATA\[)?(.*?)(?:\]\]>)?\s*</Layer>", data, re.S)
    if comp and comp.group(1) == "ASCII_READABLE":
        return sum(sum(1 for c in L.strip() if ord(c) - 48 > 0) for L in layers)
    return None


@pytest.mark.skipif(not os.path.isdir(REF_TREE), reason="reference tree not present")
def test_reference_vxa_files_load():
    # every VXA shipped with the reference either loads with one voxel per occupied
    # cell, or names a feature this engine rejects
    files = []
    for root, _, names in os.walk(REF_TREE):
        files += [os.path.join(root, n) for n in names if n.endswith(".vxa")]
    assert files
    loaded = 0
    for f in sorted(files):
        text = open(f, encoding="latin-1").read()
        try:
            v = engine.Vxa(text)
        except engine.VX3Error as e:
            # a structure cell naming a material past the palette indexes the
            # importer's material map out of range (VX_Sim.cpp:94-99): rejected here
            assert e.status == _abi.EUNSUPPORTED or (e.status == _abi.EPARSE and "not in the palette" in str(e)), \
                (f, str(e))
            continue
        n = _nonzero_cells(text)
        if n is not None:
            assert v.n_voxels == n, f
        loaded += 1
    assert loaded > 0


def _status(fit, nv=10, t=0.5, com0=(0, 0, 0), com=(1, 2, 3)):
    s = _abi.SimStatus()
    s.fitness, s.num_voxel, s.num_measured_voxel, s.current_time = fit, nv, nv - 1, t
    for k in range(3):
        s.initial_com[k], s.current_com[k] = com0[k], com[k]
    s.total_distance_of_all_voxels = 0.25
    return s


def test_report_xml_layout_and_order():
    entries = [("b.vxd", _status(0.5), 0.01, None), ("a.vxd", _status(float("nan")), 0.01, None),
               ("c.1.vxd", _status(2.0), 0.01, None)]
    x = engine.report_xml("gen_0001", entries)
    assert x.startswith('<?xml version="1.0" encoding="utf-8"?>\n<report><inputdir>gen_0001</inputdir>')
    assert "<bestfit><filename>c.1.vxd</filename><fitness_score>2</fitness_score></bestfit>" in x
    # sorted by fitness descending, NaN last; names cut at the first '.'
    assert x.index("<c>") < x.index("<b>") < x.index("<a>")
    assert "<a><currentTime>0.5</currentTime><fitness_score>nan</fitness_score><num_voxel>10</num_voxel>" in x
    assert "<num_measured_voxel>9</num_measured_voxel><voxSize>0.01</voxSize><numClosePairs>0</numClosePairs>" in x
    assert "<currentCenterOfMass><x>1</x><y>2</y><z>3</z></currentCenterOfMass>" in x
    # positions written with std::to_string (six decimals) when requested
    pos = (np.array([[0.0, 0.0, 0.0], [0.01, 0, 0]]), np.array([[0.001, 0.0, 0.0], [0.0111111, 0, 0]]), np.array([1, 2]))
    x = engine.report_xml("d", [("r.vxd", _status(1.0, nv=2), 0.01, pos)])
    assert "<init_pos>0.000000,0.000000,0.000000;0.010000,0.000000,0.000000;</init_pos>" in x
    assert "<pos>0.001000,0.000000,0.000000;0.011111,0.000000,0.000000;</pos><mats>1;2;</mats>" in x


def test_attach_keys_round_trip():
    # Sticky, EnableAttach, SafetyGuard and AttachCondition.Condition_k (VX3_SimulationManager.cu:329-343,
    # VXA/VX_Object.cpp:1436) reach the description, and the oracle run from it attaches as from the model
    import sys
    sys.path.insert(0, __import__("os").path.dirname(__file__))
    from test_oracle_kat import _sticky_bars, _links_between
    from gpuvoxels_amd.model import MT
    from gpuvoxels_amd.vxa import model_to_vxa
    cond = [[(MT["mtVAR"], 2.0), (MT["mtCONST"], 0.0115), (MT["mtSUB"], 0.0), (MT["mtEND"], 0.0)], [], [], [], []]
    m = _sticky_bars(7, conditions=cond, guard=77)
    v = engine.Vxa(model_to_vxa(m))
    d = v.desc
    assert d.enable_attach == 1 and d.safety_guard == 77 and d.n_attach_tokens[0] == 4
    assert [d.n_attach_tokens[k] for k in range(1, 5)] == [0, 0, 0, 0]
    assert d.materials[0].sticky == 1
    a, b = OSim(m), OSim.from_desc(d)
    a.step(3000)
    b.step(3000)
    assert _links_between(a, 7) == _links_between(b, 7) == [(0, 7, 2), (3, 10, 2), (6, 13, 2)]
    assert np.array_equal(a.state()["pos"], b.state()["pos"])
