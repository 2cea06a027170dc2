"""VXA documents exercising the loader: every case is built from this package's own
models and writer, so no reference input file is needed to reproduce them."""
import base64
import re

import numpy as np

from gpuvoxels_amd import _abi, vxa, workloads
from gpuvoxels_amd.model import MT, Material, Model


def _bc_cube():
    m = workloads.cube(6, steps=0)
    m.name = "cube_bcs"
    regions = [
        # bottom layer fixed (box), all six dofs
        dict(PrimType=0, X=0, Y=0, Z=0, dX=1, dY=1, dZ=0.1, DofFixed=63),
        # top corner pushed down (sphere), force shared by the voxels it touches
        dict(PrimType=2, X=1, Y=1, Z=1, Radius=0.3, DofFixed=0, ForceZ=-0.5, TorqueX=0.001),
        # a rod of voxels along x (cylinder) with a prescribed x displacement
        dict(PrimType=1, X=0, Y=0.5, Z=0.5, dX=1, dY=0, dZ=0, Radius=0.2, DofFixed=1, DisplaceX=0.0005),
        # overlaps the sphere: its voxels get the add-twice force rule
        dict(PrimType=0, X=0.5, Y=0.5, Z=0.8, dX=0.5, dY=0.5, dZ=0.2, DofFixed=0, ForceX=0.2, TorqueZ=0.002),
    ]
    return vxa.model_to_vxa(m, regions=regions)


def _data_material():
    m = workloads.cube(4, steps=0)
    m.name = "cube_data"
    m.materials = [workloads.base_material(mat_model=_abi.MDL_DATA, strain_data=[0.0, 0.01, 0.05, 0.2],
                                           stress_data=[0.0, 1e4, 3e4, 4e4], elastic_mod=1e6, matid=1)]
    m.gravity_mult = workloads.gravity_mult(-9.81)
    return vxa.model_to_vxa(m)


def _two_material_bilinear():
    m = workloads.slab(10, 6, 3, collisions=False)
    m.name = "slab_bilinear"
    return vxa.model_to_vxa(m)


def _base64_structure():
    m = workloads.robots(1, seed=7, grid=6, steps=0)[0]
    text = vxa.model_to_vxa(m)
    # re-encode the structure layers as BASE64 (the reader's default compression)
    def enc(match):
        raw = bytes(ord(c) - 48 for c in match.group(1))
        return "<Layer><![CDATA[%s]]></Layer>" % base64.b64encode(raw).decode()
    data = re.search(r"<Data>(.*?)</Data>", text, re.S).group(0)
    data2 = re.sub(r"<Layer><!\[CDATA\[([0-9:;<=>?@A-Z]*)\]\]></Layer>", enc, data)
    return text.replace('Compression="ASCII_READABLE"', 'Compression="BASE64"').replace(data, data2)


def _legacy_thermal_and_defaults():
    m = workloads.cube(3, steps=0)
    text = vxa.model_to_vxa(m)
    # legacy TempAmp (absolute) instead of TempAmplitude; no Integration/DtFrac; Damping without SlowDampingZ
    text = re.sub(r"<TempAmplitude>[^<]*</TempAmplitude>", "<TempAmp>31.5</TempAmp>", text)
    text = re.sub(r"<Integration>.*?</Integration>", "", text)
    text = re.sub(r"<SlowDampingZ>[^<]*</SlowDampingZ>", "", text)
    return text


def _robot_with_stop():
    m = workloads.robots(1, seed=9, steps=0)[0]
    # stop when t > 0.02: <mtGREATERTHAN><mtVAR>t</mtVAR><mtCONST>0.02</mtCONST></mtGREATERTHAN>
    m.stop_tokens = [(MT["mtCONST"], 0.02), (MT["mtVAR"], 4.0), (MT["mtGREATERTHAN"], 0.0), (MT["mtEND"], 0.0)]
    return vxa.model_to_vxa(m)


def _display_colors():
    # Display colours: in range, past 1 (clamped to 255), negative (clamped to 0),
    # a partial Display block (missing channels read 0.5) and none at all, and a
    # material without a Mechanical block (both blocks keep the previous material's values)
    m = workloads.robots(1, seed=11, grid=8, steps=0)[0]
    m.name = "colors"
    for mt, c in zip(m.materials, [(1.0, 0.2, 0.35, 0.8), (1.7, -0.25, 0.999, 0.0), (0.5, 0.5, 0.5, 1.0),
                                   (0.1, 0.3, 0.7, 0.9)]):
        mt.color = c
    text = vxa.model_to_vxa(m)
    mats = re.findall(r'<Material ID="\d+">.*?</Material>', text, re.S)
    text = text.replace(mats[2], re.sub(r"<Display>.*?</Display>", "", mats[2]))
    text = text.replace(mats[3], re.sub(r"<Green>[^<]*</Green><Blue>[^<]*</Blue>", "",
                                        re.sub(r"<Mechanical>.*?</Mechanical>", "", mats[3], flags=re.S)))
    # VoxCAD stop-condition keys, which the Voxelyze3 driver reads and ignores
    return text.replace("</Simulator>", "<StopCondition><StopConditionType>5</StopConditionType>"
                                        "<StopConditionValue>0.25</StopConditionValue></StopCondition>"
                                        "<EquilibriumMode><EquilibriumModeEnabled>1</EquilibriumModeEnabled>"
                                        "</EquilibriumMode></Simulator>")


def _cilia_and_force_field():
    # VX3 cilia: per-voxel base and shift forces on the structure, Cilia scale and
    # switch-on time per material; a short ShiftCiliaForce layer pads with zeros; and a
    # force field per axis
    m = workloads.robots(1, seed=17, grid=6, steps=0)[0]
    m.name = "cilia"
    rng = np.random.default_rng(5)
    m.enable_cilia = 1
    m.base_cilia = rng.normal(0, 1e-4, (m.n_voxels, 3))
    m.shift_cilia = rng.normal(0, 1e-4, (m.n_voxels, 3))
    for k, mt in enumerate(m.materials):
        mt.cilia = [0.0, 2.5, 8.0, 1.0][k]
        mt.cilia_on_after = [0.0, 0.0, 0.01, 0.0][k]
    m.force_field = [[(MT["mtCONST"], 1e-4), (MT["mtEND"], 0.0)], [],
                     [(MT["mtVAR"], 4.0), (MT["mtCONST"], -1e-3), (MT["mtMUL"], 0.0), (MT["mtEND"], 0.0)]]
    text = vxa.model_to_vxa(m)
    shift = re.search(r"<ShiftCiliaForce>.*?</ShiftCiliaForce>", text, re.S).group(0)
    layers = re.findall(r"<Layer><!\[CDATA\[(.*?)\]\]></Layer>", shift)
    short = shift.replace(layers[-1], ",".join(layers[-1].split(",")[:6]))   # whole cells: a partial one is out of range there
    return text.replace(shift, short)


CASES = {
    "robot": lambda: vxa.model_to_vxa(workloads.robots(2, seed=3, steps=0)[1]),
    "cube6": lambda: vxa.model_to_vxa(workloads.cube(6, steps=0)),
    "cube_bcs": _bc_cube,
    "cube_data": _data_material,
    "slab_bilinear": _two_material_bilinear,
    "robot_base64": _base64_structure,
    "legacy_thermal": _legacy_thermal_and_defaults,
    "robot_stop": _robot_with_stop,
    "colors": _display_colors,
    "cilia": _cilia_and_force_field,
}


def case(name):
    return CASES[name]()
