"""VXA / VXD text writers for `Model` descriptions.

The inverse of the engine's loader (vx3h_vxa_parse): a `Model` becomes a VXA
document in the layout VoxCAD and the gpuVoxels experiment scripts write
(Voxelyze3 `<VXA Version="1.1">` with Simulator, Environment and a VXC of version
0.94 whose structure layers are ASCII_READABLE, palette index + 48 per cell, see
Voxelyze3/src/VXA/VX_Object.cpp:1760-1830), and a set of replacements becomes a
VXD (children carrying replace="VXA.a.b", Voxelyze3/src/Utils/ctool.h:49-57).
MathTree token programs are written back as the element tree whose reversed
breadth-first order they are (VX3_SimulationManager.cu:157-275).
"""
from typing import Dict, List, Optional, Sequence
from xml.sax.saxutils import escape

import numpy as np

from .model import MT_OPS, MT_VARS, Model

_ARITY = {"mtEND": 0, "mtCONST": 0, "mtE": 0, "mtPI": 0, "mtVAR": 0}
for _op in ("mtADD", "mtSUB", "mtMUL", "mtDIV", "mtPOW", "mtGREATERTHAN", "mtLESSTHAN", "mtAND", "mtOR"):
    _ARITY[_op] = 2
for _op in ("mtSQRT", "mtSIN", "mtCOS", "mtTAN", "mtATAN", "mtLOG", "mtINT", "mtABS", "mtNOT", "mtNORMALCDF"):
    _ARITY[_op] = 1
_VAR_NAMES = {v: k for k, v in MT_VARS.items()}


def _attach_xml(m):
    """EnableAttach, SafetyGuard and the attach conditions (VX3_SimulationManager.cu:329-343)."""
    if not m.enable_attach:
        return ""
    conds = "".join("<Condition_%d>%s</Condition_%d>" % (k, mathtree_xml(t), k)
                    for k, t in enumerate(m.attach_conditions) if t)
    return ("<EnableAttach>1</EnableAttach><SafetyGuard>%d</SafetyGuard>" % m.safety_guard +
            ("<AttachCondition>%s</AttachCondition>" % conds if conds else ""))


def mathtree_xml(tokens: Sequence) -> str:
    """XML subtree of a token program given as (op, value) pairs ending with mtEND."""
    toks = [(MT_OPS[int(op)], float(v)) for op, v in tokens]
    if not toks or toks[-1][0] != "mtEND":
        raise ValueError("program must end with mtEND")
    bfs = list(reversed(toks[:-1]))
    if not bfs:
        return ""
    nodes = [{"op": op, "value": v, "kids": []} for op, v in bfs]
    queue, nxt = [nodes[0]], 1
    while queue:
        n = queue.pop(0)
        for _ in range(_ARITY[n["op"]]):
            if nxt >= len(nodes):
                raise ValueError("token program is not a complete tree")
            n["kids"].append(nodes[nxt])
            queue.append(nodes[nxt])
            nxt += 1
    if nxt != len(nodes):
        raise ValueError("token program has unused tokens")

    def emit(n):
        if n["op"] == "mtVAR":
            return "<mtVAR>%s</mtVAR>" % _VAR_NAMES[int(round(n["value"]))]
        if n["op"] == "mtCONST":
            return "<mtCONST>%r</mtCONST>" % n["value"]
        if not n["kids"]:
            return "<%s/>" % n["op"]
        return "<%s>%s</%s>" % (n["op"], "".join(emit(k) for k in n["kids"]), n["op"])

    return emit(nodes[0])


def _num(x) -> str:
    return repr(float(x))


def model_to_vxa(m: Model, regions: Optional[List[dict]] = None, extra_simulator: str = "") -> str:
    """VXA text of a Model.  Externals are not written; pass boundary-condition
    regions instead (dicts with PrimType, X, Y, Z, dX, dY, dZ, Radius, DofFixed,
    ForceX.. , TorqueX.., DisplaceX.., AngDisplaceX..)."""
    ijk = np.asarray(m.ijk, dtype=np.int64)
    lo = ijk.min(axis=0) if len(ijk) else np.zeros(3, np.int64)
    if (lo < 0).any():
        raise ValueError("lattice indices must be non-negative")
    dims = ijk.max(axis=0) + 1 if len(ijk) else np.ones(3, np.int64)
    XV, YV, ZV = (int(d) for d in dims)
    cells = np.zeros((ZV, YV, XV), dtype=np.int64)
    order = np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))
    if not np.array_equal(order, np.arange(len(ijk))):
        raise ValueError("voxels must be listed in structure-index order (x fastest, then y, then z)")
    cells[ijk[:, 2], ijk[:, 1], ijk[:, 0]] = np.asarray(m.mat) + 1
    if cells.max() + 48 > 255:
        raise ValueError("too many materials for ASCII_READABLE layers")
    zeta = m.materials[0] if m.materials else None
    out = ['<?xml version="1.0" encoding="ISO-8859-1"?>', '<VXA Version="1.1">', "<Simulator>",
           "<Integration><Integrator>0</Integrator><DtFrac>%s</DtFrac></Integration>" % _num(m.dt_frac)]
    if zeta is not None:
        out.append("<Damping><BondDampingZ>%s</BondDampingZ><ColDampingZ>%s</ColDampingZ><SlowDampingZ>%s</SlowDampingZ>"
                   "</Damping>" % (_num(zeta.zeta_internal), _num(zeta.zeta_collision), _num(zeta.zeta_global)))
    out.append("<Collisions><SelfColEnabled>%d</SelfColEnabled></Collisions>" % int(bool(m.collisions_enabled)))
    out.append("<Features><VolumeEffectsEnabled>%d</VolumeEffectsEnabled></Features>" % int(bool(m.volume_effects)))
    if m.stop_tokens or m.stop_condition_type:
        out.append("<StopCondition>%s%s</StopCondition>"
                   % ("<StopConditionFormula>%s</StopConditionFormula>" % mathtree_xml(m.stop_tokens) if m.stop_tokens
                      else "",
                      "<StopConditionType>%d</StopConditionType><StopConditionValue>%s</StopConditionValue>"
                      % (m.stop_condition_type, _num(m.stop_condition_value)) if m.stop_condition_type else ""))
    if m.equilibrium_mode:
        out.append("<EquilibriumMode><EquilibriumModeEnabled>1</EquilibriumModeEnabled></EquilibriumMode>")
    if m.fitness_tokens:
        out.append("<FitnessFunction>%s</FitnessFunction>" % mathtree_xml(m.fitness_tokens))
    out.append("<AttachDetach><EnableCollision>%d</EnableCollision><boundingRadius>%s</boundingRadius>"
               "<watchDistance>%s</watchDistance>%s</AttachDetach>"
               % (int(bool(m.collisions_enabled)), _num(m.bounding_radius), _num(m.watch_distance),
                  ("<EnableDetach>1</EnableDetach>" if m.enable_detach else "") + _attach_xml(m)))
    out.append("<EnableExpansion>%d</EnableExpansion>" % int(bool(m.enable_expansion)))
    out.append("<SavePositionOfAllVoxels>%d</SavePositionOfAllVoxels>" % int(bool(m.save_position_of_all_voxels)))
    if m.enable_cilia:
        out.append("<EnableCilia>1</EnableCilia>")
    if m.enable_signals:
        out.append("<EnableSignals>1</EnableSignals>")
    if m.secondary_experiment:
        out.append("<SecondaryExperiment>1</SecondaryExperiment><ReinitializeInitialPositionAfterThisManySeconds>%s"
                   "</ReinitializeInitialPositionAfterThisManySeconds>" % _num(m.reinit_after))

    if any(m.force_field):
        out.append("<ForceField>%s</ForceField>" % "".join(
            "<%s>%s</%s>" % (axis, mathtree_xml(toks), axis)
            for axis, toks in zip(("x_forcefield", "y_forcefield", "z_forcefield"), m.force_field) if toks))
    out.append(extra_simulator)
    out.append("</Simulator>")
    out.append("<Environment>")
    if regions:
        out.append("<Boundary_Conditions><NumBCs>%d</NumBCs>" % len(regions))
        for r in regions:
            out.append("<FRegion>" + "".join("<%s>%s</%s>" % (k, v, k) for k, v in r.items()) + "</FRegion>")
        out.append("</Boundary_Conditions>")
    grav_acc = getattr(m, "gravity_acc", None)
    if grav_acc is None:
        grav_acc = -float(m.gravity_mult) * 9.80665
    out.append("<Gravity><GravEnabled>%d</GravEnabled><GravAcc>%s</GravAcc><FloorEnabled>%d</FloorEnabled></Gravity>"
               % (int(m.gravity_mult != 0), _num(grav_acc), int(bool(m.floor_enabled))))
    out.append("<Thermal><TempEnabled>%d</TempEnabled><TempAmplitude>%s</TempAmplitude><TempBase>%s</TempBase>"
               "<VaryTempEnabled>%d</VaryTempEnabled><TempPeriod>%s</TempPeriod></Thermal>"
               % (int(bool(m.temp_enabled)), _num(m.temp_amplitude), _num(m.temp_base), int(bool(m.vary_temp_enabled)),
                  _num(m.temp_period)))
    out.append("</Environment>")
    out.append('<VXC Version="0.94"><Lattice><Lattice_Dim>%s</Lattice_Dim></Lattice>' % _num(m.voxel_size))
    out.append("<Voxel><Vox_Name>BOX</Vox_Name><X_Squeeze>1</X_Squeeze><Y_Squeeze>1</Y_Squeeze><Z_Squeeze>1</Z_Squeeze></Voxel>")
    out.append("<Palette>")
    for i, mt in enumerate(m.materials):
        mech = ["<MatModel>%d</MatModel>" % mt.mat_model,
                "<Elastic_Mod>%s</Elastic_Mod>" % _num(mt.elastic_mod),
                "<Plastic_Mod>%s</Plastic_Mod>" % _num(mt.plastic_mod),
                "<Yield_Stress>%s</Yield_Stress>" % _num(mt.yield_stress),
                "<Fail_Stress>%s</Fail_Stress>" % _num(mt.fail_stress),
                "<Density>%s</Density>" % _num(mt.density),
                "<Poissons_Ratio>%s</Poissons_Ratio>" % _num(mt.poissons_ratio),
                "<CTE>%s</CTE>" % _num(mt.cte),
                "<uStatic>%s</uStatic>" % _num(mt.u_static),
                "<uDynamic>%s</uDynamic>" % _num(mt.u_dynamic),
                "<isMeasured>%d</isMeasured>" % int(mt.is_measured),
                "<Fixed>%d</Fixed>" % int(mt.fixed),
                "<TurnOnThermalExpansionAfterThisManySeconds>%s</TurnOnThermalExpansionAfterThisManySeconds>"
                % _num(mt.thermal_on_after),
                "<Cilia>%s</Cilia>" % _num(mt.cilia),
                "<TurnOnCiliaAfterThisManySeconds>%s</TurnOnCiliaAfterThisManySeconds>" % _num(mt.cilia_on_after),
                "<RemoveFromSimulationAfterThisManySeconds>%s</RemoveFromSimulationAfterThisManySeconds>"
                % _num(mt.remove_after),
                "<isPaceMaker>%d</isPaceMaker><PaceMakerPeriod>%s</PaceMakerPeriod><signalValueDecay>%s</signalValueDecay>"
                "<signalTimeDelay>%s</signalTimeDelay><inactivePeriod>%s</inactivePeriod><isTarget>%d</isTarget>"
                "<Sticky>%d</Sticky>"
                % (int(mt.is_pacemaker), _num(mt.pacemaker_period), _num(mt.signal_value_decay),
                   _num(mt.signal_time_delay), _num(mt.inactive_period), int(mt.is_target), int(mt.sticky))]
        if mt.strain_data:
            mech.append("<SSData><NumDataPts>%d</NumDataPts><StrainData>%s</StrainData><StressData>%s</StressData></SSData>"
                        % (len(mt.strain_data), "".join("<Strain>%s</Strain>" % _num(e) for e in mt.strain_data),
                           "".join("<Stress>%s</Stress>" % _num(s) for s in mt.stress_data)))
        r, g, b, a = (_num(c) for c in mt.color)
        out.append('<Material ID="%d"><MatType>0</MatType><Name>%s</Name>'
                   "<Display><Red>%s</Red><Green>%s</Green><Blue>%s</Blue><Alpha>%s</Alpha></Display>"
                   "<Mechanical>%s</Mechanical></Material>" % (i + 1, escape("mat%d" % (i + 1)), r, g, b, a,
                                                             "".join(mech)))
    out.append("</Palette>")
    out.append('<Structure Compression="ASCII_READABLE"><X_Voxels>%d</X_Voxels><Y_Voxels>%d</Y_Voxels><Z_Voxels>%d</Z_Voxels>'
               % (XV, YV, ZV))
    out.append("<Data>")
    for z in range(ZV):
        out.append("<Layer><![CDATA[%s]]></Layer>" % "".join(chr(48 + int(c)) for c in cells[z].ravel()))
    out.append("</Data>")
    # PhaseOffset, BaseCiliaForce and ShiftCiliaForce are always written: when one is
    # absent, the reference's importer reads an uninitialized array pointer
    # (Voxelyze3/src/VXA/VX_Object.h:179,185,190,214-216)
    ph = np.zeros((ZV, YV, XV))
    if m.phase is not None:
        ph[ijk[:, 2], ijk[:, 1], ijk[:, 0]] = np.asarray(m.phase, dtype=np.float64)
    out.append("<PhaseOffset>")
    for z in range(ZV):
        out.append("<Layer><![CDATA[%s]]></Layer>" % ",".join(repr(float(v)) for v in ph[z].ravel()))
    out.append("</PhaseOffset>")
    for tag, vec in (("BaseCiliaForce", m.base_cilia), ("ShiftCiliaForce", m.shift_cilia)):
        cv = np.zeros((ZV, YV, XV, 3))
        if vec is not None:
            cv[ijk[:, 2], ijk[:, 1], ijk[:, 0]] = np.asarray(vec, dtype=np.float64).reshape(-1, 3)
        out.append("<%s>" % tag)
        for z in range(ZV):
            out.append("<Layer><![CDATA[%s]]></Layer>" % ",".join(repr(float(v)) for v in cv[z].ravel()))
        out.append("</%s>" % tag)
    out.append("</Structure></VXC></VXA>")
    return "\n".join(out)


def vxd(replacements: Dict[str, str], tags: Optional[Dict[str, str]] = None) -> str:
    """VXD text: each (path -> inner XML) pair becomes a child carrying replace=path."""
    tags = tags or {}
    body = []
    for path, inner in replacements.items():
        tag = tags.get(path, path.split(".")[-1])
        body.append('<%s replace="%s">%s</%s>' % (tag, escape(path), inner, tag))
    return '<?xml version="1.0" encoding="ISO-8859-1"?>\n<VXD>\n%s\n</VXD>\n' % "\n".join(body)
