"""Host-side simulation descriptions.

`Model` is the Python form of a `vx3h_sim_desc`: materials, voxels in insertion
order, phase offsets, externals and the environment/simulator switches that
CVX_Sim::Import and VX3_SimulationManager::readVXD consume
(Voxelyze3/src/VXA/VX_Sim.cpp:59-153, Voxelyze3/src/VX3/VX3_SimulationManager.cu:323-367).
`Model.desc()` packs it into the C structure; the returned object keeps every
array alive for as long as the structure is in use.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi


@dataclass
class Material:
    elastic_mod: float = 1e6
    density: float = 1e3
    poissons_ratio: float = 0.0
    cte: float = 0.0
    u_static: float = 0.0
    u_dynamic: float = 0.0
    zeta_internal: float = 1.0
    zeta_global: float = 0.0
    zeta_collision: float = 0.0
    mat_model: int = _abi.MDL_LINEAR
    plastic_mod: float = 0.0
    yield_stress: float = 0.0
    fail_stress: float = 0.0
    strain_data: List[float] = field(default_factory=list)
    stress_data: List[float] = field(default_factory=list)
    is_measured: int = 1
    fixed: int = 0
    thermal_on_after: float = 0.0
    matid: int = 0
    color: tuple = (0.5, 0.5, 0.5, 1.0)   # Display Red, Green, Blue, Alpha in [0, 1]
    cilia: float = 0.0                    # scale of the cilia force of voxels of this material
    cilia_on_after: float = 0.0           # TurnOnCiliaAfterThisManySeconds
    remove_after: float = 0.0             # RemoveFromSimulationAfterThisManySeconds (secondary experiment)
    is_pacemaker: int = 0                 # signals: fires 100 every pacemaker_period
    pacemaker_period: float = 0.0
    signal_value_decay: float = 0.9
    signal_time_delay: float = 0.03
    inactive_period: float = 0.03
    is_target: int = 0                    # contacts with other materials count as hits
    sticky: int = 0                       # surface voxels of this material attach to each other

    def to_c(self) -> _abi.MaterialDesc:
        m = _abi.MaterialDesc()
        m.mat_model = self.mat_model
        m.elastic_mod = self.elastic_mod
        m.plastic_mod = self.plastic_mod
        m.yield_stress = self.yield_stress
        m.fail_stress = self.fail_stress
        m.density = self.density
        m.poissons_ratio = self.poissons_ratio
        m.cte = self.cte
        m.u_static = self.u_static
        m.u_dynamic = self.u_dynamic
        m.zeta_internal = self.zeta_internal
        m.zeta_global = self.zeta_global
        m.zeta_collision = self.zeta_collision
        m.n_data = len(self.strain_data)
        for i, (e, s) in enumerate(zip(self.strain_data, self.stress_data)):
            m.strain_data[i] = e
            m.stress_data[i] = s
        m.is_measured = self.is_measured
        m.fixed = self.fixed
        m.thermal_on_after = self.thermal_on_after
        m.matid = self.matid
        m.cilia = self.cilia
        m.cilia_on_after = self.cilia_on_after
        m.remove_after = self.remove_after
        m.is_pacemaker = self.is_pacemaker
        m.pacemaker_period = self.pacemaker_period
        m.signal_value_decay = self.signal_value_decay
        m.signal_time_delay = self.signal_time_delay
        m.inactive_period = self.inactive_period
        m.is_target = self.is_target
        m.sticky = self.sticky
        for k in range(4):   # CVXC_Material::GetRedi etc. then CVX_Material::setRed's clamp
            m.color[k] = min(255, max(0, int(np.float32(self.color[k]) * np.float32(255))))
        return m


@dataclass
class Ext:
    voxel: int
    dof_fixed: int = 0
    translate: tuple = (0.0, 0.0, 0.0)
    rotate: tuple = (0.0, 0.0, 0.0)
    force: tuple = (0.0, 0.0, 0.0)
    moment: tuple = (0.0, 0.0, 0.0)


# MathTree token opcodes (Voxelyze3/src/Utils/VX3_MathTree.h:4-29)
MT_OPS = ["mtEND", "mtCONST", "mtE", "mtPI", "mtVAR", "mtADD", "mtSUB", "mtMUL", "mtDIV", "mtPOW", "mtSQRT", "mtSIN",
          "mtCOS", "mtTAN", "mtATAN", "mtLOG", "mtINT", "mtABS", "mtNOT", "mtGREATERTHAN", "mtLESSTHAN", "mtAND", "mtOR",
          "mtNORMALCDF"]
MT = {n: i for i, n in enumerate(MT_OPS)}
MT_VARS = {"x": 0, "y": 1, "z": 2, "hit": 3, "t": 4, "angle": 5, "targetCloseness": 6, "numClosePairs": 7, "num_voxel": 8}


def tokens(*toks):
    """Builds a token program from (op, value) pairs, e.g. tokens(("mtVAR", "x"))."""
    out = []
    for op, *v in toks:
        val = 0.0
        if v:
            val = float(MT_VARS[v[0]]) if op == "mtVAR" else float(v[0])
        out.append((MT[op], val))
    out.append((MT["mtEND"], 0.0))
    return out


# sqrt(x^2 + y^2): a CoM-displacement fitness in the reversed-BFS token order the
# reference's ParseMathTree produces for <mtSQRT><mtADD><mtMUL>x x</mtMUL><mtMUL>y y</mtMUL></mtADD></mtSQRT>
FITNESS_XY_DISTANCE = [(MT["mtVAR"], 1.0), (MT["mtVAR"], 1.0), (MT["mtVAR"], 0.0), (MT["mtVAR"], 0.0),
                       (MT["mtMUL"], 0.0), (MT["mtMUL"], 0.0), (MT["mtADD"], 0.0), (MT["mtSQRT"], 0.0),
                       (MT["mtEND"], 0.0)]


# |(x, y, z)|: 3D CoM displacement, token order of
# <mtSQRT><mtADD><mtADD><mtMUL>x x</mtMUL><mtMUL>y y</mtMUL></mtADD><mtMUL>z z</mtMUL></mtADD></mtSQRT>
FITNESS_DISTANCE_3D = [(MT["mtVAR"], 1.0), (MT["mtVAR"], 1.0), (MT["mtVAR"], 0.0), (MT["mtVAR"], 0.0),
                       (MT["mtVAR"], 2.0), (MT["mtVAR"], 2.0), (MT["mtMUL"], 0.0), (MT["mtMUL"], 0.0),
                       (MT["mtMUL"], 0.0), (MT["mtADD"], 0.0), (MT["mtADD"], 0.0), (MT["mtSQRT"], 0.0),
                       (MT["mtEND"], 0.0)]


@dataclass
class Model:
    voxel_size: float = 0.001
    materials: List[Material] = field(default_factory=list)
    ijk: Optional[np.ndarray] = None          # (V, 3) int32
    mat: Optional[np.ndarray] = None          # (V,) int32
    phase: Optional[np.ndarray] = None        # (V,) float64 (cycles)
    externals: List[Ext] = field(default_factory=list)
    gravity_mult: float = 0.0
    gravity_acc: Optional[float] = None       # GravAcc the multiplier came from (VXA writer), if known
    floor_enabled: int = 0
    temp_enabled: int = 0
    vary_temp_enabled: int = 0
    temp_base: float = 25.0
    temp_amplitude: float = 0.0
    temp_period: float = 0.1
    enable_expansion: int = 0
    dt_frac: float = 0.9
    dt_fixed: float = 0.0
    volume_effects: int = 0
    collisions_enabled: int = 0
    bounding_radius: float = 0.75
    watch_distance: float = 1.0
    stop_tokens: list = field(default_factory=list)
    fitness_tokens: list = field(default_factory=list)
    max_steps: int = 0
    save_position_of_all_voxels: int = 0
    # VoxCAD stop-condition family (SC_* of _abi) and equilibrium mode; the Voxelyze3
    # driver ignores them, so they are off unless set here
    stop_condition_type: int = 0
    stop_condition_value: float = 0.0
    equilibrium_mode: int = 0
    # VX3 cilia (surface voxels pushed along base_cilia rotated with the voxel) and
    # force field (a MathTree token list per axis, empty = none)
    enable_cilia: int = 0
    base_cilia: Optional[np.ndarray] = None   # (V, 3) N, voxel frame
    shift_cilia: Optional[np.ndarray] = None  # (V, 3)
    force_field: list = field(default_factory=lambda: [[], [], []])
    enable_detach: int = 0                    # failed links leave their voxels (VX3 detach)
    secondary_experiment: int = 0             # timed voxel removal and initial-position reset
    reinit_after: float = 0.0                 # ReinitializeInitialPositionAfterThisManySeconds
    enable_signals: int = 0                   # VX3 signals between linked voxels
    enable_attach: int = 0                    # touching sticky voxels get new links (VX3 attach)
    attach_conditions: list = field(default_factory=lambda: [[], [], [], [], []])   # MathTrees, all > 0
    safety_guard: int = 500                   # steps a new link carries 1/100 of its force
    init_velocity: Optional[np.ndarray] = None   # (V, 3) m/s: initial linear velocity per voxel
    max_dist_close_pairs: float = 0.0         # MaxDistInVoxelLengthsToCountAsPair (target pairs)
    name: str = "sim"

    @property
    def n_voxels(self) -> int:
        return 0 if self.ijk is None else int(self.ijk.shape[0])

    def desc(self):
        """Returns (desc, keepalive)."""
        d = _abi.SimDesc()
        keep = []
        mats = (_abi.MaterialDesc * len(self.materials))(*[m.to_c() for m in self.materials])
        keep.append(mats)
        d.voxel_size = self.voxel_size
        d.n_materials = len(self.materials)
        d.materials = mats
        ijk = np.ascontiguousarray(self.ijk, dtype=np.int32).reshape(-1)
        mat = np.ascontiguousarray(self.mat, dtype=np.int32)
        keep += [ijk, mat]
        d.n_voxels = self.n_voxels
        d.ijk = ijk.ctypes.data_as(C.POINTER(C.c_int))
        d.material = mat.ctypes.data_as(C.POINTER(C.c_int))
        if self.phase is not None:
            ph = np.ascontiguousarray(self.phase, dtype=np.float64)
            keep.append(ph)
            d.phase = ph.ctypes.data_as(C.POINTER(C.c_double))
        if self.externals:
            ex = (_abi.External * len(self.externals))()
            for i, e in enumerate(self.externals):
                ex[i].voxel = e.voxel
                ex[i].dof_fixed = e.dof_fixed
                for k in range(3):
                    ex[i].translate[k] = e.translate[k]
                    ex[i].rotate[k] = e.rotate[k]
                    ex[i].force[k] = e.force[k]
                    ex[i].moment[k] = e.moment[k]
            keep.append(ex)
            d.n_externals = len(self.externals)
            d.externals = ex
        d.gravity_mult = self.gravity_mult
        d.floor_enabled = self.floor_enabled
        d.temp_enabled = self.temp_enabled
        d.vary_temp_enabled = self.vary_temp_enabled
        d.temp_base = self.temp_base
        d.temp_amplitude = self.temp_amplitude
        d.temp_period = self.temp_period
        d.enable_expansion = self.enable_expansion
        d.dt_frac = self.dt_frac
        d.dt_fixed = self.dt_fixed
        d.volume_effects = self.volume_effects
        d.collisions_enabled = self.collisions_enabled
        d.bounding_radius = self.bounding_radius
        d.watch_distance = self.watch_distance
        for name, toks in (("stop", self.stop_tokens), ("fitness", self.fitness_tokens)):
            if toks:
                arr = (_abi.Token * len(toks))(*[_abi.Token(op, val) for op, val in toks])
                keep.append(arr)
                setattr(d, "n_%s_tokens" % name, len(toks))
                setattr(d, "%s_tokens" % name, arr)
        d.max_steps = self.max_steps
        d.save_position_of_all_voxels = self.save_position_of_all_voxels
        d.stop_condition_type = self.stop_condition_type
        d.stop_condition_value = self.stop_condition_value
        d.equilibrium_mode = self.equilibrium_mode
        d.enable_cilia = self.enable_cilia
        for name, arr in (("base_cilia_force", self.base_cilia), ("shift_cilia_force", self.shift_cilia)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
                if a.size != 3 * self.n_voxels:
                    raise ValueError("%s needs 3 values per voxel" % name)
                keep.append(a)
                setattr(d, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        for k, toks in enumerate(self.force_field):
            if toks:
                arr = (_abi.Token * len(toks))(*[_abi.Token(op, val) for op, val in toks])
                keep.append(arr)
                d.n_force_field_tokens[k] = len(toks)
                d.force_field_tokens[k] = C.cast(arr, C.POINTER(_abi.Token))
        d.enable_detach = self.enable_detach
        d.secondary_experiment = self.secondary_experiment
        d.reinit_after = self.reinit_after
        d.enable_signals = self.enable_signals
        d.enable_attach = self.enable_attach
        for k, toks in enumerate(self.attach_conditions):
            if toks:
                arr = (_abi.Token * len(toks))(*[_abi.Token(op, val) for op, val in toks])
                keep.append(arr)
                d.n_attach_tokens[k] = len(toks)
                d.attach_tokens[k] = C.cast(arr, C.POINTER(_abi.Token))
        d.safety_guard = self.safety_guard
        if self.init_velocity is not None:
            a = np.ascontiguousarray(self.init_velocity, dtype=np.float64).reshape(-1)
            if a.size != 3 * self.n_voxels:
                raise ValueError("init_velocity needs 3 values per voxel")
            keep.append(a)
            d.init_velocity = a.ctypes.data_as(C.POINTER(C.c_double))
        d.max_dist_close_pairs = self.max_dist_close_pairs
        return d, keep
