"""ctypes mirror of include/vx3h.h (the engine's C ABI).

The structures here must match the C header field for field; a CPU test checks
their sizes against the compiled library's own sizeof() values.
"""
import ctypes as C

MAX_CURVE = 16
MAX_TOKENS = 256

OK, EINVAL, ENOMEM, EDEVICE, EPARSE, EUNSUPPORTED, ENOTFOUND = range(7)
MDL_LINEAR, MDL_LINEAR_FAIL, MDL_BILINEAR, MDL_DATA = range(4)
DOF_TX, DOF_TY, DOF_TZ, DOF_RX, DOF_RY, DOF_RZ = 1, 2, 4, 8, 16, 32
DOF_ALL = 63
PATH_STREAM, PATH_RESIDENT, PATH_AUTO = range(3)
PREC_F32, PREC_F64 = range(2)
(SC_NONE, SC_MAX_TIME_STEPS, SC_MAX_SIM_TIME, SC_TEMP_CYCLES, SC_CONST_MAXENERGY, SC_MIN_KE,
 SC_MIN_MAXMOVE) = range(7)
SIM_RUNNING, SIM_DONE, SIM_DIVERGED, SIM_EMPTY, SIM_FAILED = range(5)


class MaterialDesc(C.Structure):
    _fields_ = [
        ("mat_model", C.c_int),
        ("elastic_mod", C.c_double),
        ("plastic_mod", C.c_double),
        ("yield_stress", C.c_double),
        ("fail_stress", C.c_double),
        ("density", C.c_double),
        ("poissons_ratio", C.c_double),
        ("cte", C.c_double),
        ("u_static", C.c_double),
        ("u_dynamic", C.c_double),
        ("zeta_internal", C.c_double),
        ("zeta_global", C.c_double),
        ("zeta_collision", C.c_double),
        ("n_data", C.c_int),
        ("strain_data", C.c_double * MAX_CURVE),
        ("stress_data", C.c_double * MAX_CURVE),
        ("is_measured", C.c_int),
        ("fixed", C.c_int),
        ("thermal_on_after", C.c_double),
        ("matid", C.c_int),
        ("color", C.c_int * 4),
        ("cilia", C.c_double),
        ("cilia_on_after", C.c_double),
        ("remove_after", C.c_double),
        ("is_pacemaker", C.c_int),
        ("pacemaker_period", C.c_double),
        ("signal_value_decay", C.c_double),
        ("signal_time_delay", C.c_double),
        ("inactive_period", C.c_double),
        ("is_target", C.c_int),
        ("sticky", C.c_int),
    ]


class External(C.Structure):
    _fields_ = [
        ("voxel", C.c_int),
        ("dof_fixed", C.c_int),
        ("translate", C.c_double * 3),
        ("rotate", C.c_double * 3),
        ("force", C.c_double * 3),
        ("moment", C.c_double * 3),
    ]


class Token(C.Structure):
    _fields_ = [("op", C.c_int), ("value", C.c_double)]


class SimDesc(C.Structure):
    _fields_ = [
        ("voxel_size", C.c_double),
        ("n_materials", C.c_int),
        ("materials", C.POINTER(MaterialDesc)),
        ("n_voxels", C.c_int),
        ("ijk", C.POINTER(C.c_int)),
        ("material", C.POINTER(C.c_int)),
        ("phase", C.POINTER(C.c_double)),
        ("n_externals", C.c_int),
        ("externals", C.POINTER(External)),
        ("gravity_mult", C.c_float),
        ("floor_enabled", C.c_int),
        ("temp_enabled", C.c_int),
        ("vary_temp_enabled", C.c_int),
        ("temp_base", C.c_double),
        ("temp_amplitude", C.c_double),
        ("temp_period", C.c_double),
        ("enable_expansion", C.c_int),
        ("dt_frac", C.c_double),
        ("dt_fixed", C.c_double),
        ("volume_effects", C.c_int),
        ("collisions_enabled", C.c_int),
        ("bounding_radius", C.c_double),
        ("watch_distance", C.c_double),
        ("n_stop_tokens", C.c_int),
        ("stop_tokens", C.POINTER(Token)),
        ("n_fitness_tokens", C.c_int),
        ("fitness_tokens", C.POINTER(Token)),
        ("max_steps", C.c_int),
        ("save_position_of_all_voxels", C.c_int),
        ("stop_condition_type", C.c_int),
        ("stop_condition_value", C.c_double),
        ("equilibrium_mode", C.c_int),
        ("enable_cilia", C.c_int),
        ("base_cilia_force", C.POINTER(C.c_double)),
        ("shift_cilia_force", C.POINTER(C.c_double)),
        ("n_force_field_tokens", C.c_int * 3),
        ("force_field_tokens", C.POINTER(Token) * 3),
        ("enable_detach", C.c_int),
        ("secondary_experiment", C.c_int),
        ("reinit_after", C.c_double),
        ("enable_signals", C.c_int),
        ("enable_attach", C.c_int),
        ("n_attach_tokens", C.c_int * 5),
        ("attach_tokens", C.POINTER(Token) * 5),
        ("safety_guard", C.c_int),
        ("init_velocity", C.POINTER(C.c_double)),
        ("max_dist_close_pairs", C.c_double),
    ]


class BatchCfg(C.Structure):
    _fields_ = [
        ("max_device_bytes", C.c_size_t),
        ("precision", C.c_int),
        ("path", C.c_int),
        ("block_size", C.c_int),
        ("steps_per_launch", C.c_int),
        ("collision_rebuild_interval", C.c_int),
        ("use_graph", C.c_int),
        ("collision_table_entries", C.c_int),
    ]


class SimStatus(C.Structure):
    _fields_ = [
        ("slot", C.c_int),
        ("state", C.c_int),
        ("steps", C.c_int),
        ("num_voxel", C.c_int),
        ("num_measured_voxel", C.c_int),
        ("current_time", C.c_double),
        ("dt", C.c_double),
        ("initial_com", C.c_double * 3),
        ("current_com", C.c_double * 3),
        ("fitness", C.c_double),
        ("total_distance_of_all_voxels", C.c_double),
        ("recent_angle", C.c_double),
        ("recommended_dt", C.c_double),
        ("collision_count", C.c_int),
        ("n_links", C.c_int),
        ("num_close_pairs", C.c_int),
        ("target_closeness", C.c_double),
    ]


class StateView(C.Structure):
    _fields_ = [
        ("pos", C.POINTER(C.c_double)),
        ("orient", C.POINTER(C.c_double)),
        ("lin_mom", C.POINTER(C.c_double)),
        ("ang_mom", C.POINTER(C.c_double)),
        ("vox_flags", C.POINTER(C.c_int)),
        ("temp", C.POINTER(C.c_double)),
        ("link_strain", C.POINTER(C.c_double)),
        ("link_max_strain", C.POINTER(C.c_double)),
        ("link_small_angle", C.POINTER(C.c_int)),
        ("init_pos", C.POINTER(C.c_double)),
        ("link_capacity", C.c_int),
    ]


class TopologyView(C.Structure):
    _fields_ = [
        ("n_voxels", C.c_int),
        ("n_links", C.c_int),
        ("link_vneg", C.POINTER(C.c_int)),
        ("link_vpos", C.POINTER(C.c_int)),
        ("link_axis", C.POINTER(C.c_int)),
        ("voxel_links", C.POINTER(C.c_int)),
    ]


class ReportEntry(C.Structure):
    _fields_ = [
        ("vxd_name", C.c_char_p),
        ("status", C.POINTER(SimStatus)),
        ("voxel_size", C.c_double),
        ("save_positions", C.c_int),
        ("init_pos", C.POINTER(C.c_double)),
        ("pos", C.POINTER(C.c_double)),
        ("mats", C.POINTER(C.c_int)),
    ]


class VxaInfo(C.Structure):
    _fields_ = [
        ("enable_attach", C.c_int), ("enable_detach", C.c_int),
        ("enable_cilia", C.c_int), ("enable_signals", C.c_int),
        ("secondary_experiment", C.c_int),
        ("record_step_size", C.c_int),
        ("max_dist_close_pairs", C.c_double),
        ("has_force_field", C.c_int),
        ("record_voxel", C.c_int), ("record_link", C.c_int),
        ("stop_condition_type", C.c_int), ("stop_condition_value", C.c_double),
        ("equilibrium_mode", C.c_int),
    ]


class ServerCfg(C.Structure):
    _fields_ = [("batch", BatchCfg), ("quantum_steps", C.c_int), ("max_sims", C.c_int)]


class ServerResult(C.Structure):
    _fields_ = [("ticket", C.c_longlong), ("status", SimStatus), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("voxel_steps", C.c_double)]


class TuneTrial(C.Structure):
    _fields_ = [("block_size", C.c_int), ("steps_per_launch", C.c_int), ("collision_rebuild_interval", C.c_int),
                ("voxel_steps_per_s", C.c_double)]
