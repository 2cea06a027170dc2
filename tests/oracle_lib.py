"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the reference's
own CPU engine (oracle/_ref/libvx_ref.so, present only where the reference tree
was available to build it).  Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

from gpuvoxels_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libvx_ref.so")

_libs = {}


def _bind(lib):
    P, D, F, I = C.c_void_p, C.c_double, C.c_float, C.c_int
    pI, pD, pF = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
    sig = {
        "vxo_create": (P, [C.POINTER(_abi.SimDesc)]),
        "vxo_free": (None, [P]),
        "vxo_last_error": (C.c_char_p, []),
        "vxo_counts": (None, [P, pI, pI]),
        "vxo_topology": (None, [P, pI, pI, pI, pI]),
        "vxo_do_time_step": (I, [P, F]),
        "vxo_recommended_dt": (F, [P]),
        "vxo_reset_time": (None, [P]),
        "vxo_set_gravity": (None, [P, F]),
        "vxo_enable_floor": (None, [P, I]),
        "vxo_enable_collisions": (None, [P, I]),
        "vxo_set_ambient_temperature": (None, [P, F, I]),
        "vxo_set_voxel_temperature": (None, [P, I, F]),
        "vxo_ext_set_force": (None, [P, I, F, F, F]),
        "vxo_ext_set_moment": (None, [P, I, F, F, F]),
        "vxo_ext_set_fixed": (None, [P, I, I]),
        "vxo_ext_set_fixed_all": (None, [P, I, I]),
        "vxo_ext_set_displacement_all": (None, [P, I, D, D, D, D, D, D]),
        "vxo_external_force": (None, [P, I, pF]),
        "vxo_material_set": (None, [P, I, I, F]),
        "vxo_replace_voxel_material": (None, [P, I, I]),
        "vxo_vx3_step": (I, [P, I]),
        "vxo_vx3_status": (None, [P, C.POINTER(_abi.SimStatus)]),
        "vxo_voxel_state": (None, [P, pD, pD, pD, pD, pI, pF]),
        "vxo_link_state": (None, [P, pF, pF, pF, pI, pI, pD, pD]),
        "vxo_corner_offset": (None, [P, I, I, pF]),
        "vxo_voxcad_history": (None, [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "vxo_local_signal": (None, [P, C.POINTER(C.c_double)]),
        "vxo_collision_pairs": (I, [P, pI, I]),
        "vxo_set_collision_interval": (None, [P, I]),
        "vxo_voxel_material_consts": (I, [P, pF, I]),
        "vxo_link_material_consts": (I, [P, pI, pF, I]),
        "vxo_mathtree_eval": (D, [C.POINTER(_abi.Token), I, D, D, D, D, D, D, D, I, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if hasattr(lib, "ref_load_vxa"):
        lib.ref_load_vxa.restype = P
        lib.ref_load_vxa.argtypes = [C.c_char_p]
        lib.ref_vxa_env.argtypes = [P, pF, pI, pD, pD, pD]
        lib.ref_vxa_voxels.argtypes = [P, pI, pI, pD]
        lib.ref_vxa_externals.argtypes = [P, pI, pD, pD, pF, pF]
    return lib


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


def load(kind="oracle"):
    if kind in _libs:
        return _libs[kind]
    path = ORACLE_SO if kind == "oracle" else REF_SO
    if kind == "oracle" and not os.path.exists(path):
        build_oracle()
    if not os.path.exists(path):
        return None
    _libs[kind] = _bind(C.CDLL(path))
    return _libs[kind]


def ref_available():
    return os.path.exists(REF_SO)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class OSim:
    """One simulation in the oracle (kind='oracle') or the reference (kind='ref')."""

    def __init__(self, model, kind="oracle", handle=None):
        self.lib = load(kind)
        if self.lib is None:
            raise RuntimeError("library for %s not built" % kind)
        self.kind = kind
        if handle is None:
            self.desc, self._keep = model.desc()
            self.h = self.lib.vxo_create(C.byref(self.desc))
            if not self.h:
                raise ValueError(self.lib.vxo_last_error().decode())
        else:
            self.h = handle
        nv, nl = C.c_int(), C.c_int()
        self.lib.vxo_counts(self.h, C.byref(nv), C.byref(nl))
        self.nv, self.nl = nv.value, nl.value

    @classmethod
    def from_desc(cls, desc, kind="oracle"):
        """A simulation built from a vx3h_sim_desc (kept alive by the caller)."""
        lib = load(kind)
        h = lib.vxo_create(C.byref(desc))
        if not h:
            raise ValueError(lib.vxo_last_error().decode())
        return cls(None, kind, handle=h)

    @classmethod
    def from_vxa(cls, text):
        lib = load("ref")
        h = lib.ref_load_vxa(text.encode())
        return cls(None, "ref", handle=h)

    def __del__(self):
        try:
            self.lib.vxo_free(self.h)
        except Exception:
            pass

    def topology(self):
        vneg = np.zeros(self.nl, np.int32)
        vpos = np.zeros(self.nl, np.int32)
        axis = np.zeros(self.nl, np.int32)
        vl = np.zeros(6 * self.nv, np.int32)
        self.lib.vxo_topology(self.h, _p(vneg, C.c_int), _p(vpos, C.c_int), _p(axis, C.c_int), _p(vl, C.c_int))
        return vneg, vpos, axis, vl.reshape(-1, 6)

    def step_raw(self, dt):
        return self.lib.vxo_do_time_step(self.h, dt)

    def recommended_dt(self):
        return self.lib.vxo_recommended_dt(self.h)

    def step(self, n):
        r = self.lib.vxo_vx3_step(self.h, n)
        nl = C.c_int()
        self.lib.vxo_counts(self.h, None, C.byref(nl))   # attach adds links
        self.nl = nl.value
        return r

    def status(self):
        st = _abi.SimStatus()
        self.lib.vxo_vx3_status(self.h, C.byref(st))
        return st

    def state(self):
        pos = np.zeros(3 * self.nv)
        ori = np.zeros(4 * self.nv)
        lm = np.zeros(3 * self.nv)
        am = np.zeros(3 * self.nv)
        fl = np.zeros(self.nv, np.int32)
        tp = np.zeros(self.nv, np.float32)
        self.lib.vxo_voxel_state(self.h, _p(pos, C.c_double), _p(ori, C.c_double), _p(lm, C.c_double), _p(am, C.c_double),
                                 _p(fl, C.c_int), _p(tp, C.c_float))
        return dict(pos=pos.reshape(-1, 3), orient=ori.reshape(-1, 4), lin_mom=lm.reshape(-1, 3),
                    ang_mom=am.reshape(-1, 3), flags=fl, temp=tp)

    def link_state(self):
        n = self.nl
        st, ms, so = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        sa, bs = np.zeros(n, np.int32), np.zeros(n, np.int32)
        rl, f = np.zeros(n), np.zeros(12 * n)
        self.lib.vxo_link_state(self.h, _p(st, C.c_float), _p(ms, C.c_float), _p(so, C.c_float), _p(sa, C.c_int),
                                _p(bs, C.c_int), _p(rl, C.c_double), _p(f, C.c_double))
        return dict(strain=st, max_strain=ms, strain_offset=so, small_angle=sa, bool_states=bs, rest_length=rl,
                    forces=f.reshape(-1, 12))

    def collision_pairs(self):
        n = self.lib.vxo_collision_pairs(self.h, None, 0)
        out = np.zeros(2 * max(n, 1), np.int32)
        self.lib.vxo_collision_pairs(self.h, _p(out, C.c_int), n)
        return out[: 2 * n].reshape(-1, 2)

    def set_collision_interval(self, n):
        """Also regenerate the collision lists every n steps (the engine's collision_rebuild_interval)."""
        self.lib.vxo_set_collision_interval(self.h, int(n))

    def voxel_material_consts(self):
        n = self.lib.vxo_voxel_material_consts(self.h, None, 0)
        out = np.zeros(14 * n, np.float32)
        self.lib.vxo_voxel_material_consts(self.h, _p(out, C.c_float), n)
        return out.reshape(-1, 14)

    def link_material_consts(self):
        n = self.lib.vxo_link_material_consts(self.h, None, None, 0)
        idx = np.zeros(self.nl, np.int32)
        out = np.zeros(20 * max(n, 1), np.float32)
        self.lib.vxo_link_material_consts(self.h, _p(idx, C.c_int), _p(out, C.c_float), n)
        return idx, out[: 20 * n].reshape(-1, 20)

    # KAT helpers (CVoxelyze-level mutation)
    def set_force(self, v, f):
        self.lib.vxo_ext_set_force(self.h, v, *[float(x) for x in f])

    def set_moment(self, v, m):
        self.lib.vxo_ext_set_moment(self.h, v, *[float(x) for x in m])

    def set_fixed(self, v, mask):
        self.lib.vxo_ext_set_fixed(self.h, v, mask)

    def set_fixed_all(self, v, fixed=True):
        self.lib.vxo_ext_set_fixed_all(self.h, v, 1 if fixed else 0)

    def set_displacement_all(self, v, t=(0, 0, 0), r=(0, 0, 0)):
        self.lib.vxo_ext_set_displacement_all(self.h, v, *[float(x) for x in t], *[float(x) for x in r])

    def external_force(self, v):
        out = np.zeros(3, np.float32)
        self.lib.vxo_external_force(self.h, v, _p(out, C.c_float))
        return out

    def corner_offset(self, v, corner):
        out = np.zeros(3, np.float32)
        self.lib.vxo_corner_offset(self.h, v, corner, _p(out, C.c_float))
        return out

    def local_signal(self):
        out = np.zeros(self.nv)
        self.lib.vxo_local_signal(self.h, _p(out, C.c_double))
        return out

    def voxcad_history(self):
        """VoxCAD stop-condition histories (kinetic, total energy, max move), newest first."""
        k, t, m = np.zeros(500), np.zeros(500), np.zeros(500)
        self.lib.vxo_voxcad_history(self.h, _p(k, C.c_double), _p(t, C.c_double), _p(m, C.c_double))
        return k, t, m

    def reset_time(self):
        self.lib.vxo_reset_time(self.h)

    def material_set(self, m, which, value):
        self.lib.vxo_material_set(self.h, m, which, value)

    def set_gravity(self, g):
        self.lib.vxo_set_gravity(self.h, g)

    def set_ambient_temperature(self, t, all_voxels=True):
        self.lib.vxo_set_ambient_temperature(self.h, t, 1 if all_voxels else 0)

    def replace_voxel_material(self, v, m):
        self.lib.vxo_replace_voxel_material(self.h, v, m)

    def enable_collisions(self, en=True):
        self.lib.vxo_enable_collisions(self.h, 1 if en else 0)


def mathtree_eval(toks, x=0.0, y=0.0, z=0.0, hit=0.0, t=0.0, angle=0.0, closeness=0.0, ncp=0, nv=0, kind="oracle"):
    lib = load(kind)
    arr = (_abi.Token * len(toks))(*[_abi.Token(op, val) for op, val in toks])
    return lib.vxo_mathtree_eval(arr, len(toks), x, y, z, hit, t, angle, closeness, ncp, nv)
