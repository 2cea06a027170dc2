#!/usr/bin/env python3
"""Writes tests/golden/vxa_<case>.npz: what the reference's own VXA loader and CPU
engine (oracle/_ref/libvx_ref.so, built from the reference sources in place) make of
each document in tests/vxa_cases.py -- voxel lattice indices, material indices,
phase offsets, per-voxel externals, environment values, derived material constants,
and the state after 200 VX3 steps at sampled voxels.

    python tests/golden/make_vxa_golden.py        (needs oracle/_ref built: make -C oracle ref)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle_lib import OSim, load  # noqa: E402
from vxa_cases import CASES  # noqa: E402

STEPS = 200


def sample_index(nv, k=96):
    return np.unique(np.linspace(0, nv - 1, min(nv, k)).astype(np.int64))


def reference_vectors(text):
    lib = load("ref")
    r = OSim.from_vxa(text)
    nv = r.nv
    ijk, mat, ph = np.zeros(3 * nv, np.int32), np.zeros(nv, np.int32), np.zeros(nv)
    lib.ref_vxa_voxels(C.c_void_p(r.h), ijk.ctypes.data_as(C.POINTER(C.c_int)), mat.ctypes.data_as(C.POINTER(C.c_int)),
                       ph.ctypes.data_as(C.POINTER(C.c_double)))
    dof, tr, ro = np.zeros(nv, np.int32), np.zeros(3 * nv), np.zeros(3 * nv)
    fo, mo = np.zeros(3 * nv, np.float32), np.zeros(3 * nv, np.float32)
    lib.ref_vxa_externals(C.c_void_p(r.h), dof.ctypes.data_as(C.POINTER(C.c_int)), tr.ctypes.data_as(C.POINTER(C.c_double)),
                          ro.ctypes.data_as(C.POINTER(C.c_double)), fo.ctypes.data_as(C.POINTER(C.c_float)),
                          mo.ctypes.data_as(C.POINTER(C.c_float)))
    g, fl = C.c_float(), C.c_int()
    dtf, ta, tp = C.c_double(), C.c_double(), C.c_double()
    lib.ref_vxa_env(C.c_void_p(r.h), C.byref(g), C.byref(fl), C.byref(dtf), C.byref(ta), C.byref(tp))
    cap = 256
    rgba, mid = np.zeros(4 * cap, np.int32), np.zeros(cap, np.int32)
    nm = lib.ref_vxa_colors(C.c_void_p(r.h), rgba.ctypes.data_as(C.POINTER(C.c_int)), mid.ctypes.data_as(C.POINTER(C.c_int)),
                            cap)
    cb, cs = np.zeros(3 * nv), np.zeros(3 * nv)
    mcil, mon = np.zeros(cap), np.zeros(cap)
    dp = C.POINTER(C.c_double)
    lib.ref_vxa_cilia(C.c_void_p(r.h), cb.ctypes.data_as(dp), cs.ctypes.data_as(dp), mcil.ctypes.data_as(dp),
                      mon.ctypes.data_as(dp))
    st_type, st_val, eq = C.c_int(), C.c_double(), C.c_int()
    lib.ref_vxa_stop(C.c_void_p(r.h), C.byref(st_type), C.byref(st_val), C.byref(eq))
    vmc = r.voxel_material_consts()
    lidx, lmc = r.link_material_consts()
    r.step(STEPS)
    st = r.state()
    s = sample_index(nv)
    return dict(nv=np.int64(nv), nl=np.int64(r.nl), ijk=ijk, mat=mat, phase=ph, ext_dof=dof, ext_trans=tr, ext_rot=ro,
                ext_force=fo, ext_moment=mo, env_grav=np.float32(g.value), env_floor=np.int32(fl.value),
                env_dtfrac=np.float64(dtf.value), env_temp_amp=np.float64(ta.value), env_temp_period=np.float64(tp.value),
                mat_color=rgba[: 4 * nm].reshape(-1, 4), mat_id=mid[:nm],
                stop=np.array([st_type.value, st_val.value, eq.value], np.float64),
                cilia_base=cb, cilia_shift=cs, mat_cilia=np.stack([mcil[:nm], mon[:nm]], axis=1), vmat_consts=vmc, lmat_index=lidx, lmat_consts=lmc, sample=s, pos=st["pos"][s], orient=st["orient"][s],
                pos_sum=np.float64(st["pos"].sum()), orient_sum=np.float64(st["orient"].sum()))


def main():
    for name, make in CASES.items():
        vec = reference_vectors(make())
        path = os.path.join(HERE, "vxa_%s.npz" % name)
        np.savez_compressed(path, **vec)
        print("wrote", path, "voxels", int(vec["nv"]))


if __name__ == "__main__":
    main()
