"""ctypes bindings of libvx3h.so (include/vx3h.h) and a small Batch wrapper.

The library is the product: there is no Python or CPU fallback.  Loading fails
loudly when gpuvoxels_amd/libvx3h.so has not been built.
"""
import ctypes as C
import os

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvx3h.so")

_lib = None


class VX3Error(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("vx3h status %d: %s" % (status, msg))
        self.status = status


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("gpuvoxels_amd/libvx3h.so is not built; run `python -m gpuvoxels_amd.build`")
    L = C.CDLL(LIB_PATH)
    S = C.c_int
    P = C.c_void_p
    sig = {
        "vx3h_last_error": (C.c_char_p, []),
        "vx3h_device_count": (S, [C.POINTER(C.c_int)]),
        "vx3h_batch_create": (S, [C.c_int, C.POINTER(_abi.BatchCfg), C.POINTER(P)]),
        "vx3h_batch_destroy": (S, [P]),
        "vx3h_sim_admit": (S, [P, C.POINTER(_abi.SimDesc), C.POINTER(C.c_int)]),
        "vx3h_sim_footprint": (S, [C.POINTER(_abi.SimDesc), C.c_int, C.POINTER(C.c_size_t)]),
        "vx3h_batch_step": (S, [P, C.c_int]),
        "vx3h_sim_step": (S, [P, C.c_int, C.c_float, C.POINTER(C.c_int)]),
        "vx3h_batch_step_noalloc": (S, [P, C.c_int]),
        "vx3h_batch_sync": (S, [P]),
        "vx3h_batch_poll": (S, [P, C.POINTER(_abi.SimStatus), C.c_int, C.POINTER(C.c_int)]),
        "vx3h_sim_retire": (S, [P, C.c_int]),
        "vx3h_sim_read_state": (S, [P, C.c_int, C.POINTER(_abi.StateView)]),
        "vx3h_build_topology": (S, [C.POINTER(_abi.SimDesc), C.POINTER(_abi.TopologyView)]),
        "vx3h_batch_kernel_times": (S, [P, C.c_int, C.POINTER(C.c_double)]),
        "vx3h_batch_set_timing": (S, [P, C.c_int]),
        "vx3h_sim_workgroups": (S, [P, C.c_int, C.POINTER(C.c_int)]),
        "vx3h_calib_copy": (S, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
        "vx3h_sim_collision_pairs": (S, [P, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
        "vx3h_server_create": (S, [C.c_int, C.POINTER(_abi.ServerCfg), C.POINTER(P)]),
        "vx3h_server_destroy": (S, [P]),
        "vx3h_server_submit": (S, [P, C.POINTER(_abi.SimDesc), C.c_int, C.POINTER(C.c_longlong)]),
        "vx3h_server_round": (S, [P, C.POINTER(C.c_int)]),
        "vx3h_server_results": (S, [P, C.POINTER(_abi.ServerResult), C.c_int, C.POINTER(C.c_int)]),
        "vx3h_server_clock": (C.c_double, [P]),
        "vx3h_batch_timing": (S, [P, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
        "vx3h_sim_history_header": (S, [P, C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "vx3h_sim_history_record": (S, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_size_t,
                                        C.POINTER(C.c_size_t)]),
        "vx3h_autotune": (S, [C.c_int, C.POINTER(C.POINTER(_abi.SimDesc)), C.c_int, C.POINTER(_abi.BatchCfg), C.c_int,
                              C.c_int, C.c_int, C.POINTER(_abi.BatchCfg), C.POINTER(_abi.TuneTrial), C.c_int,
                              C.POINTER(C.c_int)]),
    }
    sig.update({
        "vx3h_vxa_parse": (S, [C.c_char_p, C.c_char_p, C.POINTER(P)]),
        "vx3h_vxa_desc": (C.POINTER(_abi.SimDesc), [P]),
        "vx3h_vxa_get_info": (S, [P, C.POINTER(_abi.VxaInfo), C.POINTER(C.c_char_p)]),
        "vx3h_vxa_free": (None, [P]),
        "vx3h_report_xml": (S, [C.c_char_p, C.POINTER(_abi.ReportEntry), C.c_int, C.c_char_p, C.c_size_t,
                                C.POINTER(C.c_size_t)]),
    })
    optional = {}
    for name, (res, args) in list(sig.items()) + list(optional.items()):
        if not hasattr(L, name):
            if name in sig:
                raise RuntimeError("libvx3h.so lacks %s" % name)
            continue
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status):
    if status != _abi.OK:
        raise VX3Error(status, lib().vx3h_last_error().decode())


def device_count():
    n = C.c_int(0)
    check(lib().vx3h_device_count(C.byref(n)))
    return n.value


class Vxa:
    """A VXA (optionally patched by a VXD) read by the engine's loader (vx3h_vxa_parse).
    `desc` points into memory the object owns."""

    def __init__(self, vxa_text, vxd_text=None):
        self.h = C.c_void_p()
        check(lib().vx3h_vxa_parse(vxa_text.encode() if isinstance(vxa_text, str) else vxa_text,
                                   None if vxd_text is None else
                                   (vxd_text.encode() if isinstance(vxd_text, str) else vxd_text), C.byref(self.h)))
        self.desc = lib().vx3h_vxa_desc(self.h).contents
        self.desc._owner = self   # the description points into memory this object frees
        info = _abi.VxaInfo()
        raw = C.c_char_p()
        check(lib().vx3h_vxa_get_info(self.h, C.byref(info), C.byref(raw)))
        self.info = info
        self.raw_print = raw.value.decode() if raw.value else ""

    @property
    def n_voxels(self):
        return self.desc.n_voxels

    def close(self):
        if self.h:
            lib().vx3h_vxa_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def report_xml(input_dir_name, entries):
    """entries: list of (vxd_name, SimStatus, voxel_size, positions or None) where
    positions = (init_pos (V,3), pos (V,3), mats (V,)) when SavePositionOfAllVoxels."""
    arr = (_abi.ReportEntry * max(len(entries), 1))()
    keep = []
    for i, (name, st, vs, posi) in enumerate(entries):
        nm = name.encode()
        keep.append(nm)
        arr[i].vxd_name = nm
        s = _abi.SimStatus()
        C.memmove(C.byref(s), C.byref(st), C.sizeof(s))
        keep.append(s)
        arr[i].status = C.pointer(s)
        arr[i].voxel_size = vs
        if posi is not None:
            a, b, m = (np.ascontiguousarray(posi[0], np.float64), np.ascontiguousarray(posi[1], np.float64),
                       np.ascontiguousarray(posi[2], np.int32))
            keep += [a, b, m]
            arr[i].save_positions = 1
            arr[i].init_pos, arr[i].pos, arr[i].mats = _p(a, C.c_double), _p(b, C.c_double), _p(m, C.c_int)
    n = C.c_size_t(0)
    check(lib().vx3h_report_xml(input_dir_name.encode(), arr, len(entries), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value)
    check(lib().vx3h_report_xml(input_dir_name.encode(), arr, len(entries), buf, n.value, C.byref(n)))
    return buf.value.decode()


def calib_copy(nbytes, reps=1, device=0):
    """Mean ms of a plain device copy of nbytes (reads nbytes, writes nbytes)."""
    ms = C.c_double(0)
    check(lib().vx3h_calib_copy(device, int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def topology(model):
    d, keep = model.desc()
    tv = _abi.TopologyView()
    check(lib().vx3h_build_topology(C.byref(d), C.byref(tv)))
    nv, nl = tv.n_voxels, tv.n_links
    vneg, vpos, axis = np.zeros(nl, np.int32), np.zeros(nl, np.int32), np.zeros(nl, np.int32)
    vl = np.zeros(6 * nv, np.int32)
    tv.link_vneg, tv.link_vpos, tv.link_axis = _p(vneg, C.c_int), _p(vpos, C.c_int), _p(axis, C.c_int)
    tv.voxel_links = _p(vl, C.c_int)
    check(lib().vx3h_build_topology(C.byref(d), C.byref(tv)))
    return vneg, vpos, axis, vl.reshape(-1, 6)


def autotune(models, precision=_abi.PREC_F64, path=_abi.PATH_AUTO, warmup_steps=200, trial_steps=1000,
             tune_rebuild_interval=False, device=0, use_graph=1):
    """vx3h_autotune over Models or Vxas: returns (best configuration as a dict of
    Batch keyword arguments, [(block_size, steps_per_launch, rebuild_interval,
    voxel-steps/s), ...] for every candidate measured)."""
    keep, ptrs = [], []
    for m in models:
        if isinstance(m, Vxa):
            ptrs.append(C.pointer(m.desc))
        else:
            d, k = m.desc()
            keep.append((d, k))
            ptrs.append(C.pointer(d))
    arr = (C.POINTER(_abi.SimDesc) * len(ptrs))(*ptrs)
    base = _abi.BatchCfg()
    base.precision, base.path, base.use_graph = precision, path, use_graph
    best = _abi.BatchCfg()
    cap = 64
    trials = (_abi.TuneTrial * cap)()
    n = C.c_int(0)
    check(lib().vx3h_autotune(device, arr, len(ptrs), C.byref(base), int(warmup_steps), int(trial_steps),
                              int(bool(tune_rebuild_interval)), C.byref(best), trials, cap, C.byref(n)))
    cfg = dict(precision=best.precision, path=best.path, block_size=best.block_size,
               steps_per_launch=best.steps_per_launch, collision_rebuild_interval=best.collision_rebuild_interval,
               use_graph=best.use_graph)
    return cfg, [(t.block_size, t.steps_per_launch, t.collision_rebuild_interval, t.voxel_steps_per_s)
                 for t in trials[: min(n.value, cap)]]


class Batch:
    """A device-resident batch of simulations (vx3h_batch)."""

    def __init__(self, device=0, precision=_abi.PREC_F32, path=_abi.PATH_STREAM, max_device_bytes=0, block_size=0,
                 steps_per_launch=0, use_graph=1, collision_rebuild_interval=0, collision_table_entries=0):
        cfg = _abi.BatchCfg()
        cfg.max_device_bytes = max_device_bytes
        cfg.precision = precision
        cfg.path = path
        cfg.block_size = block_size
        cfg.steps_per_launch = steps_per_launch
        cfg.collision_rebuild_interval = collision_rebuild_interval
        cfg.collision_table_entries = collision_table_entries
        cfg.use_graph = use_graph
        self.h = C.c_void_p()
        check(lib().vx3h_batch_create(device, C.byref(cfg), C.byref(self.h)))
        self.nv = {}
        self.nl = {}

    def close(self):
        if self.h:
            lib().vx3h_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, model):
        """Admits a Model or a Vxa; returns the slot."""
        if isinstance(model, Vxa):
            d = model.desc
        else:
            d, keep = model.desc()
        slot = C.c_int(-1)
        check(lib().vx3h_sim_admit(self.h, C.byref(d), C.byref(slot)))
        self.nv[slot.value] = model.n_voxels
        return slot.value

    def step(self, n):
        check(lib().vx3h_batch_step(self.h, int(n)))

    def step_noalloc(self, n):
        """vx3h_batch_step_noalloc: n steps of every simulation on the per-step kernels with
        no device allocation (the step's scratch is a static device array)."""
        check(lib().vx3h_batch_step_noalloc(self.h, int(n)))

    def do_time_step(self, slot, dt=-1.0):
        """VX3_VoxelyzeKernel::doTimeStep(dt) of one simulation (vx3h_sim_step): False when
        a link diverged."""
        ok = C.c_int(0)
        check(lib().vx3h_sim_step(self.h, int(slot), float(dt), C.byref(ok)))
        return bool(ok.value)

    def sync(self):
        check(lib().vx3h_batch_sync(self.h))

    def poll(self):
        n = C.c_int(0)
        check(lib().vx3h_batch_poll(self.h, None, 0, C.byref(n)))
        arr = (_abi.SimStatus * max(n.value, 1))()
        check(lib().vx3h_batch_poll(self.h, arr, n.value, C.byref(n)))
        return {arr[i].slot: arr[i] for i in range(n.value)}

    PASSES = ("poisson", "links", "contacts", "voxels", "control", "resident")

    def kernel_times(self, n_steps):
        """Mean per-pass kernel durations (ms) over n_steps steps, from HIP events
        recorded on the batch's stream."""
        out = (C.c_double * len(self.PASSES))()
        check(lib().vx3h_batch_kernel_times(self.h, int(n_steps), out))
        return {k: out[i] for i, k in enumerate(self.PASSES) if out[i] > 0}

    def workgroups(self, slot):
        """Workgroups that step the simulation in `slot`: 0 = per-step kernels, 1 = one
        workgroup of the resident kernel, n > 1 = n tiles of the tiled resident kernel."""
        n = C.c_int(0)
        check(lib().vx3h_sim_workgroups(self.h, int(slot), C.byref(n)))
        return n.value

    def set_timing(self, enable=True):
        check(lib().vx3h_batch_set_timing(self.h, int(bool(enable))))

    def timing(self):
        """(summed ms, launches, steps) of the resident launches since the last call."""
        ms, n, st = C.c_double(0), C.c_int(0), C.c_longlong(0)
        check(lib().vx3h_batch_timing(self.h, C.byref(ms), C.byref(n), C.byref(st)))
        return ms.value, n.value, st.value

    def collision_pairs(self, slot):
        """(n, 2) sim-local voxel pairs (i < j) the simulation currently watches, ascending."""
        n = C.c_int(0)
        check(lib().vx3h_sim_collision_pairs(self.h, slot, None, 0, C.byref(n)))
        out = np.zeros(2 * max(n.value, 1), np.int32)
        check(lib().vx3h_sim_collision_pairs(self.h, slot, _p(out, C.c_int), n.value, C.byref(n)))
        return out[: 2 * n.value].reshape(-1, 2)

    def retire(self, slot):
        check(lib().vx3h_sim_retire(self.h, slot))
        self.nv.pop(slot, None)

    def _text(self, fn, *args):
        n = C.c_size_t(0)
        check(fn(*args, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(fn(*args, buf, n.value, C.byref(n)))
        return buf.value.decode()

    def history_header(self, slot):
        """The history stream's settings block (rescale, material colours, voxel size)."""
        return self._text(lib().vx3h_sim_history_header, self.h, slot)

    def history_record(self, slot, step_label, record_voxel=1, record_link=0):
        """One history line of the simulation's current state."""
        return self._text(lib().vx3h_sim_history_record, self.h, slot, step_label, record_voxel, record_link)

    def read_state(self, slot, nl=None):
        nv = self.nv[slot]
        pos, ori, lm, am = np.zeros(3 * nv), np.zeros(4 * nv), np.zeros(3 * nv), np.zeros(3 * nv)
        fl, tf, ip = np.zeros(nv, np.int32), np.zeros(nv), np.zeros(3 * nv)
        v = _abi.StateView()
        v.init_pos = _p(ip, C.c_double)
        v.pos, v.orient, v.lin_mom, v.ang_mom = (_p(pos, C.c_double), _p(ori, C.c_double), _p(lm, C.c_double),
                                                 _p(am, C.c_double))
        v.vox_flags, v.temp = _p(fl, C.c_int), _p(tf, C.c_double)
        out = {}
        if nl:
            st, ms, sa = np.zeros(nl), np.zeros(nl), np.zeros(nl, np.int32)
            v.link_strain, v.link_max_strain, v.link_small_angle = (_p(st, C.c_double), _p(ms, C.c_double),
                                                                    _p(sa, C.c_int))
            v.link_capacity = nl   # links made by attach follow the admitted ones
            out.update(strain=st, max_strain=ms, small_angle=sa)
        check(lib().vx3h_sim_read_state(self.h, slot, C.byref(v)))
        out.update(pos=pos.reshape(-1, 3), orient=ori.reshape(-1, 4), lin_mom=lm.reshape(-1, 3),
                   ang_mom=am.reshape(-1, 3), flags=fl, tfac=tf, init_pos=ip.reshape(-1, 3))
        return out


class Server:
    """Continuous-batching request loop (vx3h_server_*): submit models as they arrive,
    run rounds, collect results with submit/admit/finish times."""

    def __init__(self, device=0, precision=_abi.PREC_F64, path=_abi.PATH_AUTO, max_device_bytes=0, quantum_steps=1000,
                 max_sims=0, steps_per_launch=0):
        cfg = _abi.ServerCfg()
        cfg.batch.precision = precision
        cfg.batch.path = path
        cfg.batch.max_device_bytes = max_device_bytes
        cfg.batch.steps_per_launch = steps_per_launch
        cfg.batch.use_graph = 1
        cfg.quantum_steps = quantum_steps
        cfg.max_sims = max_sims
        self.h = C.c_void_p()
        check(lib().vx3h_server_create(device, C.byref(cfg), C.byref(self.h)))

    def submit(self, model, max_steps=0):
        """Queues a Model or Vxa (the server copies it); returns the ticket."""
        if isinstance(model, Vxa):
            d = model.desc
        else:
            d, keep = model.desc()
        t = C.c_longlong(0)
        check(lib().vx3h_server_submit(self.h, C.byref(d), int(max_steps), C.byref(t)))
        return t.value

    def round(self):
        n = C.c_int(0)
        check(lib().vx3h_server_round(self.h, C.byref(n)))
        return n.value

    def results(self):
        n = C.c_int(0)
        check(lib().vx3h_server_results(self.h, None, 0, C.byref(n)))
        if not n.value:
            return []
        arr = (_abi.ServerResult * n.value)()
        check(lib().vx3h_server_results(self.h, arr, n.value, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    def clock(self):
        return lib().vx3h_server_clock(self.h)

    def close(self):
        if self.h:
            lib().vx3h_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
