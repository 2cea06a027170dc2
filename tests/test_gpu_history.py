"""History stream (the VoxCAD playback text the reference prints for a simulation
with RecordHistory.RecordStepSize > 0, VX3_SimulationManager.cu:40-108) against the
same text formatted here from the oracle's state after the same number of steps.

The formatter below follows the reference's printf sequence; the values it prints
are pinned elsewhere (positions and orientations by the trajectory tests, corner
offsets by the oracle-vs-reference and closed-form tests).  GPU and oracle states
differ by rounding, so a printed number may differ from the oracle's by one unit of
its last printed digit at a rounding boundary; the text around the numbers must be
identical."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from gpuvoxels_amd import _abi, engine, workloads
from oracle_lib import OSim
from worker_cases import make_task

pytestmark = pytest.mark.gpu
VS = 1 / 0.001
WORKER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpuvoxels_amd", "vx3_node_worker")
NUM = re.compile(r"-?\d+(?:\.\d+)?")


def oracle_record(o, label, matids, record_voxel=1, record_link=0):
    st = o.state()
    t = o.status().current_time
    vneg, vpos, _, vl = o.topology()
    out = []
    if record_voxel:
        out.append("<<<Step%d Time:%f>>>" % (label, t))
        for v in range(o.nv):
            if (vl[v] >= 0).all():      # interior voxels are not drawn
                continue
            p, q = st["pos"][v], st["orient"][v]
            out.append("%.1f,%.1f,%.1f," % (p[0] * VS, p[1] * VS, p[2] * VS))
            out.append("%.1f,%.2f,%.2f,%.2f," % (2.0 * math.acos(min(q[0], 1.0)) * 57.29577951308232, q[1], q[2], q[3]))
            c = [float(x) * VS for x in (*o.corner_offset(v, 0), *o.corner_offset(v, 7))]
            out.append("%.1f,%.1f,%.1f,%.1f,%.1f,%.1f," % tuple(c))
            out.append("%d," % matids[v])
            out.append("%.1f," % 0.0)
            out.append(";")
        out.append("<<<>>>")
    if record_link:
        out.append("|[[[%d]]]" % label)
        for k in range(o.nl):
            a, b = st["pos"][vpos[k]], st["pos"][vneg[k]]
            out.append("%.4f,%.4f,%.4f,%.4f,%.4f,%.4f,;" % (a[0], a[1], a[2], b[0], b[1], b[2]))
        out.append("[[[]]]")
    out.append("\n")
    return "".join(out)


def expected_header(ids_colors, voxel_size):
    out = "\n{{{setting}}}<rescale>0.001</rescale>\n"
    for mid, (r, g, b, a) in ids_colors:
        out += ("{{{setting}}}<matcolor><id>%d</id><r>%.2f</r><g>%.2f</g><b>%.2f</b><a>%.2f</a></matcolor>\n"
                % (mid, r / 255., g / 255., b / 255., a / 255.))
    return out + "\n{{{setting}}}<voxel_size>%f</voxel_size>\n" % voxel_size


def _split_angles(text):
    """The text with each voxel entry's rotation angle (its fourth field) taken out,
    and the angles."""
    if not text.startswith("<<<Step"):
        return text, []
    head, rest = text.split(">>>", 1)
    body, tail = rest.split("<<<>>>", 1)
    entries, angles = [], []
    for e in body.split(";")[:-1]:
        f = e.split(",")
        angles.append(float(f[3]))
        f[3] = "A"
        entries.append(",".join(f))
    return head + ">>>" + ";".join(entries) + ";<<<>>>" + tail, angles


# FP32 trajectories against the FP64 oracle, by the number of printed decimals: the
# measured FP32 bounds after 1000 steps of the robot workload (3e-3 of the 0.12 m span
# in position, 2.5e-2 in a quaternion component, both with 2x headroom), in the print
# units: positions and corner offsets in mm with one decimal (0.36 mm), quaternion
# components with two decimals, link end positions in m with four decimals.  The
# floor's kinetic friction is discontinuous at zero horizontal velocity, so FP32
# rounding alone moves a voxel by tenths of a millimetre.
F32_TOL = {1: 0.4, 2: 0.03, 4: 4e-4}


def differing(a, b, angle_tol=None, tol=None):
    """Numbers of the two texts that differ by more than one unit of the last printed
    digit (or, with tol, by more than tol[number of decimals] where given); the text
    around the numbers must match exactly.  With angle_tol, the rotation angles
    (degrees) are compared to that tolerance instead: 2*acos(w) is ill-conditioned near
    w = 1, where a float orientation's rounding moves it by tenths of a degree."""
    if angle_tol is not None:
        a, aa = _split_angles(a)
        b, ab = _split_angles(b)
        assert len(aa) == len(ab)
        worst = max([abs(x - y) for x, y in zip(aa, ab)], default=0.0)
        assert worst <= angle_tol, worst
    assert NUM.sub("#", a) == NUM.sub("#", b)
    na, nb = NUM.findall(a), NUM.findall(b)
    bad, off = [], 0
    for x, y in zip(na, nb):
        if x == y:
            continue
        off += 1
        dec = len(x.split(".")[1]) if "." in x else 0
        lim = max(1.0001 * 10.0 ** -dec, (tol or {}).get(dec, 0.0))
        if abs(float(x) - float(y)) > lim:
            bad.append((x, y))
    return bad, off, len(na)


def colored_robot(steps=0):
    m = workloads.robots(1, seed=13, grid=8, steps=steps)[0]
    for mt, c in zip(m.materials, [(1.0, 0.2, 0.35, 0.8), (0.0, 0.6, 0.9, 1.0), (0.5, 0.5, 0.5, 1.0), (0.3, 0.3, 0.1, 0.5)]):
        mt.color = c
    return m


def ids_colors(m):
    d, keep = m.desc()
    return [(d.materials[i].matid, tuple(d.materials[i].color)) for i in range(d.n_materials)]


def test_header_lists_palette_colours():
    m = colored_robot()
    b = engine.Batch(device=0, precision=_abi.PREC_F64, path=_abi.PATH_STREAM)
    s = b.admit(m)
    got = b.history_header(s)
    b.close()
    assert got == expected_header(ids_colors(m), m.voxel_size)
    assert "<r>1.00</r><g>0.20</g><b>0.35</b><a>0.80</a>" in got   # int(c*255)/255 printed with two decimals


@pytest.mark.parametrize("precision,path", [(_abi.PREC_F64, _abi.PATH_STREAM), (_abi.PREC_F64, _abi.PATH_RESIDENT),
                                            (_abi.PREC_F32, _abi.PATH_RESIDENT)])
def test_records_match_oracle(precision, path):
    m = colored_robot()
    matids = np.array([mt.matid for mt in m.materials])[m.mat]
    o = OSim(m)
    b = engine.Batch(device=0, precision=precision, path=path)
    s = b.admit(m)
    done = 0
    for target in (1, 200, 1000):
        b.step(target - done)
        o.step(target - done)
        done = target
        got = b.history_record(s, target - 1, 1, 1)
        want = oracle_record(o, target - 1, matids, 1, 1)
        f64 = precision == _abi.PREC_F64
        bad, off, total = differing(got, want, None if f64 else 2.0, None if f64 else F32_TOL)
        print("step %d: %d numbers, %d differ by one last digit" % (target, total, off))
        assert not bad, bad[:10]
        # FP32 states drift from the FP64 oracle by about 1e-5 m over 1000 steps,
        # a tenth of the 0.1 mm print resolution
        assert off <= (0.01 if precision == _abi.PREC_F64 else 0.25) * total
    # the voxel part alone, and the link part alone
    only_voxels = b.history_record(s, 999, 1, 0)
    assert only_voxels.startswith("<<<Step999 Time:") and only_voxels.endswith("<<<>>>\n") and "|[[[" not in only_voxels
    only_links = b.history_record(s, 999, 0, 1)
    assert only_links.startswith("|[[[999]]]") and only_links.endswith("[[[]]]\n")
    assert only_links.count(";") == o.nl
    b.close()


def test_worker_prints_history_of_recorded_simulation(tmp_path):
    vxt, inp, base_text, names = make_task(str(tmp_path), n=1, stop_time=0.01)
    name = names[0]
    path = os.path.join(inp, name)
    text = open(path).read().replace("</VXD>", '<RecordHistory replace="VXA.Simulator.RecordHistory">'
                                               "<RecordStepSize>10</RecordStepSize><RecordLink>1</RecordLink>"
                                               "</RecordHistory></VXD>")
    with open(path, "w") as f:
        f.write(text)
    out = tmp_path / "report.xml"
    r = subprocess.run([WORKER, "-i", vxt, "-o", str(out)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    v = engine.Vxa(base_text, text)
    assert v.info.record_step_size == 10 and v.info.record_link == 1
    d = v.desc
    cols = [(d.materials[i].matid, tuple(d.materials[i].color)) for i in range(d.n_materials)]
    assert expected_header(cols, d.voxel_size) in r.stdout
    o = OSim.from_desc(d)
    o.step(1)
    rdt = o.status().recommended_dt
    real = int(10 / (10000 * rdt * d.dt_frac)) + 1
    line = "real_stepsize: %d ; recommendedTimeStep %f; d_v3->DtFrac %f . \n" % (real, rdt, d.dt_frac)
    assert line in r.stdout
    records = [ln for ln in r.stdout.split("\n") if ln.startswith("<<<Step")]
    total = o.step(1000000) + 1      # runs until the stop condition t > 0.01
    labels = [int(re.match(r"<<<Step(\d+) ", ln).group(1)) for ln in records]
    assert labels == [j for j in range(total) if j % real == 0]
    assert len(labels) >= 3
    # the last record against the oracle's state at that step
    o2 = OSim.from_desc(d)
    o2.step(labels[-1] + 1)
    matids = np.array([d.materials[d.material[i]].matid for i in range(d.n_voxels)])
    bad, off, n = differing(records[-1] + "\n", oracle_record(o2, labels[-1], matids, 1, 1))
    assert not bad, bad[:10]
    assert off <= 0.01 * n
