"""A task directory the way the reference's experiment scripts lay one out: base.vxa
plus one VXD per robot that replaces VXA.VXC.Structure, and a .vxt task file
(Voxelyze3/src/Executables/Voxelyze3.cpp:101-111)."""
import os
import re

from gpuvoxels_amd import vxa, workloads
from gpuvoxels_amd.model import MT


def structure_vxd(model):
    text = vxa.model_to_vxa(model)
    m = re.search(r'<Structure Compression="([^"]*)">(.*?)</Structure>', text, re.S)
    return ('<?xml version="1.0" encoding="ISO-8859-1"?>\n<VXD><Structure replace="VXA.VXC.Structure" '
            'Compression="%s">%s</Structure></VXD>\n' % (m.group(1), m.group(2)))


def make_task(root, n=3, seed=21, stop_time=0.01, save_positions=False, collisions=False):
    robots = workloads.robots(n, seed=seed, grid=8, steps=0)
    base = robots[0]
    base.stop_tokens = [(MT["mtCONST"], float(stop_time)), (MT["mtVAR"], 4.0), (MT["mtGREATERTHAN"], 0.0),
                        (MT["mtEND"], 0.0)]
    base.save_position_of_all_voxels = int(save_positions)
    base.collisions_enabled = int(collisions)
    inp = os.path.join(root, "generation_0001")
    os.makedirs(inp, exist_ok=True)
    base_text = vxa.model_to_vxa(base)
    with open(os.path.join(inp, "base.vxa"), "w") as f:
        f.write(base_text)
    names = []
    for i, r in enumerate(robots):
        name = "robot_%03d.vxd" % i
        with open(os.path.join(inp, name), "w") as f:
            f.write(structure_vxd(r))
        names.append(name)
    vxt = os.path.join(root, "task.vxt")
    with open(vxt, "w") as f:
        f.write('<?xml version="1.0" encoding="utf-8"?>\n<vxa>%s</vxa><input_dir>%s</input_dir><vxd>%s</vxd>\n'
                % (os.path.join(inp, "base.vxa"), inp, "".join("<f>%s</f>" % n for n in names)))
    return vxt, inp, base_text, names
