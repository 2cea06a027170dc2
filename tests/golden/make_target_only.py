#!/usr/bin/env python3
"""Writes tests/golden/target_only.npz: the input documents of the reference's
Experiments/xenobot/h.target_only/data1 run (base.vxa and bot_0000.vxd: a 112x112x7
lattice holding a ciliated, signalling xenobot of 2,964 voxels among sticky target
material, with collisions, attach, detach and close-pair counting on), stored
compressed so that the GPU tests can run the experiment where the reference tree is
absent.  The documents are data (inputs); the expected results come from the oracle
at test time.

    python tests/golden/make_target_only.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/Experiments/xenobot/h.target_only/data1"

if __name__ == "__main__":
    vxa = open(os.path.join(SRC, "base.vxa"), "rb").read()
    vxd = open(os.path.join(SRC, "bot_0000.vxd"), "rb").read()
    np.savez_compressed(os.path.join(HERE, "target_only.npz"), vxa=np.frombuffer(vxa, np.uint8),
                        vxd=np.frombuffer(vxd, np.uint8))
    print("wrote target_only.npz: base.vxa %d B, bot_0000.vxd %d B" % (len(vxa), len(vxd)))
