"""Committed fixtures of the far-field path (see make_golden.py)."""
import os

import numpy as np

from abt_amd.beams import BeamSet

HERE = os.path.dirname(os.path.abspath(__file__))


def load_small():
    d = np.load(os.path.join(HERE, "diff_small.npz"))
    tree = lambda p: BeamSet(d[p + "ampl"], d[p + "vk7"], d[p + "prop"], d[p + "fov"], d[p + "vert_off"],
                             d[p + "verts"])
    return d, tree("b_"), tree("e_")
