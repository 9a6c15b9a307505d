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


def load_vis():
    """[(name, VisGeometry, field_in, beam_mid_z, {mode: outputs})] of vis_small.npz."""
    from abt_amd.visibility import VisGeometry
    d = np.load(os.path.join(HERE, "vis_small.npz"))
    out = []
    for name in d["names"]:
        p = f"{name}_"
        g = VisGeometry(d[p + "v"], d[p + "f_off"], d[p + "f_vi"], d[p + "mid"], d[p + "nz"], d[p + "ap"],
                        d[p + "ap_nz"])
        res = {m: {k: d[f"{p}{m}_{k}"] for k in ("in_beam", "is_vis", "dist_beam", "id_beam", "is_shad")}
               for m in ("int", "ext", "inc")}
        out.append((str(name), g, d[p + "field_in"], float(d[p + "beam_mid_z"]), res))
    return out
