"""Writes tests/golden/vis_small.npz: facet-visibility cases (find_vis_int / find_vis_ext /
find_vis_inc, beam_loop_mod.f90:2037-2609) computed by the CPU oracle, committed so that later
builds are checked against fixed vectors.

The reference ships no fixtures for its visibility routines, so these vectors come from the
oracle, itself pinned by closed forms, brute-force occlusion and planted mistakes
(tests/test_vis_oracle.py).

Run from the repository root:  python tests/golden/make_golden_vis.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from abt_amd.visibility import VisGeometry  # noqa: E402
from abt_amd.workloads import hex_column_apertures, random_rotations  # noqa: E402


def _cases():
    """(name, geometry, field_in, beam_mid_z): a rough column in two orientations and a
    two-body layered scene."""
    out = []
    v, f, a, n = hex_column_apertures(4.0, 8.0, 3, 4, sigma=0.3, seed=5)
    topo = VisGeometry.from_mesh(v, f, a, n)
    rng = np.random.default_rng(11)
    for k, R in enumerate(random_rotations(2, 21)):
        g = topo.rotated(R)
        up = np.nonzero(g.ap_nz > 0.01)[0] + 1
        fin = np.nonzero(g.ap == rng.choice(up))[0] + 1
        out.append((f"column{k}", g, fin, g.v[:, 2].max() + 1.0))
    parts = [(2.0, 1.0, 2, 1, (2.0, 1.5, -1.0)), (6.0, 8.0, 3, 2, (0.0, 0.0, 0.0))]
    verts, faces, ap, apn, nv, na = [], [], [], [], 0, 0
    for radius, length, nfhr, nfpl, off in parts:
        v, f, a, n = hex_column_apertures(radius, length, nfhr, nfpl)
        verts.append(v + np.asarray(off))
        faces.append(f + nv)
        ap.append(a + na)
        apn.append(n)
        nv += len(v)
        na += len(n)
    c, s = np.cos(np.radians(4.0)), np.sin(np.radians(4.0))
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    g = VisGeometry.from_mesh(np.concatenate(verts), np.concatenate(faces), np.concatenate(ap),
                              np.concatenate(apn)).rotated(R)
    fin = np.nonzero(np.isin(g.ap, (8, 16)))[0] + 1
    out.append(("layered", g, fin, g.v[:, 2].max() + 1.0))
    # an external beam leaving the bottom face of an upper column onto a lower one
    parts = [(6.0, 3.0, 3, 2, (0.0, 0.0, 0.0)), (3.0, 2.0, 2, 1, (2.0, 0.5, 6.0))]
    verts, faces, ap, apn, nv, na = [], [], [], [], 0, 0
    for radius, length, nfhr, nfpl, off in parts:
        v, f, a, n = hex_column_apertures(radius, length, nfhr, nfpl)
        verts.append(v + np.asarray(off))
        faces.append(f + nv)
        ap.append(a + na)
        apn.append(n)
        nv += len(v)
        na += len(n)
    g = VisGeometry.from_mesh(np.concatenate(verts), np.concatenate(faces), np.concatenate(ap),
                              np.concatenate(apn)).rotated(R)
    fin = np.nonzero(g.ap == 15)[0] + 1
    out.append(("layered_ext", g, fin, g.v[:, 2].max() + 1.0))
    return out


def main():
    d = {}
    for name, g, fin, bz in _cases():
        d[f"{name}_v"], d[f"{name}_f_off"], d[f"{name}_f_vi"] = g.v, g.f_off, g.f_vi
        d[f"{name}_mid"], d[f"{name}_nz"], d[f"{name}_ap"], d[f"{name}_ap_nz"] = g.mid, g.nz, g.ap, g.ap_nz
        d[f"{name}_field_in"], d[f"{name}_beam_mid_z"] = fin, np.float64(bz)
        for mode in ("int", "ext", "inc"):
            r = oracle.find_vis(mode, g, None if mode == "inc" else fin, bz)
            for k, x in r.items():
                d[f"{name}_{mode}_{k}"] = x
    path = os.path.join(ROOT, "tests", "golden", "vis_small.npz")
    np.savez_compressed(path, names=np.array([c[0] for c in _cases()]), **d)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
