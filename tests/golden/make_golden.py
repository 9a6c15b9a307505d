"""Writes tests/golden/diff_small.npz: small far-field cases computed by the CPU FP64
oracle, committed so that later builds are checked against fixed vectors.

The reference ships no fixtures for the diffraction path (its only test covers
geometry input), so these vectors come from the oracle, which is itself pinned by
closed forms and invariants (tests/test_oracle_pins.py).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from abt_amd.beams import concat  # noqa: E402
from abt_amd.workloads import column_beams, euler_matrix, grid, hex_column, random_triangles  # noqa: E402

LAM = 0.532


def main():
    th, ph = grid((0, 10, 180), (0, 30, 330))            # 19 x 12
    v, f = hex_column(5.0, 10.0, 6, 12, 0.0)
    cb, ce, _ = column_beams(v, f, euler_matrix(0, 30, 0), rec=8)
    beams = concat([random_triangles(5, seed=7, box=12.0), cb.subset(np.arange(0, cb.n, 400))])
    ext = ce.subset(np.arange(0, ce.n, 100))
    rb, re = oracle.diff_main(th, ph, LAM, beams, ext, raw=True)
    fin = oracle.finalise(th, ph, LAM, rb, re)
    out = dict(theta=th, phi=ph, lam=np.float64(LAM),
               b_ampl=beams.ampl, b_vk7=beams.vk7, b_prop=beams.prop, b_fov=beams.fov,
               b_vert_off=beams.vert_off, b_verts=beams.verts,
               e_ampl=ext.ampl, e_vk7=ext.vk7, e_prop=ext.prop, e_fov=ext.fov,
               e_vert_off=ext.vert_off, e_verts=ext.verts,
               ampl_far_beam=rb, ampl_far_ext_diff=re, mueller=fin["mueller"],
               mueller_1d=fin["mueller_1d"],
               params=np.array([fin[k] for k in ("scatt", "scatt_beam", "scatt_ext_diff", "asymmetry",
                                                 "asymmetry_beam", "asymmetry_ext_diff", "back_scatt")]))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "diff_small.npz")
    np.savez_compressed(path, **out)
    print(path, beams.n, ext.n)


if __name__ == "__main__":
    main()
