"""Runs the Fortran drop-in (tests/fortran/_build/abt_driver, linking
abt_amd/fortran/diff_mod.f90 and the reference's types_mod) on a beam set and
reads its outputs back.  Test infrastructure."""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_build", "abt_driver")


def run(theta, phi, lam, verts, faces, b_fi, b_ampl, b_vk7, b_prop, b_fov, e_fi, e_ampl, e_vk7,
        e_prop, fast_diff=False, precision="fp64", workdir=None):
    """verts (nv,3); faces (nf,3) 0-based; *_fi 1-based facet ids.  Returns dict of
    complex [r, c, theta, phi] arrays ampl_far_beam, ampl_far_ext_diff, amplc
    (diffraction of beam 1) and the x, y, z bins, plus the log text."""
    nth, nph = len(theta), len(phi)
    workdir = workdir or tempfile.mkdtemp(prefix="abt_f_")
    fin = os.path.join(workdir, "in.bin")
    fout = os.path.join(workdir, "out.bin")
    flog = os.path.join(workdir, "log.txt")
    nb, ne = len(b_fi), len(e_fi)
    with open(fin, "wb") as f:
        np.array([nth, nph, len(verts), len(faces), nb, ne, int(fast_diff)], np.int64).tofile(f)
        np.array([lam], np.float64).tofile(f)
        np.asarray(theta, np.float64).tofile(f)
        np.asarray(phi, np.float64).tofile(f)
        np.asarray(verts, np.float64).T.copy().tofile(f)        # v(nv, 3), column-major
        np.full(len(faces), faces.shape[1], np.int64).tofile(f)
        (np.asarray(faces, np.int64) + 1).ravel().tofile(f)
        for fi, a, v, p, fv, n in ((b_fi, b_ampl, b_vk7, b_prop, b_fov, nb),
                                   (e_fi, e_ampl, e_vk7, e_prop, np.full(ne, 1.0), ne)):
            if n == 0:
                continue
            np.asarray(a, np.float64).ravel().tofile(f)
            np.asarray(v, np.float64).ravel().tofile(f)
            np.asarray(p, np.float64).ravel().tofile(f)
            np.asarray(fv, np.float64).ravel().tofile(f)
            np.asarray(fi, np.int64).tofile(f)
    env = dict(os.environ, ABT_PRECISION=precision)
    subprocess.run([DRIVER, fin, fout, flog], check=True, env=env, timeout=600)
    raw = np.fromfile(fout, dtype=np.float64)
    nb_ = nth * nph
    parts, off = {}, 0
    for key, size in (("ampl_far_beam", 8 * nb_), ("ampl_far_ext_diff", 8 * nb_), ("x", nb_),
                      ("y", nb_), ("z", nb_), ("amplc", 8 * nb_)):
        parts[key] = raw[off:off + size]
        off += size
    from abt_amd.lib import raw_to_complex
    out = {k: raw_to_complex(parts[k], nph, nth) for k in ("ampl_far_beam", "ampl_far_ext_diff",
                                                             "amplc")}
    out.update({k: parts[k] for k in ("x", "y", "z")})
    out["log"] = open(flog).read()
    return out
