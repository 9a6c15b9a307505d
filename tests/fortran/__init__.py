"""Runs the Fortran drop-in (tests/fortran/_build/abt_driver, linking
abt_amd/fortran/diff_mod.f90 compiled against the stand-in tests/fortran/types_mod.f90)
on a beam set and reads its outputs back.  Test infrastructure; the drivers are built
by __graft_entry__.build() (make -C tests/fortran)."""
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


VIS_DRIVER = os.path.join(HERE, "_build", "abt_vis_driver")


def run_vis(g, field_in, beam_mid_z, workdir=None):
    """Runs find_vis_int, find_vis_ext and find_vis_inc of the Fortran drop-in vis_mod on the
    geometry g (abt_amd.visibility.VisGeometry, facet normals given only by their z part: the
    x, y parts are written as 0) with the 1-based illuminating facets field_in.  Returns a
    dict per mode of in_beam, is_vis (inc only), dist_beam, id_beam, is_shad."""
    workdir = workdir or tempfile.mkdtemp(prefix="abt_v_")
    fin = os.path.join(workdir, "vin.bin")
    fout = os.path.join(workdir, "vout.bin")
    nf = g.nf
    with open(fin, "wb") as f:
        np.array([g.nv, nf, g.na, len(field_in)], np.int64).tofile(f)
        np.asarray(g.v, np.float64).T.copy().tofile(f)          # v(nv, 3), column-major
        np.diff(g.f_off).astype(np.int64).tofile(f)
        np.asarray(g.f_vi, np.int64).tofile(f)
        np.asarray(g.mid, np.float64).tofile(f)                 # mid(3, nf)
        n = np.zeros((nf, 3))
        n[:, 2] = g.nz
        n.T.copy().tofile(f)                                    # n(nf, 3)
        np.asarray(g.ap, np.int64).tofile(f)
        a = np.zeros((g.na, 3))
        a[:, 2] = g.ap_nz
        a.T.copy().tofile(f)                                    # ap normals (na, 3)
        np.asarray(field_in, np.int64).tofile(f)
        np.array([beam_mid_z], np.float64).tofile(f)
    subprocess.run([VIS_DRIVER, fin, fout], check=True, timeout=600)
    raw = open(fout, "rb").read()
    out, off = {}, 0
    for mode in ("int", "ext", "inc"):
        r = {}
        for key, dt in (("in_beam", np.int8), ("is_vis", np.int8), ("dist_beam", np.float64),
                        ("id_beam", np.int64), ("is_shad", np.int8)):
            n_b = nf * np.dtype(dt).itemsize
            r[key] = np.frombuffer(raw[off:off + n_b], dtype=dt).copy()
            off += n_b
        for key in ("in_beam", "is_vis", "is_shad"):
            r[key] = r[key] != 0
        out[mode] = r
    return out
