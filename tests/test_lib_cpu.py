"""CPU checks of the C-ABI library: it loads, exports every entry point the header
declares, and its host-side pieces (grid, per-beam frame set-up) agree with the
oracle bit for bit.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from abt_amd import lib
from abt_amd.workloads import random_triangles, stress_beams, grid, GRIDS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LAM = 0.532


def _declared():
    src = "".join(open(os.path.join(ROOT, "include", h)).read()
                  for h in sorted(os.listdir(os.path.join(ROOT, "include"))) if h.endswith(".h"))
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(abt_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    declared = _declared()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(lib.EXPORTS) == declared
    assert L.abt_version() == 1


def test_make_far_field_bins_matches_oracle_bitwise(orc):
    th, ph = grid(*GRIDS["small"])
    a = lib.make_far_field_bins(th, ph)
    b = orc.make_far_field_bins(th, ph)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("maker", ["random", "stress"])
def test_host_frames_match_oracle_bitwise(orc, maker):
    bs = random_triangles(40, seed=11) if maker == "random" else stress_beams(40, seed=5)
    recs, _ = lib.prepare_beams_host(bs, LAM)
    R = lib.load().abt_record_doubles()
    assert recs.shape == (40, R)
    for b in range(bs.n):
        o0, o1 = bs.vert_off[b], bs.vert_off[b + 1]
        fr = orc.beam_frame(bs.ampl[b], bs.vk7[b], bs.prop[b], bs.fov[b], bs.verts[o0:o1])
        r = recs[b]
        np.testing.assert_array_equal(r[0:3], fr["com"])
        np.testing.assert_array_equal(r[3:12].reshape(3, 3), _rr(fr))
        np.testing.assert_array_equal(r[12:15], fr["prop2"])
        np.testing.assert_array_equal(r[15:18], fr["incidence2"])
        sign = -1.0 if fr["anti_parallel"] else 1.0
        np.testing.assert_array_equal(r[18:26], sign * bs.ampl[b])


def _rr(fr):
    # matmul(rot2, rot) in the reference's summation order (k = 1, 2, 3)
    a, b = fr["rot2"], fr["rot"]
    out = np.empty((3, 3))
    for i in range(3):
        for j in range(3):
            out[i, j] = (a[i, 0] * b[0, j] + a[i, 1] * b[1, j]) + a[i, 2] * b[2, j]
    return out


def test_normal_incidence_frames_match_oracle(orc):
    """Apertures normal to prop: the in-plane frame angle comes from rounding noise
    (atan2 of ~1e-17 components), so agreement here needs the same libm and
    operation order as the oracle."""
    # stress_beams(2000, seed=7) beams 72, 126, 642, 1238, 1252, 1920: angles where glibc's
    # sincos and separate cos/sin calls differ in the last bit
    hard = stress_beams(2000, seed=7, min_angle=0.3).subset([72, 126, 642, 1238, 1252, 1920])
    for bs in (stress_beams(200, seed=9), hard):
        recs, _ = lib.prepare_beams_host(bs, LAM)
        agree = 0
        for b in range(bs.n):
            o0, o1 = bs.vert_off[b], bs.vert_off[b + 1]
            fr = orc.beam_frame(bs.ampl[b], bs.vk7[b], bs.prop[b], bs.fov[b], bs.verts[o0:o1])
            agree += np.array_equal(recs[b, 3:12].reshape(3, 3), _rr(fr))
        assert agree == bs.n


def test_polygon_records_need_vertex_records():
    bs = random_triangles(3, seed=1)
    # make beam 1 a planar quadrilateral
    v = bs.verts.reshape(3, 3, 3).copy()
    quad = np.vstack([v[1], (v[1][0] + v[1][2]) / 2 + (v[1][2] - v[1][1]) * 0.3])
    verts = np.vstack([v[0], quad, v[2]])
    off = np.array([0, 3, 7, 10])
    from abt_amd.beams import BeamSet
    pb = BeamSet(bs.ampl, bs.vk7, bs.prop, bs.fov, off, verts)
    recs, vrecs = lib.prepare_beams_host(pb, LAM)
    assert recs[1, 30] == 4 and recs[1, 31] == 3
    assert np.any(vrecs[3:7] != 0)


def test_bad_arguments_raise():
    from abt_amd.beams import BeamSet
    bs = random_triangles(2, seed=1)
    with pytest.raises(lib.AbtError, match="wavelength"):
        lib.prepare_beams_host(bs, -1.0)
    # diff_main needs at least one of its two outputs (rejected before any device work)
    th = np.radians([0.0, 1.0, 2.0])
    with pytest.raises(lib.AbtError, match="null argument"):
        lib.diff_main(th, th, LAM, bs, BeamSet.empty(), trees=())


def test_beam_set_files_round_trip(tmp_path):
    """npz and JSON forms of a beam set (triangles and a quadrilateral) round-trip exactly;
    the JSON amplitude matrix is written row by row: [[a11, a12], [a21, a22]]."""
    from abt_amd.beams import BeamSet, concat
    tri = random_triangles(5, seed=3, box=10.0)
    quad = BeamSet(np.arange(8.0)[None] + 0.25, [[0.0, 1.0, 0.0]], [[0.0, 0.0, -1.0]], [0.7], [0, 4],
                   [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.0]])
    bs = concat([tri, quad])
    for save, load, name in ((BeamSet.save, BeamSet.load, "b.npz"), (BeamSet.to_json, BeamSet.from_json, "b.json")):
        save(bs, tmp_path / name)
        got = load(tmp_path / name)
        for f in ("ampl", "vk7", "prop", "fov", "vert_off", "verts"):
            np.testing.assert_array_equal(getattr(got, f), getattr(bs, f))
    import json
    rec = json.load(open(tmp_path / "b.json"))["outbeams"][-1]
    a = bs.ampl_complex()[-1]
    for r in range(2):
        for c in range(2):
            assert rec["ampl"][r][c] == [a[r, c].real, a[r, c].imag]
    assert rec["prop_out"] == [0.0, 0.0, -1.0] and len(rec["verts"]) == 4
    empty = tmp_path / "e.json"
    BeamSet.empty().to_json(empty)
    assert BeamSet.from_json(empty).n == 0


def test_outbeams_from_pbt_beam_export():
    """The reference's -export_beam file (tests/golden/beam_export.json, written in its layout
    by tests/golden/make_beam_json.py from a seeded trace): one outgoing beam per illuminating
    facet of each external beam, internal beams skipped, amplitudes / e-perp / prop / facet
    vertices exactly as traced, and fov = 8 lambda / d with d the diameter of the circle of
    the beam's summed facet area (add_to_outbeam_tree, beam_loop_mod.f90:2759-2767)."""
    import json
    import os
    from abt_amd.beams import PI_R4 as PI4, outbeams_from_pbt_json
    from tests.golden.make_beam_json import case, groups
    path = os.path.join(os.path.dirname(__file__), "golden", "beam_export.json")
    V, F, b, fi = case()
    got, gfi, ids = outbeams_from_pbt_json(path, V, F, 0.532, with_ids=True)
    assert got.n == b.n
    np.testing.assert_array_equal(gfi, fi)
    np.testing.assert_array_equal(got.ampl, b.ampl)
    np.testing.assert_array_equal(got.vk7, b.vk7)
    np.testing.assert_array_equal(got.verts, b.verts)
    tree = json.load(open(path))["beam tree"]
    assert sum(1 for t in tree if t["is_int"]) > 0
    area = lambda f: 0.5 * np.linalg.norm(np.cross(V[F[f - 1][1]] - V[F[f - 1][0]], V[F[f - 1][2]] - V[F[f - 1][0]]))
    for grp in groups(b.n):
        a = sum(area(int(fi[i])) for i in grp)
        want = 0.532 / (2 * np.sqrt(a / PI4)) * 8
        for i in grp:
            np.testing.assert_allclose(got.fov[i], want, rtol=1e-14)
            np.testing.assert_array_equal(got.prop[i], b.prop[grp[0]])
    assert len(set(ids[[4, 5, 6]])) == 1 and len(set(ids[[20, 21]])) == 1 and ids[3] != ids[4]
