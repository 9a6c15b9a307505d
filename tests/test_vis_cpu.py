"""CPU checks of the visibility boundary (include/abt_vis.h): argument validation of the
C-ABI entry points (which reject bad input before any device call), the workspace size,
and the host-side geometry helpers.  No GPU calls."""
import ctypes

import numpy as np
import pytest

from abt_amd import lib
from abt_amd import visibility as V
from abt_amd.workloads import hex_column_apertures, vis_workload


def _geom():
    v, f, a, n = hex_column_apertures(3.0, 4.0, 2, 2)
    return V.VisGeometry.from_mesh(v, f, a, n)


def _call(mode, g, field_in, **over):
    L = V._lib()
    keep = dict(v=np.ascontiguousarray(g.v.T), off=np.ascontiguousarray(g.f_off, np.int64),
                vi=np.ascontiguousarray(g.f_vi, np.int64), mid=np.ascontiguousarray(g.mid),
                nz=np.ascontiguousarray(g.nz), ap=np.ascontiguousarray(g.ap, np.int64),
                apn=np.ascontiguousarray(g.ap_nz))
    keep.update(over)
    geo = V._Geo(g.nv, keep["v"].ctypes.data, g.nf, keep["off"].ctypes.data, keep["vi"].ctypes.data,
                 keep["mid"].ctypes.data, keep["nz"].ctypes.data, keep["ap"].ctypes.data, g.na,
                 keep["apn"].ctypes.data)
    fin = np.ascontiguousarray(field_in if field_in is not None else np.zeros(1), np.int64)
    out = [np.zeros(g.nf, np.uint8), np.zeros(g.nf, np.uint8), np.zeros(g.nf), np.zeros(g.nf, np.int64),
           np.zeros(g.nf, np.uint8)]
    rc = L.abt_find_vis(mode, ctypes.byref(geo), 0 if field_in is None else len(field_in), fin.ctypes.data,
                        1.0, *[o.ctypes.data for o in out])
    return rc, L.abt_last_error().decode()


@pytest.mark.parametrize("case,expect", [
    ("mode", "unknown visibility mode"),
    ("vertex", "vertex id"),
    ("aperture", "aperture id"),
    ("field_in", "field_in facet"),
    ("no_field_in", "at least one illuminating facet"),
    ("offsets", "f_off[0]"),
    ("two_vertices", "fewer than 3 vertices"),
])
def test_find_vis_rejects_bad_input(case, expect):
    g = _geom()
    f = np.array([1, 2, 3])
    mode, over = V.ABT_VIS_INT, {}
    if case == "mode":
        mode = 7
    elif case == "vertex":
        vi = g.f_vi.copy(); vi[5] = g.nv + 1; over["vi"] = vi
    elif case == "aperture":
        ap = g.ap.copy(); ap[0] = 0; over["ap"] = ap
    elif case == "field_in":
        f = np.array([1, g.nf + 1])
    elif case == "no_field_in":
        f = None
    elif case == "offsets":
        off = g.f_off.copy(); off[0] = 1; over["off"] = off
    elif case == "two_vertices":
        off = g.f_off.copy(); off[1] = 2; over["off"] = off
    rc, msg = _call(mode, g, f, **over)
    assert rc == 1 and expect in msg, (rc, msg)


def test_batched_entry_rejects_bad_input():
    L = V._lib()
    b = V._Batch()
    b.mode, b.nbeam, b.nv, b.nf, b.na = 0, 1, 12, 10, 8
    assert L.abt_find_vis_dev(ctypes.byref(b), 30, None, 0, None) == 1
    assert "null array" in L.abt_last_error().decode()
    b.mode = 5
    assert L.abt_find_vis_dev(ctypes.byref(b), 30, None, 0, None) == 1
    assert "mode" in L.abt_last_error().decode()


def test_workspace_grows_with_beams_and_facets():
    L = V._lib()
    w1 = L.abt_find_vis_workspace(1, 2112, 3 * 2112)
    w2 = L.abt_find_vis_workspace(10, 2112, 3 * 2112)
    # per beam: flags (3 nf int32), candidate lists (3 nf int2), tested-facet lists (nf int32), box data
    assert w1 >= (3 * 4 + 3 * 8 + 4) * 2112
    assert w2 >= 10 * (w1 - 8 * 256) and w2 % 256 == 0


def test_geometry_midpoints_and_normals():
    """compute_geometry_midpoints / _normals (misc_submod.f90:199-232): vertex mean and the unit
    normal of the first two edges; outward on the closed column."""
    g = _geom()
    for i in range(g.nf):
        ids = g.f_vi[g.f_off[i]:g.f_off[i + 1]] - 1
        np.testing.assert_allclose(g.mid[i], g.v[ids].mean(0), rtol=0, atol=1e-14)
    c = g.mid - g.mid.mean(0)
    n3 = np.array([np.cross(g.v[g.f_vi[g.f_off[i] + 1] - 1] - g.v[g.f_vi[g.f_off[i]] - 1],
                            g.v[g.f_vi[g.f_off[i] + 2] - 1] - g.v[g.f_vi[g.f_off[i] + 1] - 1]) for i in range(g.nf)])
    assert ((n3 * c).sum(1) > 0).all()
    np.testing.assert_allclose(g.nz, n3[:, 2] / np.linalg.norm(n3, axis=1), atol=1e-15)


def test_vis_workload_shapes():
    topo, v, mid, nz, apn, ino, fin, bmz = vis_workload(3, "ext", seed=4)
    assert v.shape == (3, 3, topo.nv) and mid.shape == (3, topo.nf, 3) and apn.shape == (3, topo.na)
    assert ino[0] == 0 and ino[-1] == len(fin) and (np.diff(ino) > 0).all()
    assert fin.min() >= 1 and fin.max() <= topo.nf
    for b in range(3):   # the illuminating facets are one down-facing aperture
        a = np.unique(topo.ap[fin[ino[b]:ino[b + 1]] - 1])
        assert len(a) == 1 and apn[b, a[0] - 1] < -0.01


def test_host_rotations_match_oracle_bitwise(orc):
    """abt_vis_rotations (host, C library) against the oracle's restatement of rotate()."""
    rng = np.random.default_rng(9)
    p = rng.normal(size=(50, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    R = V.vis_rotations(p)
    for b in range(len(p)):
        np.testing.assert_array_equal(R[b], orc.vis_rotation(p[b]))


def test_rotation_entry_points_reject_bad_input():
    L = V._lib()
    assert L.abt_vis_rotations(-1, None, None) == 1
    assert L.abt_vis_rotations(2, None, None) == 1
    assert L.abt_vis_rotate_dev(1, None, 4, None, 4, None, None, 1, None, None, None, None, None, None) == 1
    assert "null array" in L.abt_last_error().decode()
    assert L.abt_vis_rotate_dev(1, None, 0, None, 4, None, None, 1, None, None, None, None, None, None) == 1
    assert "sizes" in L.abt_last_error().decode()
