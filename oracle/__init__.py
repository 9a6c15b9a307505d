"""ctypes binding of the CPU FP64 oracle (oracle/diff_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg are the only users.  Nothing in abt_amd/ imports this.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

PI_R4 = 3.1415927410125732  # misc_submod.f90:12, REAL*4 literal widened


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64


def _setup(L):
    L.oracle_make_far_field_bins.argtypes = [_i64, _i64, _D, _D, _D, _D, _D]
    L.oracle_get_rotation_matrix2.argtypes = [_D, _D]
    L.oracle_karczewski.argtypes = [_D, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, _D, _D, _D]
    L.oracle_beam_frame.argtypes = [_D, _D, _D, ctypes.c_double, _i64, _D, _D, _D, _D, _D, _D,
                                    _D, ctypes.POINTER(ctypes.c_int)]
    L.oracle_contour_integral.argtypes = [ctypes.c_double, _D, _D, _D, _i64, _D, _i64, _D, _D,
                                          _D, _D]
    L.oracle_diffraction.argtypes = [_i64, _i64, _D, _D, _D, ctypes.c_double, _D, _D,
                                     ctypes.c_int, _D, _D, _D, ctypes.c_double, _i64, _D, _D]
    L.oracle_diff_main.argtypes = [_i64, _i64, _D, _D, ctypes.c_double, ctypes.c_int,
                                   _i64, _D, _D, _D, _D, _I64, _D,
                                   _i64, _D, _D, _D, _I64, _D, _D, _D, ctypes.c_int]
    L.oracle_ampl_to_mueller.argtypes = [_i64, _D, _D]
    L.oracle_simpne.argtypes = [_i64, _D, _D, _i64]
    L.oracle_simpne.restype = ctypes.c_double
    L.oracle_get_1d_mueller.argtypes = [_i64, _i64, _D, _D, _D]
    L.oracle_finalise.argtypes = [_i64, _i64, _D, _D, ctypes.c_double, _D, _D, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, _D, _D, _D, _D, _D, _D, _D]
    L.oracle_max_threads.restype = ctypes.c_int


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_far_field_bins(theta, phi):
    theta, phi = _f64(theta), _f64(phi)
    nphi, nth = len(phi), len(theta)
    x = np.empty(nphi * nth); y = np.empty_like(x); z = np.empty_like(x)
    lib().oracle_make_far_field_bins(nphi, nth, _p(theta), _p(phi), _p(x), _p(y), _p(z))
    return x, y, z


def get_rotation_matrix2(v0):
    """v0: (nv, 3) com-relative vertices -> (3, 3) rot (row i, column j)."""
    v0 = _f64(v0)
    out = np.empty(9)
    lib().oracle_get_rotation_matrix2(_p(v0), _p(out))
    return out.reshape(3, 3)


def karczewski(prop2, x3, y3, z3, r):
    prop2 = _f64(prop2)
    d = np.empty(4); m = np.empty(3); k = np.empty(3)
    lib().oracle_karczewski(_p(prop2), x3, y3, z3, r, _p(d), _p(m), _p(k))
    return d.reshape(2, 2), m, k


def beam_frame(ampl, vk7, prop, fov, verts):
    ampl, vk7, prop, verts = _f64(ampl), _f64(vk7), _f64(prop), _f64(verts)
    rot = np.empty(9); rot2 = np.empty(9); p2 = np.empty(3); e2 = np.empty(3); i2 = np.empty(3)
    com = np.empty(3)
    anti = ctypes.c_int(0)
    lib().oracle_beam_frame(_p(ampl), _p(vk7), _p(prop), fov, verts.shape[0], _p(verts),
                            _p(rot), _p(rot2), _p(p2), _p(e2), _p(i2), _p(com), ctypes.byref(anti))
    return dict(rot=rot.reshape(3, 3), rot2=rot2.reshape(3, 3), prop2=p2, perp2=e2,
                incidence2=i2, com=com, anti_parallel=bool(anti.value))


def contour_integral(lam, rot, rot2, v0, prop2, x3, y3, z3):
    rot, rot2, v0, prop2 = _f64(rot), _f64(rot2), _f64(v0), _f64(prop2)
    x3, y3, z3 = _f64(x3), _f64(y3), _f64(z3)
    out = np.empty(2 * len(x3))
    lib().oracle_contour_integral(lam, _p(rot), _p(rot2), _p(v0), v0.shape[0], _p(prop2),
                                  len(x3), _p(x3), _p(y3), _p(z3), _p(out))
    return out[0::2] + 1j * out[1::2]


def _ampl_to_complex(buf, nphi, nth):
    """(2*nbins*4) doubles -> complex array [r, c, theta, phi]."""
    a = buf.reshape(2, 2, nth, nphi, 2)  # [c, r, j, i, re/im]
    return (a[..., 0] + 1j * a[..., 1]).transpose(1, 0, 2, 3)


def diffraction(theta, phi, x, y, z, lam, beams, b, fast_diff=False):
    """One `diffraction` call for beam index b of a BeamSet. Returns complex [r, c, theta, phi]."""
    theta, phi = _f64(theta), _f64(phi)
    nphi, nth = len(phi), len(theta)
    out = np.empty(8 * nphi * nth)
    o0, o1 = beams.vert_off[b], beams.vert_off[b + 1]
    verts = _f64(beams.verts[o0:o1])
    lib().oracle_diffraction(nphi, nth, _p(_f64(x)), _p(_f64(y)), _p(_f64(z)), lam, _p(phi),
                             _p(theta), int(fast_diff), _p(_f64(beams.ampl[b])),
                             _p(_f64(beams.vk7[b])), _p(_f64(beams.prop[b])), float(beams.fov[b]),
                             o1 - o0, _p(verts), _p(out))
    return _ampl_to_complex(out, nphi, nth)


def diff_main(theta, phi, lam, beams, ext, fast_diff=False, nthreads=1, raw=False):
    """diff_main: returns (ampl_far_beam, ampl_far_ext_diff), complex [r, c, theta, phi]
    (or the raw Fortran-layout double buffers when raw=True)."""
    theta, phi = _f64(theta), _f64(phi)
    nphi, nth = len(phi), len(theta)
    ob = np.empty(8 * nphi * nth); oe = np.empty_like(ob)
    lib().oracle_diff_main(nphi, nth, _p(theta), _p(phi), lam, int(fast_diff),
                           beams.n, _p(beams.ampl), _p(beams.vk7), _p(beams.prop), _p(beams.fov),
                           _p(beams.vert_off, _I64), _p(beams.verts),
                           ext.n, _p(ext.ampl), _p(ext.vk7), _p(ext.prop),
                           _p(ext.vert_off, _I64), _p(ext.verts), _p(ob), _p(oe), nthreads)
    if raw:
        return ob, oe
    return _ampl_to_complex(ob, nphi, nth), _ampl_to_complex(oe, nphi, nth)


def complex_to_raw(a):
    """complex [r, c, theta, phi] -> Fortran-layout (phi, theta, 2, 2) (re, im) doubles."""
    t = np.ascontiguousarray(a.transpose(1, 0, 2, 3))  # [c, r, j, i]
    out = np.empty(t.shape + (2,))
    out[..., 0] = t.real
    out[..., 1] = t.imag
    return out.reshape(-1)


def ampl_to_mueller(raw, nbins):
    m = np.empty(16 * nbins)
    lib().oracle_ampl_to_mueller(nbins, _p(_f64(raw)), _p(m))
    return m.reshape(16, nbins)


def simpne(x, y):
    x, y = _f64(x), _f64(y)
    return lib().oracle_simpne(len(x), _p(x), _p(y), 1)


def get_1d_mueller(mueller, nphi, nth, phi):
    out = np.empty(16 * nth)
    lib().oracle_get_1d_mueller(nphi, nth, _p(_f64(mueller)), _p(_f64(phi)), _p(out))
    return out.reshape(16, nth)


def finalise(theta, phi, la, raw_beam, raw_ext, scaling=False, beam_energy_out=1.0,
             ext_energy_out=1.0):
    """finalise (outputs_mod.f90:63-278).  raw_* are Fortran-layout buffers (copied)."""
    theta, phi = _f64(theta), _f64(phi)
    nphi, nth = len(phi), len(theta)
    nb = nphi * nth
    rb = np.array(raw_beam, dtype=np.float64, copy=True)
    re = np.array(raw_ext, dtype=np.float64, copy=True)
    m = [np.empty(16 * nb) for _ in range(3)]
    m1 = [np.empty(16 * nth) for _ in range(3)]
    params = np.zeros(9)
    lib().oracle_finalise(nphi, nth, _p(theta), _p(phi), la, _p(rb), _p(re), int(scaling),
                          beam_energy_out, ext_energy_out, _p(m[0]), _p(m[1]), _p(m[2]),
                          _p(m1[0]), _p(m1[1]), _p(m1[2]), _p(params))
    keys = ["scatt", "scatt_beam", "scatt_ext_diff", "asymmetry", "asymmetry_beam",
            "asymmetry_ext_diff", "back_scatt", "scale_beam", "scale_ext"]
    return dict(mueller=m[0].reshape(16, nb), mueller_beam=m[1].reshape(16, nb),
                mueller_ext=m[2].reshape(16, nb), mueller_1d=m1[0].reshape(16, nth),
                mueller_1d_beam=m1[1].reshape(16, nth), mueller_1d_ext=m1[2].reshape(16, nth),
                ampl_beam=rb, ampl_ext=re, **dict(zip(keys, params)))


# ---- facet visibility (vis_oracle.c) ----

def _setup_vis(L):
    vp = ctypes.c_void_p
    L.oracle_vis_boxes.argtypes = [_i64, vp, vp, vp]
    L.oracle_find_vis.argtypes = [ctypes.c_int, _i64, vp, _i64, vp, vp, vp, vp, vp, _i64, vp, _i64, vp,
                                  ctypes.c_double, vp, vp, vp, vp, vp, vp]
    L.oracle_find_vis_batch.argtypes = [ctypes.c_int, _i64, _i64, vp, _i64, vp, vp, vp, vp, vp, _i64, vp,
                                        vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int]


_vis_ready = False


def _vlib():
    global _vis_ready
    L = lib()
    if not _vis_ready:
        _setup_vis(L)
        _vis_ready = True
    return L


def _a(x, dt):
    return np.ascontiguousarray(x, dtype=dt)


def vis_boxes(v):
    """Box midpoints of beam_aligned_bounding_boxes + compute_geometry_midpoints: (bx[16], by[16])."""
    vc = _a(np.asarray(v).T, np.float64)
    bx, by = np.empty(16), np.empty(16)
    _vlib().oracle_vis_boxes(vc.shape[1], vc.ctypes.data, bx.ctypes.data, by.ctypes.data)
    return bx, by


def find_vis(mode, geom, field_in=None, beam_mid_z=0.0, with_counts=False):
    """find_vis_int (mode 0) / find_vis_ext (1) / find_vis_inc (2) of one beam; geom has
    v (nv,3), f_off, f_vi (1-based), mid (nf,3), nz, ap (1-based), ap_nz."""
    mode = {"int": 0, "ext": 1, "inc": 2}.get(mode, mode)
    vc = _a(np.asarray(geom.v).T, np.float64)
    off, vi, mid = _a(geom.f_off, np.int64), _a(geom.f_vi, np.int64), _a(geom.mid, np.float64)
    nz, ap, apn = _a(geom.nz, np.float64), _a(geom.ap, np.int64), _a(geom.ap_nz, np.float64)
    fin = _a(field_in if field_in is not None else [0], np.int64)
    nfin = 0 if field_in is None else len(field_in)
    nf = len(off) - 1
    inb, vis, shad = np.zeros(nf, np.uint8), np.zeros(nf, np.uint8), np.zeros(nf, np.uint8)
    dist, idb, cnt = np.zeros(nf), np.zeros(nf, np.int64), np.zeros(2, np.int64)
    _vlib().oracle_find_vis(mode, vc.shape[1], vc.ctypes.data, nf, off.ctypes.data, vi.ctypes.data,
                            mid.ctypes.data, nz.ctypes.data, ap.ctypes.data, len(apn), apn.ctypes.data,
                            nfin, fin.ctypes.data, float(beam_mid_z), inb.ctypes.data, vis.ctypes.data,
                            dist.ctypes.data, idb.ctypes.data, shad.ctypes.data, cnt.ctypes.data)
    out = dict(in_beam=inb.astype(bool), is_vis=vis.astype(bool), dist_beam=dist, id_beam=idb,
               is_shad=shad.astype(bool))
    if with_counts:
        out["counts"] = cnt
    return out


def find_vis_batch(mode, topo, v, mid, nz, ap_nz, in_off, field_in, beam_mid_z=None, nthreads=0):
    """Many beams over one topology (OpenMP over beams): v (nbeam,3,nv), mid (nbeam,nf,3),
    nz (nbeam,nf), ap_nz (nbeam,na), in_off (nbeam+1), field_in (1-based).  Returns the
    outputs as (nbeam, nf) arrays and the per-beam counts (nbeam, 2)."""
    mode = {"int": 0, "ext": 1, "inc": 2}.get(mode, mode)
    v, mid, nz, apn = _a(v, np.float64), _a(mid, np.float64), _a(nz, np.float64), _a(ap_nz, np.float64)
    off, vi, ap = _a(topo.f_off, np.int64), _a(topo.f_vi, np.int64), _a(topo.ap, np.int64)
    ino, fin = _a(in_off, np.int64), _a(field_in if len(field_in) else [0], np.int64)
    nbeam, nvv = v.shape[0], v.shape[2]
    nf = len(off) - 1
    bmz = None if beam_mid_z is None else _a(beam_mid_z, np.float64)
    inb, vis, shad = (np.zeros((nbeam, nf), np.uint8) for _ in range(3))
    dist, idb, cnt = np.zeros((nbeam, nf)), np.zeros((nbeam, nf), np.int64), np.zeros((nbeam, 2), np.int64)
    _vlib().oracle_find_vis_batch(mode, nbeam, nvv, v.ctypes.data, nf, off.ctypes.data, vi.ctypes.data,
                                  mid.ctypes.data, nz.ctypes.data, ap.ctypes.data, apn.shape[1],
                                  apn.ctypes.data, ino.ctypes.data, fin.ctypes.data,
                                  None if bmz is None else bmz.ctypes.data, inb.ctypes.data,
                                  vis.ctypes.data, dist.ctypes.data, idb.ctypes.data, shad.ctypes.data,
                                  cnt.ctypes.data, int(nthreads))
    return dict(in_beam=inb.astype(bool), is_vis=vis.astype(bool), dist_beam=dist, id_beam=idb,
                is_shad=shad.astype(bool)), cnt


def vis_rotation(prop):
    """rotate()'s rotation (beam_loop_mod.f90:1488-1503) taking prop to -z: (3, 3)."""
    L = _vlib()
    L.oracle_vis_rotation.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    p = _a(prop, np.float64)
    R = np.empty((3, 3))
    L.oracle_vis_rotation(p.ctypes.data, R.ctypes.data)
    return R


def vis_rotate(R, geom):
    """The geometry fields the visibility routines read, rotated by R as rotate() does
    (beam_loop_mod.f90:1505-1524): (v (nv,3), mid (nf,3), nz (nf), ap_nz (na))."""
    L = _vlib()
    vp = ctypes.c_void_p
    L.oracle_vis_rotate.argtypes = [vp, _i64, vp, _i64, vp, vp, _i64, vp, vp, vp, vp, vp]
    Rm = _a(R, np.float64)
    v0 = _a(np.asarray(geom.v).T, np.float64)
    mid0, n0, apn0 = _a(geom.mid, np.float64), _a(geom.n, np.float64), _a(geom.ap_n, np.float64)
    nv, nf, na = v0.shape[1], len(mid0), len(apn0)
    v, mid, nz, apz = np.empty((3, nv)), np.empty((nf, 3)), np.empty(nf), np.empty(na)
    L.oracle_vis_rotate(Rm.ctypes.data, nv, v0.ctypes.data, nf, mid0.ctypes.data, n0.ctypes.data, na,
                        apn0.ctypes.data, v.ctypes.data, mid.ctypes.data, nz.ctypes.data, apz.ctypes.data)
    return v.T.copy(), mid, nz, apz
