"""Facet visibility of the near-field tracer on the GPU (include/abt_vis.h).

Mirrors the reference's three visibility subroutines of beam_loop_mod:
  find_vis_int(beam, rot_geometry)  -> in_beam, dist_beam, id_beam, is_shad   (beam_loop_mod.f90:2037-2247)
  find_vis_ext(beam, rot_geometry)  -> in_beam, dist_beam, id_beam, is_shad   (:2249-2459)
  find_vis_inc(beam, geometry, beam_geometry)
                                    -> in_beam, is_vis, dist_beam, id_beam, is_shad   (:2461-2609)
`beam` is given here by its illuminating facet ids (beam%field_in(:)%fi, 1-based) and
`beam_geometry` by beam_geometry%f(1)%mid(3).  Every call runs on the HIP library; there
is no CPU fallback.  `find_vis_batch_dev` evaluates many beams (one rotated geometry each,
shared facet topology) in one launch sequence on device tensors.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import lib

ABT_VIS_INT, ABT_VIS_EXT, ABT_VIS_INC = 0, 1, 2
MODES = {"int": ABT_VIS_INT, "ext": ABT_VIS_EXT, "inc": ABT_VIS_INC}


class _Geo(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int64), ("v", ctypes.c_void_p), ("nf", ctypes.c_int64),
                ("f_off", ctypes.c_void_p), ("f_vi", ctypes.c_void_p), ("f_mid", ctypes.c_void_p),
                ("f_nz", ctypes.c_void_p), ("f_ap", ctypes.c_void_p), ("na", ctypes.c_int64),
                ("ap_nz", ctypes.c_void_p)]


class _Batch(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("nbeam", ctypes.c_int64), ("nv", ctypes.c_int64),
                ("nf", ctypes.c_int64), ("na", ctypes.c_int64)] + \
               [(n, ctypes.c_void_p) for n in ("f_off", "f_vi", "f_ap", "v", "f_mid", "f_nz", "ap_nz",
                                              "in_off", "field_in", "beam_mid_z", "in_beam", "is_vis",
                                              "dist_beam", "id_beam", "is_shad", "counts")] + \
               [("phases", ctypes.c_int)]


_bound = False


def _lib():
    global _bound
    L = lib.load()
    if not _bound:
        vp = ctypes.c_void_p
        L.abt_find_vis.argtypes = [ctypes.c_int, ctypes.POINTER(_Geo), ctypes.c_int64, vp, ctypes.c_double,
                                   vp, vp, vp, vp, vp]
        L.abt_find_vis_workspace.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
        L.abt_find_vis_workspace.restype = ctypes.c_size_t
        L.abt_find_vis_dev.argtypes = [ctypes.POINTER(_Batch), ctypes.c_int64, vp, ctypes.c_size_t, vp]
        i64 = ctypes.c_int64
        L.abt_vis_rotations.argtypes = [i64, vp, vp]
        L.abt_vis_rotate_dev.argtypes = [i64, vp, i64, vp, i64, vp, vp, i64, vp, vp, vp, vp, vp, vp]
        _bound = True
    return L


@dataclass
class VisGeometry:
    """The geometry_type fields (types_mod.f90:156-190) the visibility routines read.

    v (nv, 3); facets as CSR: f_off (nf+1), f_vi 1-based vertex ids; mid (nf, 3);
    nz (nf) = n(f%ni, 3); ap (nf) 1-based aperture ids; ap_nz (na) = ap(:)%n(3); optionally
    the full facet normals n (nf, 3) and aperture normals ap_n (na, 3), which rotations need."""
    v: np.ndarray
    f_off: np.ndarray
    f_vi: np.ndarray
    mid: np.ndarray
    nz: np.ndarray
    ap: np.ndarray
    ap_nz: np.ndarray
    n: np.ndarray = None
    ap_n: np.ndarray = None

    @property
    def nf(self):
        return len(self.f_off) - 1

    @property
    def nv(self):
        return len(self.v)

    @property
    def na(self):
        return len(self.ap_nz)

    @staticmethod
    def from_mesh(verts, faces, ap, ap_normals):
        """verts (nv, 3); faces: a list of 0-based vertex-index sequences (or an (nf, k)
        array); ap: 1-based aperture id per facet; ap_normals (na, 3).  Midpoints and
        normals as compute_geometry_midpoints / compute_geometry_normals
        (misc_submod.f90:199-232): the vertex sum over n, and the normalised cross product
        of the first two edges."""
        verts = np.ascontiguousarray(verts, dtype=np.float64)
        faces = [np.asarray(f, dtype=np.int64) for f in faces]
        off = np.zeros(len(faces) + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(f) for f in faces])
        vi = np.concatenate(faces) + 1
        return VisGeometry._derive(verts, off, vi, np.asarray(ap, dtype=np.int64),
                                   np.asarray(ap_normals, dtype=np.float64))

    @staticmethod
    def _derive(verts, off, vi, ap, ap_normals):
        nvs = np.diff(off)
        s = np.zeros((len(nvs), 3))
        for k in range(int(nvs.max())):          # vertex sums in vertex order
            has = k < nvs
            s[has] = s[has] + verts[vi[off[:-1][has] + k] - 1]
        mid = s / nvs[:, None]
        p0, p1, p2 = (verts[vi[off[:-1] + q] - 1] for q in range(3))
        c = np.cross(p1 - p0, p2 - p1)
        n = c / np.sqrt((c * c).sum(1))[:, None]
        return VisGeometry(verts, off, vi, mid, n[:, 2].copy(), ap, ap_normals[:, 2].copy(), n,
                           np.asarray(ap_normals, dtype=np.float64))

    def rotated(self, R):
        """The geometry seen in a frame rotated by R (vertices, normals, aperture normals,
        midpoints recomputed), as the tracer does before find_vis_int / find_vis_ext."""
        R = np.asarray(R, dtype=np.float64)
        return VisGeometry._derive(self.v @ R.T, self.f_off, self.f_vi, self.ap, self.ap_n @ R.T)


def _host_outputs(nf):
    return (np.zeros(nf, np.uint8), np.zeros(nf, np.uint8), np.zeros(nf), np.zeros(nf, np.int64),
            np.zeros(nf, np.uint8))


def find_vis(mode, geom: VisGeometry, field_in=None, beam_mid_z=0.0):
    """One call of find_vis_int / _ext / _inc on the GPU.  Returns a dict of numpy arrays
    in_beam, is_vis, dist_beam, id_beam, is_shad (booleans as bool)."""
    L = _lib()
    mode = MODES.get(mode, mode)
    keep = [np.ascontiguousarray(geom.v.T, dtype=np.float64), np.ascontiguousarray(geom.f_off, np.int64),
            np.ascontiguousarray(geom.f_vi, np.int64), np.ascontiguousarray(geom.mid, np.float64),
            np.ascontiguousarray(geom.nz, np.float64), np.ascontiguousarray(geom.ap, np.int64),
            np.ascontiguousarray(geom.ap_nz, np.float64)]
    g = _Geo(geom.nv, keep[0].ctypes.data, geom.nf, keep[1].ctypes.data, keep[2].ctypes.data,
             keep[3].ctypes.data, keep[4].ctypes.data, keep[5].ctypes.data, geom.na, keep[6].ctypes.data)
    fin = np.ascontiguousarray(field_in if field_in is not None else np.zeros(0), dtype=np.int64)
    inb, vis, dist, idb, shad = _host_outputs(geom.nf)
    lib._check(L.abt_find_vis(int(mode), ctypes.byref(g), len(fin), fin.ctypes.data if len(fin) else None,
                              float(beam_mid_z), inb.ctypes.data, vis.ctypes.data, dist.ctypes.data,
                              idb.ctypes.data, shad.ctypes.data))
    return dict(in_beam=inb.astype(bool), is_vis=vis.astype(bool), dist_beam=dist, id_beam=idb,
                is_shad=shad.astype(bool))


def find_vis_int(field_in, rot_geometry: VisGeometry):
    r = find_vis(ABT_VIS_INT, rot_geometry, field_in)
    return r["in_beam"], r["dist_beam"], r["id_beam"], r["is_shad"]


def find_vis_ext(field_in, rot_geometry: VisGeometry):
    r = find_vis(ABT_VIS_EXT, rot_geometry, field_in)
    return r["in_beam"], r["dist_beam"], r["id_beam"], r["is_shad"]


def find_vis_inc(geometry: VisGeometry, beam_mid_z):
    r = find_vis(ABT_VIS_INC, geometry, None, beam_mid_z)
    return r["in_beam"], r["is_vis"], r["dist_beam"], r["id_beam"], r["is_shad"]


class VisBatch:
    """Device tensors for many beams over one facet topology (abt_find_vis_dev).

    topo: a VisGeometry whose f_off, f_vi, ap give the shared topology; v (nbeam, 3, nv),
    mid (nbeam, nf, 3), nz (nbeam, nf), ap_nz (nbeam, na), in_off (nbeam+1) and field_in
    (1-based) are torch tensors on the device; beam_mid_z (nbeam) for mode "inc"."""

    def __init__(self, mode, topo: VisGeometry, v, mid, nz, ap_nz, in_off=None, field_in=None,
                 beam_mid_z=None):
        import torch
        self.mode = MODES.get(mode, mode)
        dev = v.device
        self.nbeam = v.shape[0]
        self.nv, self.nf, self.na = topo.nv, topo.nf, topo.na
        self.nvref = int(topo.f_off[-1])
        self.f_off = torch.from_numpy(np.ascontiguousarray(topo.f_off, np.int64)).to(dev)
        self.f_vi = torch.from_numpy(np.ascontiguousarray(topo.f_vi, np.int64)).to(dev)
        self.f_ap = torch.from_numpy(np.ascontiguousarray(topo.ap, np.int64)).to(dev)
        self.v, self.mid, self.nz, self.ap_nz = v.contiguous(), mid.contiguous(), nz.contiguous(), ap_nz.contiguous()
        if in_off is None:
            in_off = torch.zeros(self.nbeam + 1, dtype=torch.int64, device=dev)
        if field_in is None:
            field_in = torch.zeros(1, dtype=torch.int64, device=dev)
        self.in_off, self.field_in = in_off.contiguous(), field_in.contiguous()
        self.beam_mid_z = beam_mid_z
        n = self.nbeam * self.nf
        self.in_beam = torch.empty(n, dtype=torch.uint8, device=dev)
        self.is_vis = torch.empty(n, dtype=torch.uint8, device=dev)
        self.dist_beam = torch.empty(n, dtype=torch.float64, device=dev)
        self.id_beam = torch.empty(n, dtype=torch.int64, device=dev)
        self.is_shad = torch.empty(n, dtype=torch.uint8, device=dev)
        ws = _lib().abt_find_vis_workspace(self.nbeam, self.nf, self.nvref)
        self.work = torch.empty(ws // 8 + 64, dtype=torch.float64, device=dev)
        self.counts = torch.zeros(2, dtype=torch.int64, device=dev)

    def _struct(self, counts, phases):
        p = lambda t: None if t is None else t.data_ptr()  # noqa: E731
        return _Batch(self.mode, self.nbeam, self.nv, self.nf, self.na, p(self.f_off), p(self.f_vi),
                      p(self.f_ap), p(self.v), p(self.mid), p(self.nz), p(self.ap_nz), p(self.in_off),
                      p(self.field_in), p(self.beam_mid_z), p(self.in_beam), p(self.is_vis),
                      p(self.dist_beam), p(self.id_beam), p(self.is_shad),
                      p(self.counts) if counts else None, phases)

    def run(self, stream=None, counts=False, phases=0):
        """Launch on `stream` (default: torch's current stream).  counts=True adds the
        reference's work (candidate pairs at the depth test, blocker edges) to self.counts;
        phases 1 / 2 run only the set-up / only the blocker scan (0: both)."""
        b = self._struct(counts, phases)
        wp = self.work.data_ptr()
        wp_al = (wp + 255) // 256 * 256
        nbytes = self.work.numel() * 8 - (wp_al - wp)
        lib._check(_lib().abt_find_vis_dev(ctypes.byref(b), self.nvref, ctypes.c_void_p(wp_al), nbytes,
                                           lib._stream_ptr(stream)))

    def results(self):
        sh = (self.nbeam, self.nf)
        return dict(in_beam=self.in_beam.view(sh).bool(), is_vis=self.is_vis.view(sh).bool(),
                    dist_beam=self.dist_beam.view(sh), id_beam=self.id_beam.view(sh),
                    is_shad=self.is_shad.view(sh).bool())


def vis_rotations(props):
    """rotate()'s rotation (beam_loop_mod.f90:1488-1503) for each propagation direction:
    props (nbeam, 3) -> R (nbeam, 3, 3) taking each to -z (host, C library)."""
    p = np.ascontiguousarray(props, dtype=np.float64).reshape(-1, 3)
    R = np.empty((len(p), 3, 3))
    lib._check(_lib().abt_vis_rotations(len(p), p.ctypes.data, R.ctypes.data))
    return R


def rotate_dev(R, geom: VisGeometry, stream=None):
    """The particle geometry rotated into each beam's frame on the GPU, as rotate() does
    (beam_loop_mod.f90:1467-1526): R a (nbeam, 3, 3) device tensor; returns device tensors
    v (nbeam, 3, nv), mid (nbeam, nf, 3), nz (nbeam, nf), ap_nz (nbeam, na), the inputs of
    VisBatch."""
    import torch
    dev = R.device
    nb = R.shape[0]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)  # noqa: E731
    v0, mid0, n0, apn0 = t(geom.v.T), t(geom.mid), t(geom.n), t(geom.ap_n)
    v = torch.empty((nb, 3, geom.nv), dtype=torch.float64, device=dev)
    mid = torch.empty((nb, geom.nf, 3), dtype=torch.float64, device=dev)
    nz = torch.empty((nb, geom.nf), dtype=torch.float64, device=dev)
    apnz = torch.empty((nb, geom.na), dtype=torch.float64, device=dev)
    Rc = R.contiguous()
    lib._check(_lib().abt_vis_rotate_dev(nb, Rc.data_ptr(), geom.nv, v0.data_ptr(), geom.nf, mid0.data_ptr(),
                                         n0.data_ptr(), geom.na, apn0.data_ptr(), v.data_ptr(), mid.data_ptr(),
                                         nz.data_ptr(), apnz.data_ptr(), lib._stream_ptr(stream)))
    return v, mid, nz, apnz
