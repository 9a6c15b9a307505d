"""Beam-set container: the flattened form of PBT's outgoing-beam trees.

A beam set holds what `diffraction` (diff_mod.f90:446-672) reads from each
`outbeamtype` record (types_mod.f90:8-17) plus the facet vertices it looks up
through `fi` (diff_mod.f90:529-531):

    ampl      (nb, 8)  float64  re/im of ampl(1,1), ampl(2,1), ampl(1,2), ampl(2,2)
                                (Fortran column-major order of complex(8) ampl(2,2))
    vk7       (nb, 3)  float64  e-perp vector
    prop      (nb, 3)  float64  outgoing propagation vector (prop_out)
    fov       (nb,)    float64  field of view (only read when fast_diff is on)
    vert_off  (nb+1,)  int64    offsets into verts, in vertices
    verts     (nv, 3)  float64  facet vertices, in the facet's own order

The same layout crosses the C-ABI (include/abt_diff.h, struct abt_beams).
Two file forms keep beam sets reproducible across hosts: `save`/`load` (numpy .npz,
no pickle) and `to_json`/`from_json`, a text form whose records carry the
`outbeamtype` field names (ampl, vk7, prop_out, fov, verts) with the amplitude
matrix written row by row as [[[re, im], [re, im]], [[re, im], [re, im]]], the
order of the reference's own beam export (beam_loop_mod.f90:254-263).  Floats
are written with Python's shortest round-trip repr, so a JSON round trip is exact.

`outbeams_from_pbt_json` reads the reference's own export, the near-field beam tree that
`-export_beam` writes (beam.json, beam_loop_mod.f90:136-340), and forms the outgoing beams
the way add_to_outbeam_tree does (beam_loop_mod.f90:2745-2790).
"""
from __future__ import annotations

import dataclasses
import json

import numpy as np

PI_R4 = float(np.float32(3.14159265358979))


@dataclasses.dataclass
class BeamSet:
    ampl: np.ndarray
    vk7: np.ndarray
    prop: np.ndarray
    fov: np.ndarray
    vert_off: np.ndarray
    verts: np.ndarray

    def __post_init__(self):
        self.ampl = np.ascontiguousarray(self.ampl, dtype=np.float64).reshape(-1, 8)
        nb = self.ampl.shape[0]
        self.vk7 = np.ascontiguousarray(self.vk7, dtype=np.float64).reshape(nb, 3)
        self.prop = np.ascontiguousarray(self.prop, dtype=np.float64).reshape(nb, 3)
        self.fov = np.ascontiguousarray(self.fov, dtype=np.float64).reshape(nb)
        self.vert_off = np.ascontiguousarray(self.vert_off, dtype=np.int64).reshape(nb + 1)
        self.verts = np.ascontiguousarray(self.verts, dtype=np.float64).reshape(-1, 3)
        if self.vert_off[0] != 0 or self.vert_off[-1] != self.verts.shape[0]:
            raise ValueError("vert_off must start at 0 and end at len(verts)")
        if nb and np.any(np.diff(self.vert_off) < 3):
            raise ValueError("every aperture needs at least 3 vertices")

    @property
    def n(self) -> int:
        return self.ampl.shape[0]

    @property
    def nv(self) -> np.ndarray:
        return np.diff(self.vert_off)

    @staticmethod
    def empty() -> "BeamSet":
        return BeamSet(np.zeros((0, 8)), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0),
                       np.zeros(1, np.int64), np.zeros((0, 3)))

    @staticmethod
    def from_triangles(ampl, vk7, prop, fov, tri_verts) -> "BeamSet":
        """tri_verts: (nb, 3, 3) vertex coordinates of triangular apertures."""
        tri_verts = np.asarray(tri_verts, dtype=np.float64)
        nb = tri_verts.shape[0]
        return BeamSet(ampl, vk7, prop, fov, np.arange(nb + 1, dtype=np.int64) * 3,
                       tri_verts.reshape(-1, 3))

    def ampl_complex(self) -> np.ndarray:
        """(nb, 2, 2) complex view in row/column order ampl[b, r, c]."""
        a = self.ampl.reshape(-1, 2, 2, 2)  # [b, c, r, re/im]
        return (a[..., 0] + 1j * a[..., 1]).transpose(0, 2, 1)

    def subset(self, idx) -> "BeamSet":
        idx = np.asarray(idx, dtype=np.int64)
        nv = self.nv[idx]
        off = np.zeros(len(idx) + 1, np.int64)
        np.cumsum(nv, out=off[1:])
        verts = np.concatenate([self.verts[self.vert_off[i]:self.vert_off[i + 1]] for i in idx]) \
            if len(idx) else np.zeros((0, 3))
        return BeamSet(self.ampl[idx], self.vk7[idx], self.prop[idx], self.fov[idx], off, verts)

    def save(self, path):
        np.savez_compressed(path, ampl=self.ampl, vk7=self.vk7, prop=self.prop, fov=self.fov,
                            vert_off=self.vert_off, verts=self.verts)

    @staticmethod
    def load(path) -> "BeamSet":
        with np.load(path, allow_pickle=False) as z:
            return BeamSet(z["ampl"], z["vk7"], z["prop"], z["fov"], z["vert_off"], z["verts"])

    def to_json(self, path):
        a = self.ampl
        beams = []
        for b in range(self.n):
            v = self.verts[self.vert_off[b]:self.vert_off[b + 1]]
            # ampl(r, c) at pair index r + 2c of the Fortran column-major order
            rows = [[[float(a[b, 2 * (r + 2 * c)]), float(a[b, 2 * (r + 2 * c) + 1])] for c in range(2)]
                    for r in range(2)]
            beams.append({"ampl": rows, "vk7": [float(x) for x in self.vk7[b]],
                          "prop_out": [float(x) for x in self.prop[b]], "fov": float(self.fov[b]),
                          "verts": [[float(x) for x in p] for p in v]})
        with open(path, "w") as f:
            json.dump({"outbeams": beams}, f)

    @staticmethod
    def from_json(path) -> "BeamSet":
        with open(path) as f:
            beams = json.load(f)["outbeams"]
        n = len(beams)
        ampl = np.zeros((n, 8))
        for b, rec in enumerate(beams):
            for r in range(2):
                for c in range(2):
                    ampl[b, 2 * (r + 2 * c):2 * (r + 2 * c) + 2] = rec["ampl"][r][c]
        nv = [len(rec["verts"]) for rec in beams]
        off = np.zeros(n + 1, np.int64)
        np.cumsum(nv, out=off[1:])
        verts = np.array([p for rec in beams for p in rec["verts"]], dtype=np.float64).reshape(-1, 3)
        return BeamSet(ampl, [rec["vk7"] for rec in beams] if n else np.zeros((0, 3)),
                       [rec["prop_out"] for rec in beams] if n else np.zeros((0, 3)),
                       [rec["fov"] for rec in beams], off, verts)


def concat(sets) -> BeamSet:
    sets = list(sets)
    if not sets:
        return BeamSet.empty()
    off = [np.zeros(1, np.int64)]
    base = 0
    for s in sets:
        off.append(s.vert_off[1:] + base)
        base += s.verts.shape[0]
    return BeamSet(np.concatenate([s.ampl for s in sets]), np.concatenate([s.vk7 for s in sets]),
                   np.concatenate([s.prop for s in sets]), np.concatenate([s.fov for s in sets]),
                   np.concatenate(off), np.concatenate([s.verts for s in sets]))


def outbeams_from_pbt_json(path, verts, faces, lam, with_ids=False):
    """The outgoing beams of a PBT beam.json (`-export_beam`, beam_loop_mod.f90:136-340).

    The file holds the near-field beam tree: per beam its propagation vector `prop`,
    `is_int`, and `field_in` (per illuminating facet: `fi`, the amplitude matrix row by
    row, `e_perp`).  add_to_outbeam_tree (beam_loop_mod.f90:2745-2790) makes one outgoing
    beam per field_in facet of an external beam that reaches the far field:
        ampl = field_in%ampl, prop_out = beam%prop, vk7 = field_in%e_perp,
        aperture = the facet fi (vertices from the particle geometry),
        fov = 8 lambda / d, d = 2 sqrt(A / pi), A the summed area of the beam's field_in facets.
    The file does not record which facets reach the far field (the `is_outgoing` flag that
    recursion_ext clears when an external beam meets the particle again,
    beam_loop_mod.f90:884), so every field_in facet of every external beam (is_int false) is
    taken: the outgoing set of a particle whose external beams all escape (convex
    particles), a superset otherwise.  Internal beams are skipped.

    verts (nv, 3) and faces (nf, k) 0-based: the particle geometry the tree was traced on
    (geometry%v and geometry%f(:)%vi), facet fi of the file is faces[fi - 1].  Returns a
    BeamSet (and, with_ids, the 1-based facet ids and the file's beam ids)."""
    with open(path) as f:
        doc = json.load(f)
    tree = doc["beam tree"]
    V = np.asarray(verts, dtype=np.float64)
    F = [np.asarray(fc, dtype=np.int64) for fc in faces]

    def area(fi):
        p = V[F[fi - 1]]
        n = np.zeros(3)
        for j in range(len(p)):   # Newell's formula, any planar polygon
            n += np.cross(p[j], p[(j + 1) % len(p)])
        return 0.5 * np.linalg.norm(n)

    ampl, vk7, prop, fov, vo, vs, fis, ids = [], [], [], [], [0], [], [], []
    for beam in tree:
        if beam["is_int"]:
            continue
        fields = beam["field_in"]
        if not fields:
            continue
        a_tot = 0.0
        for fd in fields:
            a_tot = a_tot + area(int(fd["fi"]))
        fv = (lam / (2.0 * np.sqrt(a_tot / PI_R4))) * 8.0     # the REAL*4 pi of misc_submod.f90:12
        for fd in fields:
            fi = int(fd["fi"])
            rows = fd["ampl"]          # [[a11, a12], [a21, a22]], each [re, im]
            a = np.empty(8)
            for r in range(2):
                for c in range(2):
                    a[2 * (r + 2 * c):2 * (r + 2 * c) + 2] = rows[r][c]
            ampl.append(a)
            vk7.append(fd["e_perp"])
            prop.append(beam["prop"])
            fov.append(fv)
            p = V[F[fi - 1]]
            vs.append(p)
            vo.append(vo[-1] + len(p))
            fis.append(fi)
            ids.append(int(beam["beam id"]))
    if not ampl:
        out = BeamSet.empty()
    else:
        out = BeamSet(np.array(ampl), np.array(vk7, dtype=np.float64), np.array(prop, dtype=np.float64),
                      np.array(fov), np.array(vo, dtype=np.int64), np.concatenate(vs))
    if with_ids:
        return out, np.array(fis, dtype=np.int64), np.array(ids, dtype=np.int64)
    return out
