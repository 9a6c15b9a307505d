"""PyTorch harness: the orientation-batched far field as a differentiable function of
every aperture's vertices and amplitude matrix (shape and roughness retrieval).

`FarFieldSystem` holds what stays fixed (grid, per-orientation beam offsets, each
beam's propagation vector, e-perp vector and field of view, the batched forward plan
and the adjoint plan).  `far_field(system, verts_b, ampl_b, verts_e, ampl_e)` returns
the per-orientation total amplitudes X_o = k (ampl_far_beam + ampl_far_ext_diff),
shape (norient, 8, nbins) in the Fortran pair layout; its backward runs the hand-written
adjoint kernel (abt_adjoint_run).  Losses of the orientation-averaged Mueller matrix
are plain torch code on X (`mueller_11_12`, `retrieval_loss`).
"""
from __future__ import annotations

import numpy as np
import torch

from . import lib


class FarFieldSystem:
    """beams / ext: dicts of device tensors with at least prop (n,3), vk7 (n,3) and
    optionally fov (n,); beam_off / ext_off: norient+1 offsets (host)."""

    def __init__(self, grid: lib.DeviceGrid, beam_off, ext_off, beams: dict, ext: dict,
                 precision="mixed", adjoint_precision=None):
        self.grid = grid
        self.lam = grid.lam
        self.beam_off = np.asarray(beam_off, dtype=np.int64)
        self.ext_off = np.asarray(ext_off, dtype=np.int64)
        self.norient = len(self.beam_off) - 1
        self.nb = int(self.beam_off[-1])
        self.ne = int(self.ext_off[-1])
        self.beams, self.ext = beams, ext
        self.device = beams["prop"].device
        self.precision = precision
        self.fwd = lib.BatchPlan(grid, self.beam_off, self.ext_off, precision)
        adj = adjoint_precision or ("fp64" if precision.startswith("fp64") else "mixed")
        self.adj = lib.AdjointPlan(grid, self.beam_off, self.ext_off, adj)
        R = lib.load().abt_record_doubles()
        n = grid.nbins
        f64 = dict(dtype=torch.float64, device=self.device)
        self.rec_b = torch.empty((max(self.nb, 1), R), **f64)
        self.rec_e = torch.empty((max(self.ne, 1), R), **f64)
        self.mueller = torch.empty(16 * n, **f64)
        self.integrals = torch.empty(6 * self.norient, **f64)
        tri = lambda m: torch.arange(m + 1, dtype=torch.int64, device=self.device) * 3
        self.voff_b, self.voff_e = tri(self.nb), tri(self.ne)
        # optional HIP-event timing of the library calls (forward kernel, adjoint kernel)
        self.timing = False
        self.events = {}

    def _event_pair(self, name):
        if not self.timing:
            return None, None
        e = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        self.events[name] = e
        return e

    def kernel_ms(self, name):
        e = self.events.get(name)
        return e[0].elapsed_time(e[1]) if e else None

    def _prepare(self, verts, ampl, tree):
        d = self.beams if tree == 0 else self.ext
        n = self.nb if tree == 0 else self.ne
        if n == 0:
            return
        dev = {"ampl": ampl.detach().contiguous(), "vk7": d["vk7"], "prop": d["prop"],
               "vert_off": self.voff_b if tree == 0 else self.voff_e,
               "verts": verts.detach().reshape(-1, 3).contiguous()}
        per_beam_fov = tree == 0 and "fov" in d
        if per_beam_fov:
            dev["fov"] = d["fov"]
        out = self.rec_b if tree == 0 else self.rec_e
        # ext-diff beams use fov = pi/2 (diff_mod.f90:797); fov only matters with fast_diff
        lib.prepare_beams_dev(dev, n, self.lam, out, fov_value=None if per_beam_fov else lib.PI_R4 / 2)

    def forward_amplitudes(self, verts_b, ampl_b, verts_e, ampl_e):
        """X (norient, 8, nbins) for the given vertices and amplitudes (no autograd)."""
        self._prepare(verts_b, ampl_b, 0)
        self._prepare(verts_e, ampl_e, 1)
        X = torch.empty((self.norient, 8, self.grid.nbins), dtype=torch.float64, device=self.device)
        e0, e1 = self._event_pair("forward")
        if e0 is not None:
            e0.record()
        self.fwd.run(self.rec_b if self.nb else None, self.rec_e if self.ne else None, self.mueller,
                     self.integrals, ampl=X)
        if e1 is not None:
            e1.record()
        return X


class _FarField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, system, verts_b, ampl_b, verts_e, ampl_e):
        X = system.forward_amplitudes(verts_b, ampl_b, verts_e, ampl_e)
        ctx.system = system
        # the records of this forward: the adjoint needs the same frames
        ctx.save_for_backward(system.rec_b.clone(), system.rec_e.clone())
        return X

    @staticmethod
    def backward(ctx, gX):
        s = ctx.system
        rb, re = ctx.saved_tensors
        f64 = dict(dtype=torch.float64, device=gX.device)
        gvb = torch.empty((s.nb, 3, 3), **f64)
        gab = torch.empty((s.nb, 8), **f64)
        gve = torch.empty((s.ne, 3, 3), **f64)
        gae = torch.empty((s.ne, 8), **f64)
        g = gX.contiguous()
        e0, e1 = s._event_pair("adjoint")
        if e0 is not None:
            e0.record()
        s.adj.run(rb if s.nb else None, re if s.ne else None, g,
                  gvb if s.nb else None, gab if s.nb else None, gve if s.ne else None,
                  gae if s.ne else None)
        if e1 is not None:
            e1.record()
        return None, gvb, gab, gve, gae


def far_field(system: FarFieldSystem, verts_b, ampl_b, verts_e, ampl_e):
    """Differentiable per-orientation total amplitudes X (norient, 8, nbins)."""
    return _FarField.apply(system, verts_b, ampl_b, verts_e, ampl_e)


class _GatherRows(torch.autograd.Function):
    """V.index_select(0, vid) with a backward that sums each row of V's gradient from a
    padded table of the rows that gathered it: one gather and a dense sum, no atomics and
    no sort per call, so the result does not depend on the order of the additions' threads."""

    @staticmethod
    def forward(ctx, V, vid, table):
        ctx.save_for_backward(table)
        return V.index_select(0, vid)

    @staticmethod
    def backward(ctx, g):
        table, = ctx.saved_tensors
        gz = torch.cat([g, g.new_zeros((1,) + tuple(g.shape[1:]))])
        return gz[table].sum(1), None, None


class ApertureGather:
    """Aperture vertices of every beam, R_o V[faces[fi - 1]], shape (n, 3, 3), as a
    differentiable function of the particle vertices V (nv, 3).  faces: (nfaces, 3) vertex
    indices; fi: 1-based facet id per beam (the reference's outbeamtype%fi, types_mod.f90:8-17,
    whose vertices diff_mod.f90:529-531 looks up); rot: (norient, 3, 3); orient: orientation
    index per beam.  The facet ids are fixed at construction, where the backward's row table
    (rows of each vertex, padded with a zero row) is built once."""

    def __init__(self, faces, fi, nv):
        vid = faces[fi - 1].reshape(-1)
        n = vid.numel()
        order = torch.argsort(vid, stable=True)
        counts = torch.bincount(vid, minlength=nv)
        start = torch.cumsum(counts, 0) - counts
        sv = vid[order]
        pos = torch.arange(n, device=vid.device) - start[sv]
        width = max(int(counts.max()) if n else 0, 1)
        table = torch.full((nv, width), n, dtype=torch.long, device=vid.device)
        table[sv, pos] = order
        self.vid, self.table = vid, table

    def __call__(self, V, rot, orient):
        P = _GatherRows.apply(V, self.vid, self.table).reshape(-1, 3, 3)
        return torch.einsum("bij,bvj->bvi", rot[orient], P)


def aperture_vertices(V, faces, fi, rot, orient):
    """One-off form of ApertureGather (builds the row table on every call)."""
    return ApertureGather(faces, fi, V.shape[0])(V, rot, orient)


def mueller_11_12(X):
    """Orientation-averaged S11 and S12 per bin (ampl_to_mueller, outputs_mod.f90:489-543,
    then the mean over orientations, divide_by_num_orientations outputs_mod.f90:13-43)."""
    a = X * X
    s11 = a[:, 0] + a[:, 1]     # |S(1,1)|^2
    s21 = a[:, 2] + a[:, 3]
    s12 = a[:, 4] + a[:, 5]
    s22 = a[:, 6] + a[:, 7]
    m11 = 0.5 * (s11 + s12 + s21 + s22)
    m12 = 0.5 * (s11 - s12 + s21 - s22)
    return m11.mean(0), m12.mean(0)


def retrieval_loss(X, target11, target_pol, weights):
    """Weighted L2 between the orientation-averaged S11 (relative to the target) and
    -S12/S11 and a target pattern:
        L = sum_bins w [((S11 - T11) / T11)^2 + (-S12/S11 - Tpol)^2]."""
    m11, m12 = mueller_11_12(X)
    pol = -m12 / m11
    return (weights * (((m11 - target11) / target11) ** 2 + (pol - target_pol) ** 2)).sum()


def solid_angle_weights(theta, phi, device):
    """sin(theta) dtheta dphi per bin (phi fastest), normalised to sum 1."""
    th = np.asarray(theta)
    ph = np.asarray(phi)
    dth = np.gradient(th) if len(th) > 1 else np.ones(1)
    dph = np.gradient(ph) if len(ph) > 1 else np.ones(1)
    w = (np.sin(th) * dth)[:, None] * dph[None, :]
    w = np.maximum(w, 0.0)
    w /= w.sum()
    return torch.from_numpy(w.reshape(-1)).to(device)
