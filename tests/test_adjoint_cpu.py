"""The adjoint bodies (abt_amd/csrc/adjbody.h) on the CPU, pinned by finite differences.

The loss is L = Re sum_bins conj(Lam) A(v, C) for random Lam, with A the FP64
far-field amplitude of one triangular beam (the kernel's FP64 body with the
cancellation-free far-field phase, run through the CPU emulator).  Central
differences with the 4th-order stencil, step 1e-4 um for vertices: the adjoint's
vertex and amplitude gradients agree within 1e-7 of the largest component.  The
mixed-precision body agrees with the FP64 body within 2e-5 (norm-relative).
"""
import numpy as np
import pytest

from abt_amd import lib
from abt_amd.beams import BeamSet
from abt_amd.workloads import grid, hex_column, random_rotations, to_beamset, trace_orientations
from tests import emu

LAM = 0.532


def _beams():
    v, f = hex_column(10.0, 20.0, 8, 14, 0.5, seed=0)
    t = trace_orientations(v, f, random_rotations(1, 3), rec=3)
    b = to_beamset(t["beams"])
    e = to_beamset(t["ext"])
    return b.subset(np.arange(0, b.n, max(1, b.n // 5))[:5]), e.subset(np.arange(2))


def _lam_prime(Lam, th, ph):
    cp = np.tile(np.cos(ph), len(th))
    sp = np.tile(np.sin(ph), len(th))
    T = np.zeros((len(cp), 2, 2))
    T[:, 0, 0] = -sp; T[:, 0, 1] = -cp; T[:, 1, 0] = cp; T[:, 1, 1] = -sp
    Lp = np.einsum("brc,bdc->brd", Lam, T)
    return np.stack([Lp[:, 0, 0].real, Lp[:, 0, 0].imag, Lp[:, 1, 0].real, Lp[:, 1, 0].imag,
                     Lp[:, 0, 1].real, Lp[:, 0, 1].imag, Lp[:, 1, 1].real, Lp[:, 1, 1].imag], 1)


def _loss(bs, th, ph, Lam):
    recs, _ = lib.prepare_beams_host(bs, LAM)
    A = emu.far_field(3, th, ph, LAM, recs, bs.n)        # [r, c, theta, phi]
    A = A.transpose(2, 3, 0, 1).reshape(-1, 2, 2)
    return float((Lam.conj() * A).real.sum())


@pytest.mark.parametrize("gspec", [((0, 5, 180), (0, 10, 350)), ((0, 0.5, 3, 1, 20), (0, 30, 330))])
def test_adjoint_body_matches_finite_differences(gspec):
    th, ph = grid(*gspec)
    B, E = _beams()
    rng = np.random.default_rng(11)
    nbins = len(th) * len(ph)
    bins = emu.adjoint_bins(th, ph, LAM)
    for bs in (B, E):
        for b in range(bs.n):
            one = bs.subset([b])
            Lam = rng.normal(size=(nbins, 2, 2)) + 1j * rng.normal(size=(nbins, 2, 2))
            recs, _ = lib.prepare_beams_host(one, LAM)
            gv, ga = emu.adjoint(recs, bins, _lam_prime(Lam, th, ph), LAM)
            h = 1e-4
            fd = np.zeros((3, 3))
            for l in range(3):
                for a in range(3):
                    f = {}
                    for m in (-2, -1, 1, 2):
                        p = BeamSet(one.ampl, one.vk7, one.prop, one.fov, one.vert_off, one.verts.copy())
                        p.verts[l, a] += m * h
                        f[m] = _loss(p, th, ph, Lam)
                    fd[l, a] = (-f[2] + 8 * f[1] - 8 * f[-1] + f[-2]) / (12 * h)
            assert np.abs(gv[0] - fd).max() <= 1e-7 * np.abs(fd).max(), (b, gv[0], fd)
            # amplitude gradient: L is linear in ampl, so differences are exact up to rounding
            fa = np.zeros(8)
            for q in range(8):
                f = []
                for m in (-1, 1):
                    p = BeamSet(one.ampl.copy(), one.vk7, one.prop, one.fov, one.vert_off, one.verts)
                    p.ampl[0, q] += m * 1e-3
                    f.append(_loss(p, th, ph, Lam))
                fa[q] = (f[1] - f[0]) / 2e-3
            assert np.abs(ga[0] - fa).max() <= 1e-8 * np.abs(fa).max(), (b, ga[0], fa)


def test_mixed_adjoint_body_matches_fp64_body():
    th, ph = grid((0, 0.5, 180), (0, 4, 356))
    B, _ = _beams()
    rng = np.random.default_rng(12)
    nbins = len(th) * len(ph)
    Lam = rng.normal(size=(nbins, 2, 2)) + 1j * rng.normal(size=(nbins, 2, 2))
    lp = _lam_prime(Lam, th, ph)
    recs, _ = lib.prepare_beams_host(B, LAM)
    bins = emu.adjoint_bins(th, ph, LAM)
    gv, ga = emu.adjoint(recs, bins, lp, LAM)
    gvm, gam = emu.adjoint(recs, bins, lp, LAM, mixed=True)
    assert np.linalg.norm(gvm - gv) / np.linalg.norm(gv) <= 2e-5
    assert np.linalg.norm(gam - ga) / np.linalg.norm(ga) <= 2e-5


def test_model_forward_matches_oracle():
    """The FP64 model the adjoint differentiates (frame-free, cancellation-free phase)
    equals the reference's diffraction (the oracle) to 5e-9 of the largest bin, off the
    exact forward directions where the reference formula itself loses accuracy.  The
    bound is the reference's own rounding: its phase k|x - c| ~ 1.2e7 rad carries
    ~1e-9 rad per evaluation."""
    import oracle as O
    from tests.test_gpu_parity import _forward_mask
    th, ph = grid((0, 5, 180), (0, 10, 350))
    B, E = _beams()
    for bs in (B, E):
        recs, _ = lib.prepare_beams_host(bs, LAM)
        got = emu.far_field(3, th, ph, LAM, recs, bs.n)
        want, _ = O.diff_main(th, ph, LAM, bs, BeamSet.empty())
        err = np.abs(got - want).max(axis=(0, 1)) / np.abs(want).max()
        fwd = _forward_mask(th, ph, [bs])
        assert err[~fwd].max() <= 5e-9, err[~fwd].max()


def test_aperture_vertices_gather_and_gradient():
    """The index_select gather of aperture vertices equals advanced indexing, values and
    gradient (torch on the CPU)."""
    import torch
    from abt_amd import autograd as AG
    rng = np.random.default_rng(3)
    V = torch.tensor(rng.normal(size=(12, 3)), requires_grad=True)
    faces = torch.tensor(rng.integers(0, 12, size=(7, 3)))
    fi = torch.tensor(rng.integers(1, 8, size=40))
    R = torch.tensor(rng.normal(size=(4, 3, 3)))
    o = torch.tensor(rng.integers(0, 4, size=40))
    a = AG.aperture_vertices(V, faces, fi, R, o)
    b = torch.einsum("bij,bvj->bvi", R[o], V[faces[fi - 1]])
    np.testing.assert_allclose(a.detach().numpy(), b.detach().numpy(), rtol=1e-14, atol=1e-14)
    G = torch.tensor(rng.normal(size=tuple(a.shape)))
    ga, = torch.autograd.grad((a * G).sum(), V)
    gb, = torch.autograd.grad((b * G).sum(), V)
    np.testing.assert_allclose(ga.numpy(), gb.numpy(), rtol=1e-12, atol=1e-12)
