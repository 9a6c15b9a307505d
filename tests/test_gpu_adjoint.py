"""Adjoint (abt_adjoint_*) and its torch.autograd harness.

Checks
  * the GPU adjoint kernel against the same bodies run on the CPU (tests/emu), to 1e-10;
  * gradients of the retrieval loss (weighted L2 of the orientation-averaged S11 and
    -S12/S11 against a target from the particle with perturbed vertices) with respect to
    particle vertices against central finite differences, step 1e-4 um, on 50 random
    vertices: the 4th-order stencil within 1e-6 of the largest gradient component.  The
    forward used for the differences is the FP64 model (ABT_FP64_ACCURATE: the
    reference's formula in rotation-invariant form with the cancellation-free far-field
    phase, smooth to ~1e-15; the reference's direct k|x - c| carries ~1e-9 rad of rounding
    per bin, which a 1e-4 um step cannot resolve).  Vertices are drawn among those whose
    apertures keep every bin away from the singular direction of the reference's
    polarisation (see _smooth_vertices);
  * amplitude-matrix gradients against finite differences;
  * the mixed-precision adjoint against the FP64 adjoint (gradient vectors within 1e-5
    in norm, relative);
  * the full adjoint workload (64 orientations, recursion depth 10): both adjoints against
    directional finite differences on 50 vertices.
"""
import numpy as np
import pytest

from abt_amd import autograd as AG
from abt_amd import lib
from abt_amd.workloads import (GRIDS, grid, hex_column, random_rotations, random_triangles,
                               trace_orientations)
from tests import emu

pytestmark = pytest.mark.gpu
LAM = 0.532


def _particle_case(norient, seed=0, rec=4, gspec=GRIDS["adjoint"], sigma=0.5):
    import torch
    dev = torch.device("cuda:0")
    verts, faces = hex_column(10.0, 20.0, 8, 14, sigma, seed=seed)
    rots = random_rotations(norient, seed + 1)
    t = trace_orientations(verts, faces, rots, rec=rec, device=dev)
    th, ph = grid(*gspec)
    g = lib.DeviceGrid(th, ph, LAM)
    return dict(verts=verts, faces=faces, rots=rots, t=t, th=th, ph=ph, g=g, dev=dev)


def _system(case, precision):
    t = case["t"]
    beams = {k: t["beams"][k] for k in ("prop", "vk7")}
    ext = {k: t["ext"][k] for k in ("prop", "vk7")}
    return AG.FarFieldSystem(case["g"], t["beam_off"].cpu().numpy(), t["ext_off"].cpu().numpy(),
                             beams, ext, precision=precision)


def _aperture_verts(V, case, tree):
    """Lab-frame aperture vertices of every beam of one tree from particle vertices V."""
    import torch
    t = case["t"][tree]
    faces = torch.as_tensor(case["faces"], device=V.device)
    R = torch.as_tensor(case["rots"], dtype=torch.float64, device=V.device)
    fv = V[faces[t["fi"] - 1]]                   # (n, 3, 3)
    return torch.einsum("bij,bvj->bvi", R[t["orient"]], fv)


def test_adjoint_kernel_matches_cpu_bodies():
    import torch
    case = _particle_case(3, rec=2, gspec=((0, 5, 180), (0, 10, 350)))
    s = _system(case, "fp64")
    t = case["t"]
    X = s.forward_amplitudes(t["beams"]["verts"].reshape(-1, 3, 3), t["beams"]["ampl"],
                             t["ext"]["verts"].reshape(-1, 3, 3), t["ext"]["ampl"])
    rng = np.random.default_rng(0)
    G = torch.from_numpy(rng.normal(size=tuple(X.shape))).to(case["dev"])
    nb, ne = s.nb, s.ne
    gvb = torch.empty((nb, 9), dtype=torch.float64, device=case["dev"])
    gab = torch.empty((nb, 8), dtype=torch.float64, device=case["dev"])
    gve = torch.empty((ne, 9), dtype=torch.float64, device=case["dev"])
    gae = torch.empty((ne, 8), dtype=torch.float64, device=case["dev"])
    s.adj.run(s.rec_b, s.rec_e, G, gvb, gab, gve, gae)
    torch.cuda.synchronize()
    # the same bodies on the CPU, orientation by orientation
    bins = emu.adjoint_bins(case["th"], case["ph"], LAM)
    nph = len(case["ph"])
    cp = np.tile(np.cos(case["ph"]), len(case["th"]))
    sp = np.tile(np.sin(case["ph"]), len(case["th"]))
    k = 2 * lib.PI_R4 / LAM
    Gh = G.cpu().numpy()
    for tree, (recs, off, gv, ga) in enumerate([(s.rec_b, s.beam_off, gvb, gab), (s.rec_e, s.ext_off, gve, gae)]):
        rh = recs.cpu().numpy()
        for o in range(s.norient):
            b0, b1 = off[o], off[o + 1]
            if b1 == b0:
                continue
            L = k * Gh[o]                      # [8][nbins]
            lp = np.empty((L.shape[1], 8))
            for r in range(2):
                for ri in range(2):
                    l1, l2 = L[2 * r + ri], L[2 * (r + 2) + ri]
                    lp[:, 2 * r + ri] = -sp * l1 - cp * l2
                    lp[:, 2 * (r + 2) + ri] = cp * l1 - sp * l2
            wv, wa = emu.adjoint(rh[b0:b1], bins, lp, LAM)
            # same bodies; the device build contracts a*b+c into FMAs, the host build does
            # not, and gradient sums over bins cancel to ~1e-8 of their largest terms
            np.testing.assert_allclose(gv[b0:b1].cpu().numpy(), wv.reshape(-1, 9), rtol=1e-8,
                                       atol=1e-8 * np.abs(wv).max())
            np.testing.assert_allclose(ga[b0:b1].cpu().numpy(), wa, rtol=1e-8,
                                       atol=1e-8 * np.abs(wa).max())


def _retrieval_setup(norient, precision, rec=4):
    import torch
    case = _particle_case(norient, rec=rec)
    s = _system(case, precision)
    # target: the particle with vertices perturbed by N(0, 0.05 um), traced again
    rng = np.random.default_rng(77)
    vt = case["verts"] + rng.normal(scale=0.05, size=case["verts"].shape)
    tt = trace_orientations(vt, case["faces"], case["rots"], rec=rec, device=case["dev"])
    st = AG.FarFieldSystem(case["g"], tt["beam_off"].cpu().numpy(), tt["ext_off"].cpu().numpy(),
                           {k: tt["beams"][k] for k in ("prop", "vk7")},
                           {k: tt["ext"][k] for k in ("prop", "vk7")}, precision="fp64_accurate")
    Xt = st.forward_amplitudes(tt["beams"]["verts"].reshape(-1, 3, 3), tt["beams"]["ampl"],
                               tt["ext"]["verts"].reshape(-1, 3, 3), tt["ext"]["ampl"])
    t11, t12 = AG.mueller_11_12(Xt)
    w = AG.solid_angle_weights(case["th"], case["ph"], case["dev"])
    return case, s, t11, -t12 / t11, w


def _loss_of(case, s, V, ab, ae, T11, Tp, w):
    import torch
    vb = _aperture_verts(V, case, "beams")
    ve = _aperture_verts(V, case, "ext")
    X = AG.far_field(s, vb, ab, ve, ae)
    return AG.retrieval_loss(X, T11, Tp, w)


def _smooth_vertices(case, s, threshold=1e-5):
    """Particle vertices whose apertures have every bin direction at 1 - |khat.e_y| >
    threshold (e_y: the aperture frame's in-plane axis, the record's second frame row)."""
    v = _vertex_margins(case, s)
    return np.nonzero(np.isfinite(v) & (v > threshold))[0]


def _vertex_margins(case, s):
    """Per particle vertex, the smallest 1 - |khat.e_y| over the bins and the apertures
    that use it (inf for vertices no aperture uses).
    The reference's polarisation (karczewski, diff_mod.f90:145-181) is singular at
    khat = +-e_y, where m = (e_y - k_y khat) / sqrt(1 - k_y^2) has no limit, and is
    clamped at |k_y| = 0.999999 (diff_mod.f90:151), so within ~0.3 degrees of that axis
    the loss is not smooth on the 1e-4 um scale of the differences and they cannot
    check the gradient there."""
    import torch
    x, y, z = lib.make_far_field_bins(case["th"], case["ph"])
    X = torch.from_numpy(np.stack([x, y, z], 0)).to(case["dev"])
    X = X / X.norm(dim=0, keepdim=True)
    faces = torch.as_tensor(case["faces"], device=case["dev"])
    vmin = torch.full((len(case["verts"]),), float("inf"), dtype=torch.float64, device=case["dev"])
    for tree, rec in (("beams", s.rec_b), ("ext", s.rec_e)):
        fi = case["t"][tree]["fi"] - 1
        n = fi.shape[0]
        ey = rec[:n, 6:9]
        m = torch.cat([(ey[i:i + 2048] @ X).abs().max(1).values for i in range(0, n, 2048)])
        d = (1 - m)[:, None].expand(n, 3).reshape(-1)
        vmin.scatter_reduce_(0, faces[fi].reshape(-1), d, reduce="amin")
    return vmin.cpu().numpy()


def test_vertex_gradient_matches_finite_differences():
    import torch
    case, s, T11, Tp, w = _retrieval_setup(6, "fp64_accurate")
    t = case["t"]
    V = torch.tensor(case["verts"], dtype=torch.float64, device=case["dev"], requires_grad=True)
    ab, ae = t["beams"]["ampl"], t["ext"]["ampl"]
    L = _loss_of(case, s, V, ab, ae, T11, Tp, w)
    L.backward()
    grad = V.grad.cpu().numpy()
    smooth = _smooth_vertices(case, s)
    assert len(smooth) >= 20, len(smooth)
    rng = np.random.default_rng(5)
    pick = rng.choice(smooth, min(50, len(smooth)), replace=False)
    h = 1e-4
    got, fd4, fd2 = [], [], []
    with torch.no_grad():
        for vi in pick:
            for a in range(3):
                f = {}
                for m in (-2, -1, 1, 2):
                    Vp = V.detach().clone()
                    Vp[vi, a] += m * h
                    f[m] = float(_loss_of(case, s, Vp, ab, ae, T11, Tp, w))
                fd4.append((-f[2] + 8 * f[1] - 8 * f[-1] + f[-2]) / (12 * h))
                fd2.append((f[1] - f[-1]) / (2 * h))
                got.append(grad[vi, a])
    got, fd4, fd2 = map(np.array, (got, fd4, fd2))
    scale = np.abs(fd4).max()
    assert np.abs(got - fd4).max() <= 1e-6 * scale, (np.abs(got - fd4).max() / scale, got, fd4)
    # the 2nd-order stencil carries the (q h)^2 / 6 truncation of the form-factor phases
    # (|q| up to 2k = 23.6 /um): a few 1e-4 of the largest component
    assert np.abs(got - fd2).max() <= 5e-4 * scale


def test_amplitude_gradient_matches_finite_differences():
    import torch
    case, s, T11, Tp, w = _retrieval_setup(3, "fp64_accurate")
    t = case["t"]
    V = torch.tensor(case["verts"], dtype=torch.float64, device=case["dev"])
    ab = t["beams"]["ampl"].clone().requires_grad_(True)
    ae = t["ext"]["ampl"].clone().requires_grad_(True)
    L = _loss_of(case, s, V, ab, ae, T11, Tp, w)
    L.backward()
    rng = np.random.default_rng(6)
    for which, A in (("beams", ab), ("ext", ae)):
        g = A.grad.cpu().numpy()
        idx = rng.choice(A.shape[0], 4, replace=False)
        for b in idx:
            q = int(rng.integers(8))
            hh = 1e-6 * max(1e-12, float(np.abs(A[b].detach().cpu().numpy()).max()))
            f = []
            for m in (-1, 1):
                Ap = A.detach().clone()
                Ap[b, q] += m * hh
                args = (Ap, ae.detach()) if which == "beams" else (ab.detach(), Ap)
                with torch.no_grad():
                    f.append(float(_loss_of(case, s, V, *args, T11, Tp, w)))
            fd = (f[1] - f[0]) / (2 * hh)
            assert abs(fd - g[b, q]) <= 1e-6 * max(abs(fd), np.abs(g).max() * 1e-3), (which, b, q, fd, g[b, q])


def test_mixed_adjoint_matches_fp64_adjoint():
    import torch
    case, s64, T11, Tp, w = _retrieval_setup(4, "fp64_accurate")
    sm = _system(case, "mixed")
    t = case["t"]
    grads = []
    for s in (s64, sm):
        V = torch.tensor(case["verts"], dtype=torch.float64, device=case["dev"], requires_grad=True)
        L = _loss_of(case, s, V, t["beams"]["ampl"], t["ext"]["ampl"], T11, Tp, w)
        L.backward()
        grads.append(V.grad.cpu().numpy().ravel())
    g64, gm = grads
    rel = np.linalg.norm(gm - g64) / np.linalg.norm(g64)
    assert rel <= 1e-5, rel


def _directional_fd(case, s, V, ab, ae, T11, Tp, w, vi, d, h=1e-4):
    """4th- and 2nd-order central differences of the loss along d at vertex vi."""
    import torch
    f = {}
    with torch.no_grad():
        for m in (-2, -1, 1, 2):
            Vp = V.detach().clone()
            Vp[vi] += (m * h) * d
            f[m] = float(_loss_of(case, s, Vp, ab, ae, T11, Tp, w))
    return (-f[2] + 8 * f[1] - 8 * f[-1] + f[-2]) / (12 * h), (f[1] - f[-1]) / (2 * h)


def test_full_adjoint_workload_matches_finite_differences():
    """The adjoint workload at its full size: 64 orientations of the rough column traced to
    recursion depth 10, 361 x 90 bins, the retrieval loss against the particle with
    vertices perturbed by N(0, 0.05 um).  On 50 random particle vertices, the derivative
    along a random unit direction by central differences of the FP64 model's loss, step
    h = 1e-6 um: the FP64 adjoint within 1e-6 of the largest of the 50 (4th-order stencil);
    the mixed adjoint (FP64 phases, FP32 edge terms and polarisation, FP32 sums over groups
    of 8 bins) within 2e-5 of it; the two gradient vectors within 1e-5 in norm.

    The step is 1e-6 um, not the 1e-4 um of the smaller check above: at this size every
    vertex belongs to ~1800 apertures, and each has bins within 1e-6 (in 1 - |khat.e_y|) of
    the singular axis of the reference's polarisation, inside its clamp |k_y| <= 0.999999
    (diff_mod.f90:151), where the loss has kinks.  A kink inside the stencil shows as
    disagreement between the 2nd-order differences with steps h and 2h (their truncation is
    ~1e-10 here, their rounding ~1e-8 of the largest); such vertices (a few of the 50; the
    count falls with h) are left out, and at least 40 must remain."""
    import torch
    case, s64, T11, Tp, w = _retrieval_setup(64, "fp64_accurate", rec=10)
    t = case["t"]
    assert t["beam_off"][-1] > 64 * 2000
    ab, ae = t["beams"]["ampl"], t["ext"]["ampl"]
    grads = {}
    for name, s in (("fp64", s64), ("mixed", _system(case, "mixed"))):
        V = torch.tensor(case["verts"], dtype=torch.float64, device=case["dev"], requires_grad=True)
        L = _loss_of(case, s, V, ab, ae, T11, Tp, w)
        L.backward()
        grads[name] = V.grad.cpu().numpy()
    g64, gm = grads["fp64"], grads["mixed"]
    rel = np.linalg.norm(gm - g64) / np.linalg.norm(g64)
    assert rel <= 1e-5, rel
    used = np.nonzero(np.isfinite(_vertex_margins(case, s64)))[0]
    rng = np.random.default_rng(11)
    pick = rng.choice(used, 50, replace=False)
    V = torch.tensor(case["verts"], dtype=torch.float64, device=case["dev"])
    h = 1e-6
    fd4, fd2h, fd22h, d64, dm = [], [], [], [], []
    with torch.no_grad():
        for vi in pick:
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            dt = torch.tensor(d, dtype=torch.float64, device=case["dev"])
            f = {}
            for m in (-2, -1, 1, 2):
                Vp = V.clone()
                Vp[vi] += (m * h) * dt
                f[m] = float(_loss_of(case, s64, Vp, ab, ae, T11, Tp, w))
            fd4.append((-f[2] + 8 * f[1] - 8 * f[-1] + f[-2]) / (12 * h))
            fd2h.append((f[1] - f[-1]) / (2 * h))
            fd22h.append((f[2] - f[-2]) / (4 * h))
            d64.append(g64[vi] @ d)
            dm.append(gm[vi] @ d)
    fd4, fd2h, fd22h, d64, dm = map(np.array, (fd4, fd2h, fd22h, d64, dm))
    scale = np.abs(fd4).max()
    smooth = np.abs(fd2h - fd22h) <= 1e-6 * scale
    assert smooth.sum() >= 40, int(smooth.sum())
    e64 = np.abs(d64 - fd4)[smooth].max() / scale
    em = np.abs(dm - fd4)[smooth].max() / scale
    assert e64 <= 1e-6, (e64, em, int(smooth.sum()))
    assert em <= 2e-5, (e64, em, int(smooth.sum()))
