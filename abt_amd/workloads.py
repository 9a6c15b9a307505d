"""Seeded workloads: angular grids and synthetic beam sets.

`make_angles` restates the reference's grid builder (input_mod.f90:114-196)
and its degree-to-radian conversion with the REAL*4 pi (input_mod.f90:742,789,
misc_submod.f90:12), so grids match what the reference's `-theta`/`-phi` flags
produce.

Beam-set generators:
  * `random_triangles`    -- generic seeded triangular apertures (tests)
  * `stress_beams`        -- the seeded 200,000-aperture diffraction stress
                             test (planar triangles normal to their propagation
                             vector, centroids in a 100 um cube, edge directions
                             uniform in the aperture plane)
  * `column_beams`        -- beam sets shaped like the tracer's output for a
                             hexagonal column (ext-diff beams on illuminated
                             facets plus reflected / refracted outgoing beams
                             from a centroid-ray trace of the convex column).
                             The reference's near-field tracer is upstream of
                             the diffraction path and is not re-implemented;
                             this generator only reproduces the shape of its
                             output (beam counts, directions, amplitudes).
"""
from __future__ import annotations

import numpy as np

from .beams import BeamSet, concat

PI_R4 = float(np.float32(3.14159265358979))


def make_angles(points, splits):
    """input_mod.f90:114-196: piecewise-uniform angles in degrees."""
    out = []
    n = len(points)
    for i in range(n - 1):
        jmax = int(np.rint((points[i + 1] - points[i]) / splits[i]))
        for j in range(jmax + 1):
            if i != n - 2 and j == jmax:
                continue
            out.append(points[i] + j * splits[i])
    return np.array(out, dtype=np.float64)


def deg2rad_ref(a):
    return np.asarray(a, dtype=np.float64) * PI_R4 / 180.0


def grid(theta_spec, phi_spec):
    """theta/phi specs as the CLI takes them: (a0, s0, a1, s1, ..., an) in degrees."""
    def parse(spec):
        spec = list(spec)
        return make_angles(spec[0::2], spec[1::2])
    return deg2rad_ref(parse(theta_spec)), deg2rad_ref(parse(phi_spec))


# grids of the benchmark configurations (theta spec, phi spec), degrees
GRIDS = {
    "small": ((0, 0.5, 180), (0, 2, 358)),              # 361 x 180
    "stress": ((0, 0.1, 180), (0, 1, 359)),             # 1801 x 360
    "orient": ((0, 0.25, 180), (0, 2, 358)),            # 721 x 180
    "adjoint": ((0, 0.5, 180), (0, 4, 356)),            # 361 x 90
    "large": ((0, 0.05, 5, 0.5, 180), (0, 1, 359)),     # 451 x 360
}


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _perp_basis(p, rng):
    a = _unit(np.cross(p, rng.normal(size=p.shape)))
    b = np.cross(p, a)
    return a, b


def random_triangles(n, seed=0, box=10.0, size=(0.5, 4.0), amp_decades=0.0) -> BeamSet:
    rng = np.random.default_rng(seed)
    c = rng.uniform(-box / 2, box / 2, (n, 3))
    e1 = rng.normal(size=(n, 3)) * rng.uniform(*size, (n, 1)) / np.sqrt(3)
    e2 = rng.normal(size=(n, 3)) * rng.uniform(*size, (n, 1)) / np.sqrt(3)
    v1 = c - (e1 + e2) / 3
    tri = np.stack([v1, v1 + e1, v1 + e2], 1)
    prop = _unit(rng.normal(size=(n, 3)))
    vk7, _ = _perp_basis(prop, rng)
    a = rng.normal(size=(n, 8)) * (10.0 ** rng.uniform(-amp_decades, 0, (n, 1)))
    return BeamSet.from_triangles(a, vk7, prop, np.full(n, np.pi / 2), tri)


def reflected_facets(n, seed=0, box=40.0, size=(1.0, 4.0)) -> BeamSet:
    """External reflections of the incident wave (0,0,-1): n triangular facets with seeded
    normals (n_z >= 0.2, lit from above), each aperture the facet itself and its beam the
    mirror image prop = inc - 2 (inc.n) n, e-perp unit(n x prop).  The forward direction
    (0,0,-1) is the mirror image of every beam's prop in its aperture plane, where the
    Karczewski-Wolf amplitude of every beam vanishes."""
    rng = np.random.default_rng(seed)
    nrm = _unit(rng.normal(size=(n, 3)))
    nrm[:, 2] = np.abs(nrm[:, 2])
    nrm = _unit(np.where(nrm[:, 2:3] < 0.2, nrm + [0.0, 0.0, 0.4], nrm))
    inc = np.array([0.0, 0.0, -1.0])
    prop = inc - 2 * (nrm @ inc)[:, None] * nrm
    vk7 = _unit(np.cross(nrm, prop))
    a, b = _perp_basis(nrm, rng)
    c = rng.uniform(-box / 2, box / 2, (n, 3))
    ang = rng.uniform(0, 2 * np.pi, (n, 1))
    r = rng.uniform(*size, (n, 1))
    tri = np.stack([c + r * (np.cos(ang + t) * a + np.sin(ang + t) * b)
                    for t in (0.0, 2.1, 4.2)], 1)
    amp = rng.normal(size=(n, 8))
    return BeamSet.from_triangles(amp, vk7, prop, np.full(n, np.pi / 2), tri)


def stress_beams(n=200_000, seed=20240601, min_angle=0.0) -> BeamSet:
    """Diffraction stress test: n planar triangles, centroids uniform in a 100 um cube,
    two edge vectors of length U(0.3, 5) um in the plane normal to the propagation
    vector (uniform on the sphere), each with a uniform direction in that plane (so
    sliver triangles occur as often as the stated distribution makes them), amplitude
    entries (N(0,1) + i N(0,1)) * 10^U(-6,0), e-perp a seeded unit vector normal to prop.
    min_angle > 0 instead keeps the angle between the two edges in
    [min_angle, pi - min_angle] (no slivers; the earlier form of this generator, which
    some tests use for its particular apertures).  Generator: numpy PCG64 with `seed`."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 100.0, (n, 3))
    prop = _unit(rng.normal(size=(n, 3)))
    a, b = _perp_basis(prop, rng)
    ang1 = rng.uniform(0, 2 * np.pi, (n, 1))
    if min_angle > 0:
        ang2 = ang1 + rng.uniform(min_angle, np.pi - min_angle, (n, 1)) * rng.choice([-1.0, 1.0], (n, 1))
    else:
        ang2 = rng.uniform(0, 2 * np.pi, (n, 1))
    l1 = rng.uniform(0.3, 5.0, (n, 1))
    l2 = rng.uniform(0.3, 5.0, (n, 1))
    e1 = l1 * (np.cos(ang1) * a + np.sin(ang1) * b)
    e2 = l2 * (np.cos(ang2) * a + np.sin(ang2) * b)
    v1 = c - (e1 + e2) / 3
    tri = np.stack([v1, v1 + e1, v1 + e2], 1)
    scale = 10.0 ** rng.uniform(-6, 0, (n, 1))
    ampl = rng.normal(size=(n, 8)) * scale
    vk7, _ = _perp_basis(prop, rng)
    return BeamSet.from_triangles(ampl, vk7, prop, np.full(n, np.pi / 2), tri)


# --------------------------------------------------------------------------
# hexagonal column geometry and a centroid-ray beam generator

def hex_column(radius=5.0, length=10.0, nfhr=6, nfpl=12, sigma=0.0, seed=0):
    """Triangulated hexagonal column centred at the origin, axis along z.
    Prism faces: nfhr x nfpl grid of quads split into 2 triangles each;
    basal faces: fan of 6*nfhr... triangles.  Optional Gaussian roughness
    `sigma` moves interior vertices along the face normal (the surface stays
    closed).  Returns (verts (nv,3), faces (nf,3) int, counter-clockwise seen
    from outside)."""
    rng = np.random.default_rng(seed)
    ang = np.arange(6) * np.pi / 3
    corners = np.stack([radius * np.cos(ang), radius * np.sin(ang)], 1)
    verts = []
    index = {}

    def vid(p, jitter_dir=None):
        key = tuple(np.round(p, 9))
        if key in index:
            return index[key]
        q = np.array(p, float)
        if jitter_dir is not None and sigma > 0:
            q = q + rng.normal() * sigma * jitter_dir
        index[key] = len(verts)
        verts.append(q)
        return index[key]

    faces = []
    zs = np.linspace(-length / 2, length / 2, nfpl + 1)
    for s in range(6):
        c0, c1 = corners[s], corners[(s + 1) % 6]
        nrm = np.array([*(c0 + c1) / 2, 0.0])
        nrm /= np.linalg.norm(nrm)
        for a in range(nfhr):
            for b in range(nfpl):
                def P(ia, ib):
                    t = ia / nfhr
                    xy = c0 + t * (c1 - c0)
                    interior = 0 < ia < nfhr and 0 < ib < nfpl
                    return vid([xy[0], xy[1], zs[ib]], nrm if interior else None)
                p00, p10, p11, p01 = P(a, b), P(a + 1, b), P(a + 1, b + 1), P(a, b + 1)
                faces.append((p00, p10, p11))
                faces.append((p00, p11, p01))
    for zsgn in (-1, 1):
        z = zsgn * length / 2
        nrm = np.array([0, 0, zsgn], float)
        for s in range(6):
            c0, c1 = corners[s], corners[(s + 1) % 6]
            for r in range(nfhr):
                # ring strip between radius fractions r/nfhr and (r+1)/nfhr
                f0, f1 = r / nfhr, (r + 1) / nfhr
                n_out = r + 1
                for t in range(n_out):
                    def Q(f, i, n):
                        if f == 0:
                            return vid([0.0, 0.0, z], nrm)
                        xy = f * (c0 + (i / n) * (c1 - c0))
                        interior = f < 1
                        return vid([xy[0], xy[1], z], nrm if interior else None)
                    a0 = Q(f1, t, n_out); a1 = Q(f1, t + 1, n_out)
                    b0 = Q(f0, t, max(r, 1)) if r > 0 else Q(0, 0, 1)
                    tri = (b0, a0, a1)
                    faces.append(tri if zsgn > 0 else tri[::-1])
                    if r > 0 and t < r:
                        b1 = Q(f0, t + 1, r)
                        tri = (b0, a1, b1)
                        faces.append(tri if zsgn > 0 else tri[::-1])
    return np.array(verts), np.array(faces, dtype=np.int64)


def euler_matrix(alpha, beta, gamma):
    """zyz rotation (degrees), the usual Mishchenko particle-orientation convention."""
    a, b, g = np.deg2rad([alpha, beta, gamma])
    Rz = lambda t: np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    Ry = lambda t: np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    return Rz(a) @ Ry(b) @ Rz(g)


def _fresnel(cos_i, m):
    sin_i2 = np.clip(1 - cos_i ** 2, 0, None)
    cos_t = np.sqrt(1 - sin_i2 / m ** 2 + 0j)
    rs = (cos_i - m * cos_t) / (cos_i + m * cos_t)
    rp = (m * cos_i - cos_t) / (cos_t + m * cos_i)
    ts = 2 * cos_i / (cos_i + m * cos_t)
    tp = 2 * cos_i / (cos_t + m * cos_i)
    return rs, rp, ts, tp, cos_t


def _fresnel_t(cos_i, m):
    """Fresnel coefficients (torch, complex128) for a real cosine and complex index m."""
    import torch
    ci = cos_i.to(torch.complex128)
    sin_i2 = torch.clamp(1 - cos_i ** 2, min=0).to(torch.complex128)
    cos_t = torch.sqrt(1 - sin_i2 / m ** 2)
    rs = (ci - m * cos_t) / (ci + m * cos_t)
    rp = (m * ci - cos_t) / (cos_t + m * ci)
    ts = 2 * ci / (ci + m * cos_t)
    tp = 2 * ci / (cos_t + m * ci)
    return rs, rp, ts, tp, cos_t


def trace_orientations(verts, faces, rots, la=0.532, m=1.31 + 0j, rec=8, device="cpu",
                       max_elems=1 << 25):
    """Beam sets shaped like the tracer's output, for many orientations of one particle.

    Incident plane wave along (0,0,-1) in the lab frame, e-perp (1,0,0), amplitude
    matrix I.  For orientation o the particle is rotated by rots[o] (lab = R . particle);
    the trace runs in the particle frame (incidence -R[2,:]) so every ray of every
    orientation tests the same facet planes, and outputs are rotated to the lab frame.
      ext-diff beams: one per illuminated facet, prop (0,0,-1), vk7 = unit(n x prop),
        amplitude = basis rotation times the path phase from the top of the particle
        (as at beam_loop_mod.f90:502-608);
      outgoing beams: per illuminated facet the external reflection, then the refracted
        centroid ray through the body: at each internal hit (the nearest facet plane
        ahead whose triangle contains the hit point) the transmitted part leaves
        through that facet (one outgoing beam, aperture = that facet) unless totally
        reflected, and the reflected part continues, `rec` times.
    Beams are ordered by (orientation, entry facet, depth), ext beams by (orientation,
    facet).  The reference's near-field tracer is upstream of the diffraction path and
    is not re-implemented; this reproduces the shape of its output (beam counts,
    apertures, directions, amplitudes).

    Returns a dict of torch tensors on `device`: "beams" and "ext" (each with ampl
    (n,8), vk7, prop, fov, vert_off, verts, fi (1-based facet ids), orient), and
    "beam_off", "ext_off" (norient+1 offsets)."""
    import torch
    dt = torch.float64
    dev = torch.device(device)
    V = torch.as_tensor(np.asarray(verts), dtype=dt, device=dev)
    F = torch.as_tensor(np.asarray(faces), dtype=torch.int64, device=dev)
    R = torch.as_tensor(np.asarray(rots), dtype=dt, device=dev).reshape(-1, 3, 3)
    O = R.shape[0]
    unit = lambda v: v / torch.linalg.norm(v, dim=-1, keepdim=True)
    dot = lambda a, b: (a * b).sum(-1)
    P = V[F]
    cen = P.mean(1)
    nrm = unit(torch.linalg.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]))
    flip = dot(nrm, cen - V.mean(0)) < 0
    nrm[flip] *= -1
    d0 = dot(P[:, 0], nrm)
    nf = F.shape[0]
    inc = -R[:, 2, :]
    ex, ey = R[:, 0, :], R[:, 1, :]
    k = 2 * np.pi / la
    ztop = (-(V @ inc.T)).max(0).values
    oi, fi = torch.nonzero((inc @ nrm.T) < -1e-9, as_tuple=True)
    nr = oi.shape[0]
    n = nrm[fi]
    ic = inc[oi]
    vk = unit(torch.linalg.cross(n, ic, dim=-1))
    phase = torch.exp(1j * k * (ztop[oi] + dot(cen[fi], ic)))
    c, s = dot(vk, ex[oi]), dot(vk, ey[oi])
    amp = torch.stack([torch.stack([c, s], -1), torch.stack([-s, c], -1)], 1).to(torch.complex128)
    amp = amp * phase[:, None, None]
    cos_i = -dot(n, ic)
    rs, rp, ts, tp, cos_t = _fresnel_t(cos_i, m)
    refl = ic - 2 * dot(ic, n)[:, None] * n
    ray = torch.arange(nr, device=dev)
    keys, fis, amps, vks, props = [ray * (rec + 2)], [fi], [amp * torch.stack([rp, rs], 1)[:, :, None]], \
        [vk], [unit(refl)]
    eta = 1 / m.real
    tdir = unit(eta * ic + (eta * cos_i - cos_t.real)[:, None] * n)
    field = amp * torch.stack([tp, ts], 1)[:, :, None]
    pos = cen[fi].clone()
    vkr = vk
    blk = max(1, max_elems // max(nf, 1))
    for depth in range(rec):
        nl = ray.shape[0]
        if nl == 0:
            break
        j = torch.empty(nl, dtype=torch.int64, device=dev)
        tj = torch.empty(nl, dtype=dt, device=dev)
        for b0 in range(0, nl, blk):
            td, ps = tdir[b0:b0 + blk], pos[b0:b0 + blk]
            denom = td @ nrm.T
            ok = denom > 1e-12
            tt = torch.where(ok, (d0[None, :] - ps @ nrm.T) / torch.where(ok, denom, torch.ones_like(denom)),
                             torch.full_like(denom, float("inf")))
            tt = torch.where(tt < 1e-9, torch.full_like(tt, float("inf")), tt)
            tmin, jj = tt.min(1)
            # coplanar facets share a plane distance: take the lowest-index facet whose
            # triangle contains the hit point (the nearest plane's if none does)
            cand = (tt <= tmin[:, None] * (1 + 1e-9) + 1e-9) & torch.isfinite(tmin)[:, None]
            ri, fj = torch.nonzero(cand, as_tuple=True)
            hit = ps[ri] + tmin[ri, None] * td[ri]
            a_, b_, c_ = P[fj, 0], P[fj, 1], P[fj, 2]
            nn = nrm[fj]
            inside = ((dot(torch.linalg.cross(b_ - a_, hit - a_, dim=-1), nn) >= -1e-9) &
                      (dot(torch.linalg.cross(c_ - b_, hit - b_, dim=-1), nn) >= -1e-9) &
                      (dot(torch.linalg.cross(a_ - c_, hit - c_, dim=-1), nn) >= -1e-9))
            first = torch.full((td.shape[0],), nf, dtype=torch.int64, device=dev)
            first.scatter_reduce_(0, ri[inside], fj[inside], reduce="amin")
            jj = torch.where(first < nf, first, jj)
            j[b0:b0 + blk] = jj
            tj[b0:b0 + blk] = tt.gather(1, jj[:, None])[:, 0]
        live = torch.isfinite(tj)
        ray, j, tj, tdir, field, pos, vkr = (x[live] for x in (ray, j, tj, tdir, field, pos, vkr))
        pos = pos + tj[:, None] * tdir
        n2 = nrm[j]
        ci = dot(tdir, n2)
        rs2, rp2, ts2, tp2, ct2 = _fresnel_t(ci, 1 / m)
        vk2 = torch.linalg.cross(n2, tdir, dim=-1)
        nv2 = torch.linalg.norm(vk2, dim=-1, keepdim=True)
        vk2 = torch.where(nv2 > 1e-12, vk2 / torch.where(nv2 > 1e-12, nv2, torch.ones_like(nv2)), vkr)
        field = field * torch.exp(1j * k * m * tj)[:, None, None]
        exits = (ct2.imag.abs() < 1e-12) & (ct2.real > 0)
        out_dir = unit(m.real * tdir + (ct2.real - m.real * ci)[:, None] * n2)
        keys.append(ray[exits] * (rec + 2) + depth + 1)
        fis.append(j[exits])
        amps.append((field * torch.stack([tp2, ts2], 1)[:, :, None])[exits])
        vks.append(vk2[exits])
        props.append(out_dir[exits])
        field = field * torch.stack([rp2, rs2], 1)[:, :, None]
        tdir = unit(tdir - 2 * ci[:, None] * n2)
    key = torch.cat(keys)
    order = torch.sort(key, stable=True).indices
    bray = torch.div(key[order], rec + 2, rounding_mode="floor")

    def pack(A, vkp, pp, facet, orient):
        nb = facet.shape[0]
        Ro = R[orient]
        a8 = torch.stack([A[:, 0, 0].real, A[:, 0, 0].imag, A[:, 1, 0].real, A[:, 1, 0].imag,
                          A[:, 0, 1].real, A[:, 0, 1].imag, A[:, 1, 1].real, A[:, 1, 1].imag], 1)
        off = torch.zeros(O + 1, dtype=torch.int64, device=dev)
        off[1:] = torch.cumsum(torch.bincount(orient, minlength=O), 0)
        return {"ampl": a8.contiguous(),
                "vk7": torch.einsum("bij,bj->bi", Ro, vkp).contiguous(),
                "prop": torch.einsum("bij,bj->bi", Ro, pp).contiguous(),
                "fov": torch.full((nb,), np.pi / 2, dtype=dt, device=dev),
                "vert_off": torch.arange(nb + 1, dtype=torch.int64, device=dev) * 3,
                "verts": torch.einsum("bij,bvj->bvi", Ro, P[facet]).reshape(-1, 3).contiguous(),
                "fi": facet + 1, "orient": orient}, off

    beams, boff = pack(torch.cat(amps)[order], torch.cat(vks)[order], torch.cat(props)[order],
                       torch.cat(fis)[order], oi[bray])
    ext, eoff = pack(amp, vk, ic, fi, oi)
    return {"beams": beams, "ext": ext, "beam_off": boff, "ext_off": eoff}


def to_beamset(d) -> BeamSet:
    """A dict of tensors from trace_orientations (one tree) as a host BeamSet."""
    g = lambda x: x.detach().cpu().numpy()
    return BeamSet(g(d["ampl"]), g(d["vk7"]), g(d["prop"]), g(d["fov"]), g(d["vert_off"]), g(d["verts"]))


def column_beams(verts, faces, rot, la=0.532, m=1.31 + 0j, rec=8, seed=0, return_fi=False):
    """One orientation of trace_orientations as host beam sets.
    Returns (beams, ext, geometry verts rotated to the lab frame), plus the 1-based
    facet ids of the beams and of the ext-diff beams when return_fi is set."""
    t = trace_orientations(verts, faces, np.asarray(rot)[None], la=la, m=m, rec=rec)
    V = np.asarray(verts) @ np.asarray(rot).T
    b, e = to_beamset(t["beams"]), to_beamset(t["ext"])
    bfi, efi = t["beams"]["fi"].numpy(), t["ext"]["fi"].numpy()
    # apertures gathered from the rotated geometry itself, so that they are bit-identical
    # to what a caller looking vertices up through the facet ids gets (diff_mod.f90:529-531)
    F = np.asarray(faces)
    b.verts = np.ascontiguousarray(V[F[bfi - 1]].reshape(-1, 3))
    e.verts = np.ascontiguousarray(V[F[efi - 1]].reshape(-1, 3))
    if return_fi:
        return b, e, V, bfi, efi
    return b, e, V


def random_rotations(n, seed):
    """n rotation matrices uniform on SO(3) (unit quaternions from a seeded normal)."""
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                     np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                     np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


# particles of the multi-orientation workloads: (radius, length, nfhr, nfpl, sigma, m, rec)
PARTICLES = {
    "rough_column": (10.0, 20.0, 8, 14, 0.5, 1.31 + 0j, 10),     # 2112 facets
    "large_column": (50.0, 100.0, 24, 48, 0.0, 1.31 + 0.01j, 12),  # 20736 facets
}


def orientation_workload(norient, particle="rough_column", seed=0, la=0.532, device="cpu"):
    """Beams of `norient` orientations (uniform on SO(3), seeded) of a PARTICLES entry."""
    radius, length, nfhr, nfpl, sigma, m, rec = PARTICLES[particle]
    verts, faces = hex_column(radius, length, nfhr, nfpl, sigma, seed=seed)
    rots = random_rotations(norient, seed + 1)
    t = trace_orientations(verts, faces, rots, la=la, m=m, rec=rec, device=device)
    t["verts0"], t["faces"], t["rots"] = verts, faces, rots
    return t


# ---- facet-visibility workloads (the near-field tracer's find_vis_* inputs) ----

def hex_column_apertures(radius=5.0, length=10.0, nfhr=6, nfpl=12, sigma=0.0, seed=0):
    """hex_column plus its apertures: the 6 prism faces (ids 1..6), the basal faces
    z = -length/2 (7) and z = +length/2 (8).  Returns (verts, faces, ap, ap_normals)."""
    verts, faces = hex_column(radius, length, nfhr, nfpl, sigma, seed=seed)
    ang = np.arange(6) * np.pi / 3 + np.pi / 6
    side = 2 * nfhr * nfpl
    base = 6 * nfhr * nfhr
    ap = np.concatenate([np.repeat(np.arange(1, 7), side), np.full(base, 7), np.full(base, 8)])
    assert len(ap) == len(faces)
    normals = np.concatenate([np.stack([np.cos(ang), np.sin(ang), np.zeros(6)], 1),
                              [[0.0, 0.0, -1.0], [0.0, 0.0, 1.0]]])
    return verts, faces, ap, normals


def vis_workload(nbeam, mode="int", particle="rough_column", seed=0):
    """`nbeam` beams for find_vis_*: the particle in a random orientation per beam (uniform on
    SO(3), seeded) seen along -z, and for internal (external) beams the facets of one
    up-facing (down-facing) aperture, picked at random, as the illuminating facets.
    Returns (topology VisGeometry, v (nbeam,3,nv), mid (nbeam,nf,3), nz (nbeam,nf),
    ap_nz (nbeam,na), in_off (nbeam+1), field_in (1-based), beam_mid_z (nbeam))."""
    from .visibility import VisGeometry
    radius, length, nfhr, nfpl, sigma = PARTICLES[particle][:5]
    verts, faces, ap, apn = hex_column_apertures(radius, length, nfhr, nfpl, sigma, seed=seed)
    topo = VisGeometry.from_mesh(verts, faces, ap, apn)
    rots = random_rotations(nbeam, seed + 7)
    rng = np.random.default_rng(seed + 11)
    nf, nv, na = topo.nf, topo.nv, topo.na
    v = np.empty((nbeam, 3, nv)); mid = np.empty((nbeam, nf, 3))
    nz = np.empty((nbeam, nf)); ap_nz = np.empty((nbeam, na))
    in_off = np.zeros(nbeam + 1, np.int64); fins = []; bmz = np.empty(nbeam)
    for b in range(nbeam):
        g = topo.rotated(rots[b])
        v[b], mid[b], nz[b], ap_nz[b] = g.v.T, g.mid, g.nz, g.ap_nz
        bmz[b] = g.v[:, 2].max() + 1.0
        if mode == "inc":
            fin = np.zeros(0, np.int64)
        else:
            want = g.ap_nz > 0.01 if mode == "int" else g.ap_nz < -0.01
            a = rng.choice(np.nonzero(want)[0]) + 1
            fin = np.nonzero(topo.ap == a)[0] + 1
        fins.append(fin)
        in_off[b + 1] = in_off[b] + len(fin)
    field_in = np.concatenate(fins) if fins else np.zeros(0, np.int64)
    return topo, v, mid, nz, ap_nz, in_off, field_in, bmz
