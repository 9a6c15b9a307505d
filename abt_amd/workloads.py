"""Seeded workloads: angular grids and synthetic beam sets.

`make_angles` restates the reference's grid builder (input_mod.f90:114-196)
and its degree-to-radian conversion with the REAL*4 pi (input_mod.f90:742,789,
misc_submod.f90:12), so grids match what the reference's `-theta`/`-phi` flags
produce.

Beam-set generators:
  * `random_triangles`    -- generic seeded triangular apertures (tests)
  * `stress_beams`        -- the seeded 200,000-aperture diffraction stress
                             test (planar triangles normal to their propagation
                             vector, centroids in a 100 um cube)
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


def stress_beams(n=200_000, seed=20240601) -> BeamSet:
    """Diffraction stress test: n planar triangles, centroids uniform in a 100 um cube,
    two edge vectors of length U(0.3, 5) um in the plane normal to the propagation
    vector (uniform on the sphere) at an angle in [0.3, pi - 0.3] rad to each other,
    amplitude entries (N(0,1) + i N(0,1)) * 10^U(-6,0), e-perp a seeded unit vector
    normal to prop.  Generator: numpy PCG64 with `seed`."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 100.0, (n, 3))
    prop = _unit(rng.normal(size=(n, 3)))
    a, b = _perp_basis(prop, rng)
    ang1 = rng.uniform(0, 2 * np.pi, (n, 1))
    # angle between the two edges in [0.3, pi - 0.3] rad either way round: no sliver triangles
    ang2 = ang1 + rng.uniform(0.3, np.pi - 0.3, (n, 1)) * rng.choice([-1.0, 1.0], (n, 1))
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


def column_beams(verts, faces, rot, la=0.532, m=1.31 + 0j, rec=8, seed=0, return_fi=False):
    """Beam sets shaped like the tracer's output for a convex particle.

    Incident plane wave along (0,0,-1), e-perp (1,0,0), amplitude matrix I.
    ext-diff beams: one per illuminated facet, prop (0,0,-1), vk7 = unit(n x prop),
    amplitude = basis rotation times the path phase (as at beam_loop_mod.f90:502-608).
    Outgoing beams: the external reflection from each illuminated facet, then the
    refracted field followed by a centroid ray through the convex body: at each
    internal hit the transmitted part leaves through that facet (one outgoing
    beam, aperture = that facet) and the reflected part continues, `rec` times.
    Returns (beams, ext, geometry verts rotated), plus the 1-based facet ids of
    the beams and of the ext-diff beams when return_fi is set."""
    rng = np.random.default_rng(seed)
    V = verts @ rot.T
    F = faces
    P = V[F]                                  # (nf, 3, 3)
    cen = P.mean(1)
    nrm = _unit(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]))
    # outward orientation: the particle is centred at the origin and convex-ish
    flip = np.einsum("ij,ij->i", nrm, cen - V.mean(0)) < 0
    nrm[flip] *= -1
    k = 2 * np.pi / la
    inc = np.array([0.0, 0, -1.0])
    ill = nrm @ inc < -1e-9
    ztop = V[:, 2].max()
    ext_list, beam_list = [], []

    beam_fi, ext_fi = [], []

    def rec_entry(i, amp, vk, prop):
        tri = P[i]
        beam_list.append((amp, vk, prop, tri))
        beam_fi.append(i + 1)

    for i in np.nonzero(ill)[0]:
        n = nrm[i]
        vk = _unit(np.cross(n, inc))
        phase = np.exp(1j * k * (ztop - cen[i, 2]))
        c, s = vk[0], vk[1]  # rotation from incident e-perp (1,0,0) to vk about z
        R = np.array([[c, s], [-s, c]])
        amp = R * phase
        ext_list.append((amp, vk, inc.copy(), P[i]))
        ext_fi.append(i + 1)
        cos_i = -n @ inc
        rs, rp, ts, tp, cos_t = _fresnel(cos_i, m)
        refl = inc - 2 * (inc @ n) * n
        rec_entry(i, np.diag([rp, rs]) @ amp, vk, _unit(refl))
        # refracted direction
        eta = 1 / m.real
        tdir = _unit(eta * inc + (eta * cos_i - np.real(cos_t)) * n)
        field = np.diag([tp, ts]) @ amp
        pos = cen[i]
        for depth in range(rec):
            # exit facet: first plane hit along tdir from pos (convex hull)
            denom = nrm @ tdir
            tt = np.einsum("ij,ij->i", P[:, 0] - pos, nrm) / np.where(denom > 1e-12, denom, np.nan)
            tt[~(denom > 1e-12)] = np.inf
            tt[tt < 1e-9] = np.inf
            j = int(np.argmin(tt))
            if not np.isfinite(tt[j]):
                break
            pos = pos + tt[j] * tdir
            n2 = nrm[j]
            ci = tdir @ n2
            rs2, rp2, ts2, tp2, ct2 = _fresnel(ci, 1 / m)
            vk2 = np.cross(n2, tdir)
            vk2 = _unit(vk2) if np.linalg.norm(vk2) > 1e-12 else vk
            field = field * np.exp(1j * k * m * tt[j])
            if np.iscomplexobj(ct2) and abs(np.imag(ct2)) < 1e-12 and np.real(ct2) > 0:
                out_dir = _unit(m.real * tdir + (np.real(ct2) - m.real * ci) * n2)
                rec_entry(j, np.diag([tp2, ts2]) @ field, vk2, out_dir)
            field = np.diag([rp2, rs2]) @ field
            tdir = _unit(tdir - 2 * (tdir @ n2) * n2)

    def pack(lst):
        if not lst:
            return BeamSet.empty()
        amp = np.array([a for a, *_ in lst])  # (nb, 2, 2) complex, [r, c]
        a8 = np.stack([amp[:, 0, 0].real, amp[:, 0, 0].imag, amp[:, 1, 0].real, amp[:, 1, 0].imag,
                       amp[:, 0, 1].real, amp[:, 0, 1].imag, amp[:, 1, 1].real, amp[:, 1, 1].imag], 1)
        vk = np.array([v for _, v, _, _ in lst])
        pr = np.array([p for _, _, p, _ in lst])
        tri = np.array([t for *_, t in lst])
        return BeamSet.from_triangles(a8, vk, pr, np.full(len(lst), np.pi / 2), tri)

    if return_fi:
        return (pack(beam_list), pack(ext_list), V, np.array(beam_fi, dtype=np.int64),
                np.array(ext_fi, dtype=np.int64))
    return pack(beam_list), pack(ext_list), V


def column_orientations(n, seed=0, radius=10.0, length=20.0, nfhr=8, nfpl=14, sigma=0.5,
                        rec=10, la=0.532):
    """Beam sets of `n` orientations drawn uniformly on SO(3) (seeded) of one rough column."""
    verts, faces = hex_column(radius, length, nfhr, nfpl, sigma, seed=seed)
    rng = np.random.default_rng(seed + 1)
    out = []
    for _ in range(n):
        qv = _unit(rng.normal(size=4))
        w, x, y, z = qv
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        b, e, _ = column_beams(verts, faces, R, la=la, rec=rec)
        out.append((b, e))
    return out
