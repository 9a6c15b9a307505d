"""Pins of the CPU oracle (oracle/diff_oracle.c) against checks that do not
restate the reference's formulas: closed forms, invariants, textbook
reductions and brute-force quadrature.  CPU only."""
import numpy as np
import pytest

from abt_amd.beams import BeamSet
from tests.formfactor import (ff_divided_difference, ff_quadrature, jones_to_mueller, kw_matrix,
                              tri_area)

LAM = 0.532
PI_R4 = 3.1415927410125732
K = 2 * PI_R4 / LAM
D2R = PI_R4 / 180


def _unit(v):
    return v / np.linalg.norm(v)


def test_pi_and_constants():
    # misc_submod.f90:12 declares pi with a default-kind literal
    assert PI_R4 == float(np.float32(3.14159265358979))
    assert K == pytest.approx(11.810499026363056, rel=1e-15)


def test_make_far_field_bins_closed_form(orc):
    theta = np.linspace(0, np.pi, 37)
    phi = np.linspace(0, 2 * np.pi, 24, endpoint=False)
    x, y, z = orc.make_far_field_bins(theta, phi)
    TH, PH = np.meshgrid(theta, phi)  # (nphi, ntheta), phi fastest in memory order F
    np.testing.assert_allclose(x, (1e6 * np.sin(TH) * np.cos(PH)).ravel(order="F"), atol=1e-8)
    np.testing.assert_allclose(y, (1e6 * np.sin(TH) * np.sin(PH)).ravel(order="F"), atol=1e-8)
    np.testing.assert_allclose(z, (-1e6 * np.cos(TH)).ravel(order="F"), atol=1e-8)


@pytest.mark.parametrize("seed", range(6))
def test_rotation_matrix_invariants(orc, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(3, 3)) * 4
    v0 = v - v.mean(0)
    R = orc.get_rotation_matrix2(v0)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)
    n = _unit(np.cross(v0[0], v0[1]))
    # aperture lies in the xy plane, normal mapped to -z (the 180 degree flip convention)
    np.testing.assert_allclose(R @ n, [0, 0, -1], atol=1e-13)
    np.testing.assert_allclose((v0 @ R.T)[:, 2], 0, atol=1e-13)


@pytest.mark.parametrize("seed", range(6))
def test_beam_frame_matches_independent_construction(orc, seed):
    """rot2.rot is fixed by the facet normal and prop alone: z' = -n, x' along prop's
    in-plane part, y' = z' x x'.  The anti-parallel flag is the sign of (vk7 x prop).n."""
    rng = np.random.default_rng(100 + seed)
    v = rng.normal(size=(3, 3)) * 3
    prop = _unit(rng.normal(size=3))
    vk7 = _unit(np.cross(prop, rng.normal(size=3)))
    fr = orc.beam_frame(np.r_[1, 0, 0, 0, 0, 0, 1, 0.], vk7, prop, 1.0, v)
    RR = fr["rot2"] @ fr["rot"]
    v0 = v - v.mean(0)
    n = _unit(np.cross(v0[0], v0[1]))
    ez = -n
    ex = _unit(prop - prop.dot(ez) * ez)
    RRi = np.array([ex, np.cross(ez, ex), ez])
    np.testing.assert_allclose(RR, RRi, atol=1e-13)
    np.testing.assert_allclose(fr["prop2"], RRi @ prop, atol=1e-13)
    np.testing.assert_allclose(fr["incidence2"], RRi @ [0, 0, -1.0], atol=1e-13)
    assert fr["anti_parallel"] == (np.cross(vk7, prop).dot(n) < 0)
    np.testing.assert_allclose(fr["com"], v.mean(0), rtol=1e-15)


@pytest.mark.parametrize("seed", range(8))
def test_karczewski_against_vector_kirchhoff(orc, seed):
    rng = np.random.default_rng(200 + seed)
    a = rng.uniform(0.1, 1.4)
    Kv = np.array([np.sin(a), 0.0, -np.cos(a)])
    kh = _unit(rng.normal(size=3))
    r = 1e6
    d, m, k = orc.karczewski(Kv, *(kh * r), r)
    np.testing.assert_allclose(k, kh, atol=1e-15)
    np.testing.assert_allclose(m, _unit(np.array([0, 1.0, 0]) - kh[1] * kh), atol=1e-13)
    np.testing.assert_allclose(d, kw_matrix(kh, Kv), atol=1e-13)


def test_karczewski_forward_limit_and_clamp(orc):
    Kv = np.array([np.sin(0.3), 0.0, -np.cos(0.3)])
    d, m, k = orc.karczewski(Kv, *(Kv * 1e6), 1e6)
    # forward: polarisation preserved, scaled by the obliquity factor -K_z
    np.testing.assert_allclose(d, -Kv[2] * np.eye(2), atol=1e-15)
    # |k_y| above the REAL*4 0.999999 is clamped to it, sign kept
    kh = _unit(np.array([1e-5, -1.0, 2e-5]))
    d, m, k = orc.karczewski(Kv, *(kh * 1e6), 1e6)
    assert k[1] == -float(np.float32(0.999999))


def _aperture_case(rng, nb_bins=40, special=None):
    v = rng.normal(size=(3, 3)) * 2.5
    prop = _unit(rng.normal(size=3))
    if special is not None:
        prop = special
    vk7 = _unit(np.cross(prop, rng.normal(size=3)))
    return v, prop, vk7


@pytest.mark.parametrize("seed", range(10))
def test_contour_integral_is_form_factor(orc, seed):
    """Green's-theorem reduction: the reference's edge sum equals +i F(q) for the
    clockwise aperture frame, F(q) = int exp(i q.r) dA, so
    area_facs2 = exp(i k |x3|) i F(q) / lambda with q = k (prop2 - x3/|x3|)_xy."""
    rng = np.random.default_rng(300 + seed)
    v, prop, vk7 = _aperture_case(rng)
    fr = orc.beam_frame(np.r_[1, 0, 0, 0, 0, 0, 1, 0.], vk7, prop, 1.0, v)
    RR = fr["rot2"] @ fr["rot"]
    v0 = v - v.mean(0)
    dirs = rng.normal(size=(40, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    x3 = (dirs * 1e6 - fr["com"]) @ RR.T
    got = orc.contour_integral(LAM, fr["rot"], fr["rot2"], v0, fr["prop2"], x3[:, 0], x3[:, 1],
                               x3[:, 2])
    p2 = (v0 @ RR.T)[:, :2]
    orient = np.sign(np.cross(np.r_[p2[1] - p2[0], 0], np.r_[p2[2] - p2[0], 0])[2])
    assert orient < 0  # the aperture frame traverses vertices clockwise
    for b in range(40):
        # |x3| summed in the same order as the oracle, so the 1.2e7 rad phase agrees
        r = np.sqrt(x3[b, 0] * x3[b, 0] + x3[b, 1] * x3[b, 1] + x3[b, 2] * x3[b, 2])
        q = K * (fr["prop2"][:2] - x3[b, :2] / r)
        F = ff_divided_difference(q, p2)
        Fq = ff_quadrature(q, p2)
        assert abs(F - Fq) <= 1e-10 * max(abs(F), 1e-3)
        want = np.exp(1j * K * r) * 1j * F / LAM
        scale = max(abs(want), 1e-3 * tri_area(p2) / LAM)
        assert abs(got[b] - want) <= 1e-9 * scale, (b, got[b], want)


@pytest.mark.parametrize("edge", ["vertical", "horizontal"])
def test_contour_integral_axis_aligned_edges(orc, edge):
    """Edges with dx or dy zero to rounding in the aperture frame go through the slope
    clamps (diff_mod.f90:379-382); the result is still the form factor."""
    # the aperture frame puts x' along prop's in-plane part; prop along +x, normal +z
    if edge == "vertical":
        v = np.array([[0.0, 0, 0], [0, 3.0, 0], [2.0, 0, 0]])
    else:
        v = np.array([[0.0, 0, 0], [3.0, 0, 0], [0, 2.0, 0]])
    prop = _unit(np.array([0.3, 0.0, -1.0]))
    vk7 = np.array([0.0, 1.0, 0.0])
    fr = orc.beam_frame(np.r_[1, 0, 0, 0, 0, 0, 1, 0.], vk7, prop, 1.0, v)
    RR = fr["rot2"] @ fr["rot"]
    v0 = v - v.mean(0)
    p2 = (v0 @ RR.T)[:, :2]
    d = np.diff(np.vstack([p2, p2[:1]]), axis=0)
    # an edge is axis-aligned to rounding, so a slope clamp (|m| or |1/m| < 1e-9) fires
    assert np.any(np.abs(d) < 1e-12)
    rng = np.random.default_rng(7)
    dirs = rng.normal(size=(30, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    x3 = (dirs * 1e6 - fr["com"]) @ RR.T
    got = orc.contour_integral(LAM, fr["rot"], fr["rot2"], v0, fr["prop2"], x3[:, 0], x3[:, 1],
                               x3[:, 2])
    for b in range(30):
        r = np.sqrt(x3[b, 0] * x3[b, 0] + x3[b, 1] * x3[b, 1] + x3[b, 2] * x3[b, 2])
        q = K * (fr["prop2"][:2] - x3[b, :2] / r)
        want = np.exp(1j * K * r) * 1j * ff_quadrature(q, p2) / LAM
        assert abs(got[b] - want) <= 1e-9 * max(abs(want), 1e-3 * tri_area(p2) / LAM)


@pytest.mark.parametrize("seed", range(5))
def test_ampl_to_mueller_is_kronecker_form(orc, seed):
    rng = np.random.default_rng(400 + seed)
    nb = 7
    J = rng.normal(size=(nb, 2, 2)) + 1j * rng.normal(size=(nb, 2, 2))
    # [r, c, theta=1, phi=nb]
    a = J.transpose(1, 2, 0)[:, :, None, :]
    M = orc.ampl_to_mueller(orc.complex_to_raw(a), nb)
    for b in range(nb):
        want = jones_to_mueller(J[b]).ravel()
        np.testing.assert_allclose(M[:, b], want, rtol=1e-13, atol=1e-13 * abs(want).max())
        # pure (non-depolarising) Mueller matrix identity: sum of squares = 4 M11^2
        S = M[:, b]
        assert np.sum(S ** 2) == pytest.approx(4 * S[0] ** 2, rel=1e-12)


@pytest.mark.parametrize("n", [3, 4, 7, 10])
def test_simpne_exact_on_quadratics(orc, n):
    rng = np.random.default_rng(n)
    x = np.sort(rng.uniform(0, 3, n))
    c = rng.normal(size=3)
    y = c[0] + c[1] * x + c[2] * x ** 2
    F = lambda t: c[0] * t + c[1] * t ** 2 / 2 + c[2] * t ** 3 / 3
    assert orc.simpne(x, y) == pytest.approx(F(x[-1]) - F(x[0]), rel=1e-12)


def test_simpne_rejects_two_points(orc):
    assert np.isnan(orc.simpne(np.array([0.0, 1.0]), np.array([1.0, 1.0])))


def _single_beam(v, prop, vk7, ampl=None):
    if ampl is None:
        ampl = np.r_[1, 0, 0, 0, 0, 0, 1, 0.]
    return BeamSet.from_triangles(np.asarray(ampl)[None], [vk7], [prop], [1.0], np.asarray(v)[None])


@pytest.mark.parametrize("tilt_deg", [0.0, 25.0, 50.0])
def test_forward_amplitude_is_obliquity_times_area(orc, tilt_deg):
    """At the bin along prop the aperture integral is the area A and the KW matrix
    is -K_z times a rotation, so |amplC|_F = |cos(tilt)| A sqrt(2) / lambda
    (Fraunhofer forward amplitude)."""
    s = 20.0
    v = np.array([[0, s / np.sqrt(3), 0], [-s / 2, -s / (2 * np.sqrt(3)), 0],
                  [s / 2, -s / (2 * np.sqrt(3)), 0]])
    t = np.deg2rad(tilt_deg)
    Ry = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    v = v @ Ry.T  # tilted aperture, centred at the origin
    theta = np.array([0.0, 2.0, 5.0]) * D2R
    phi = np.array([0.0, 90.0, 180.0, 270.0]) * D2R
    x, y, z = orc.make_far_field_bins(theta, phi)
    bs = _single_beam(v, (0, 0, -1.0), (1.0, 0, 0))
    a = orc.diffraction(theta, phi, x, y, z, LAM, bs, 0)
    norm = np.sqrt((abs(a[:, :, 0, :]) ** 2).sum(axis=(0, 1)))
    want = abs(np.cos(t)) * (s * s * np.sqrt(3) / 4) * np.sqrt(2) / LAM
    # the reference's split formula loses ~1e-7..1e-6 relative at q -> 0 (clamped k)
    np.testing.assert_allclose(norm, want, rtol=2e-6)


def test_forward_polarisation_is_phi_independent(orc):
    """Near the forward direction the diffracted field of a normally illuminated
    aperture is non-depolarising and, to O(theta^2), independent of the
    scattering-plane angle: the normalised Mueller matrix is diag(1, -1, -1, 1) in
    the reference's amplitude-basis convention at every phi.  A wrong
    scattering-plane rotation (e.g. the untransposed temp_rot1 of
    diff_mod.f90:646-650) modulates S22 and S33 with phi at O(1)."""
    s = 20.0
    v = np.array([[0, s / np.sqrt(3), 0], [-s / 2, -s / (2 * np.sqrt(3)), 0],
                  [s / 2, -s / (2 * np.sqrt(3)), 0]])
    theta = np.array([0.0, 0.5, 1.5, 2.0]) * D2R
    phi = np.arange(0, 360, 15) * D2R
    x, y, z = orc.make_far_field_bins(theta, phi)
    for vk7 in [(1.0, 0, 0), (0, 1.0, 0), (np.sqrt(0.5), np.sqrt(0.5), 0)]:
        bs = _single_beam(v, (0, 0, -1.0), vk7)
        a = orc.diffraction(theta, phi, x, y, z, LAM, bs, 0)
        nb = len(theta) * len(phi)
        M = orc.ampl_to_mueller(orc.complex_to_raw(a), nb).reshape(16, len(theta), len(phi))
        Mn = M[:, 1:] / M[0, 1:]
        want = np.diag([1.0, -1, -1, 1]).ravel()
        dev = np.abs(Mn - want[:, None, None]).max(axis=(0, 2))
        # vector (Karczewski-Wolf) corrections are second order in the scattering angle
        assert np.all(dev <= 1.1 * theta[1:] ** 2 + 1e-6), dev


def test_fast_diff_masks_outside_fov(orc):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(3, 3))
    prop = _unit(np.array([0.2, 0.1, -1.0]))
    vk7 = _unit(np.cross(prop, [1.0, 0, 0]))
    theta = np.linspace(0, np.pi, 19)
    phi = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    x, y, z = orc.make_far_field_bins(theta, phi)
    bs = BeamSet.from_triangles(np.r_[1, 0, 0, 0, 0, 0, 1, 0.][None], [vk7], [prop], [0.4], v[None])
    full = orc.diffraction(theta, phi, x, y, z, LAM, bs, 0)
    fast = orc.diffraction(theta, phi, x, y, z, LAM, bs, 0, fast_diff=True)
    X = np.stack([x, y, z], 1) - v.mean(0)
    inside = ((X / np.linalg.norm(X, axis=1)[:, None]) @ prop >= np.cos(0.4)).reshape(len(theta), len(phi))
    assert 0 < inside.sum() < inside.size
    np.testing.assert_array_equal(fast[:, :, ~inside], 0)
    np.testing.assert_array_equal(fast[:, :, inside], full[:, :, inside])


def test_diff_main_is_ordered_sum_and_thread_independent(orc):
    from abt_amd.workloads import random_triangles
    bs = random_triangles(9, seed=3, box=20.0)
    ext = random_triangles(4, seed=4, box=20.0)
    theta = np.linspace(0, np.pi, 13)
    phi = np.linspace(0, 2 * np.pi, 8, endpoint=False)
    b1, e1 = orc.diff_main(theta, phi, LAM, bs, ext, nthreads=1, raw=True)
    b4, e4 = orc.diff_main(theta, phi, LAM, bs, ext, nthreads=4, raw=True)
    np.testing.assert_array_equal(b1, b4)
    np.testing.assert_array_equal(e1, e4)
    x, y, z = orc.make_far_field_bins(theta, phi)
    s = 0
    for b in range(bs.n):
        s = s + orc.diffraction(theta, phi, x, y, z, LAM, bs, b)
    np.testing.assert_array_equal(orc.complex_to_raw(s), b1)


def _planted(tmp_path, old, new):
    """The oracle rebuilt from its source with one planted mistake, to show that a pin
    catches it (compiled like oracle/Makefile, into tmp_path)."""
    import ctypes
    import os
    import subprocess
    import oracle as O
    src_path = os.path.join(os.path.dirname(O.__file__), "diff_oracle.c")
    src = open(src_path).read()
    assert old in src
    p = tmp_path / "planted.c"
    p.write_text(src.replace(old, new))
    out = tmp_path / "libplanted.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
                    str(p), "-o", str(out), "-lm"], check=True)
    L = ctypes.CDLL(str(out))
    O._setup(L)
    return L


def _reversal_negates(orc):
    """Max over beams of |A(reversed vertex order) + A| / max|A| and whether the
    anti-parallel flag toggled for every beam."""
    from abt_amd.workloads import grid, random_triangles
    th, ph = grid((0, 10, 180), (0, 20, 340))
    x, y, z = orc.make_far_field_bins(th, ph)
    bs = random_triangles(12, seed=4, box=5.0)
    worst, toggles = 0.0, True
    for b in range(bs.n):
        v = bs.verts[bs.vert_off[b]:bs.vert_off[b + 1]]
        rv = BeamSet(bs.ampl[b:b + 1], bs.vk7[b:b + 1], bs.prop[b:b + 1], bs.fov[b:b + 1], [0, 3], v[::-1].copy())
        a0 = orc.diffraction(th, ph, x, y, z, LAM, bs, b)
        a1 = orc.diffraction(th, ph, x, y, z, LAM, rv, 0)
        f0 = orc.beam_frame(bs.ampl[b], bs.vk7[b], bs.prop[b], bs.fov[b], v)["anti_parallel"]
        f1 = orc.beam_frame(bs.ampl[b], bs.vk7[b], bs.prop[b], bs.fov[b], v[::-1].copy())["anti_parallel"]
        toggles &= f0 != f1
        worst = max(worst, np.abs(a0 + a1).max() / np.abs(a0).max())
    return worst, toggles


def test_vertex_order_reversal_negates_amplitude(orc, tmp_path, monkeypatch):
    """Reversing an aperture's vertex order flips the aperture frame (get_rotation_matrix2,
    diff_mod.f90:274-287) and toggles the anti-parallel sign flip of the amplitude matrix
    (diff_mod.f90:561-571); together they negate the diffracted amplitude at every bin.
    Without the sign flip the amplitude would be unchanged instead: planted in the oracle,
    that mistake fails this check."""
    worst, toggles = _reversal_negates(orc)
    assert toggles and worst <= 1e-8, worst
    L = _planted(tmp_path, "for (int q = 0; q < 8; q++) bf->ampl[q] = bf->anti_parallel ? -ampl[q] : ampl[q];",
                 "for (int q = 0; q < 8; q++) bf->ampl[q] = ampl[q];")
    import oracle as O
    monkeypatch.setattr(O, "_lib", L)
    worst, _ = _reversal_negates(O)
    assert worst > 1.0, worst


def _fallback_edge_jump(orc):
    """Largest second difference along theta at the edge of the forward-fallback cone
    (|cos theta| = 0.999, theta = 2.5625 deg), over that elsewhere in the band."""
    from abt_amd.workloads import random_triangles
    th_deg = np.r_[0.0, 0.5, 1.0, 1.5, 2.0, np.arange(2.40, 2.73, 0.02)]
    th = th_deg * D2R
    ph = np.array([0.0, 37.0, 123.0, 250.0]) * D2R
    x, y, z = orc.make_far_field_bins(th, ph)
    bs = random_triangles(8, seed=8, box=20.0)
    bs.prop[:] = _unit(np.array([0.05, -0.03, -1.0]))
    band = th_deg >= 2.39
    ib = np.searchsorted(th_deg[band], np.degrees(np.arccos(0.9990000128746033)))
    ratio = 0.0
    for b in range(bs.n):
        a = orc.diffraction(th, ph, x, y, z, LAM, bs, b)[:, :, band, :]
        d2 = np.diff(a, n=2, axis=2)
        edge = np.abs(d2[:, :, ib - 2:ib + 1]).max()
        other = np.abs(np.delete(d2, [ib - 2, ib - 1, ib], axis=2)).max()
        ratio = max(ratio, edge / other)
    return ratio


def test_forward_fallback_is_continuous(orc, tmp_path, monkeypatch):
    """Inside 2.56 deg of the incidence axis the scattering-plane vector hc comes from the bin
    (same phi, theta_i) (diff_mod.f90:630-635).  That bin lies in the same phi plane, so hc
    barely changes across the cone's edge and the far field stays smooth there: the second
    differences along theta that touch the edge are no larger than elsewhere.  A fallback to
    the wrong bin (phi index 0), planted in the oracle, breaks the smoothness."""
    assert _fallback_edge_jump(orc) <= 1.2
    L = _planted(tmp_path, "int64_t bt = i + nphi * theta_i;", "int64_t bt = nphi * theta_i;")
    import oracle as O
    monkeypatch.setattr(O, "_lib", L)
    assert _fallback_edge_jump(O) > 3.0
