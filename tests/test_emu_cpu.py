"""CPU checks of the kernel bodies (abt_amd/csrc/ffbody.h) through the host
emulator: the FP64 body reproduces the oracle and the mixed-precision body
meets the same criteria as on the GPU.  No GPU needed."""
import numpy as np
import pytest

import oracle as O
from abt_amd import lib
from abt_amd.beams import BeamSet
from abt_amd.workloads import grid, random_triangles, stress_beams
from tests import emu
from tests.gpu_util import ampl_rel_err, summary
from tests.test_gpu_parity import (LAM, _check_fp64, _coarse, _column_case, _forward_mask,
                                   check_mixed)


def _run(prec, th, ph, bs, fov=None):
    recs, vr = lib.prepare_beams_host(bs, LAM, fov_value=fov)
    return emu.far_field(prec, th, ph, LAM, recs, bs.n, vr)


def test_emu_fp64_body_matches_oracle(orc):
    """The plain body against the oracle: the kernels' bound (1e-9 per bin, floored), and 5e-11
    here, where the host's reciprocals are the IEEE ones.  The plain body keeps the reference's
    operations wherever their rounding shows (the phases and the cancelling edge terms), and
    reaches 1.7e-11 on this case; summing the edge terms unrotated and rotating once per beam, a
    reassociation that leaves every term unchanged, gives 1.1e-10."""
    th, ph = _coarse()
    b, e = _column_case()
    rb, re = orc.diff_main(th, ph, LAM, b, e, nthreads=8)
    for got, ref, bs in ((_run(0, th, ph, b), rb, b), (_run(0, th, ph, e, O.PI_R4 / 2), re, e)):
        fwd = _forward_mask(th, ph, [bs])
        _check_fp64(got, ref, fwd)
        err = ampl_rel_err(got, ref)
        assert err[~fwd].max() <= 5e-11, summary(err[~fwd])


def test_table_sincos_within_half_ulp():
    """sincos_tab64, the plain body's sine and cosine, against x87 extended precision over the
    ranges the body sees (edge phases to ~1e4 rad, the far-field phase k|x - c| ~ 1.2e7 rad at
    0.532 um) and beyond, where the table index passes 2^31 and wraps in the low bits of the
    shifted sum (4e10 rad): within 6e-17 absolute, i.e. about half an ulp.  A table without its
    low parts errs by up to 1.7e-16, a one-constant reduction by ~5e-10 at 1.2e7 rad."""
    rng = np.random.default_rng(3)
    for span in (10.0, 3e3, 1.3e7, 4e10):
        x = rng.uniform(-span, span, 200_000)
        s, c = emu.sincos_tab(x)
        xl = x.astype(np.longdouble)
        assert float(np.abs(s - np.sin(xl)).max()) <= 6e-17, span
        assert float(np.abs(c - np.cos(xl)).max()) <= 6e-17, span


@pytest.mark.parametrize("lam", [0.2, 1.5])
def test_emu_fp64_body_other_wavelengths(orc, lam):
    """The FP64 bodies against the oracle at a short wavelength (k |x - c| ~ 3e7 rad, table
    indices past 2^31) and a long one.  Plain: within 1e-9 of the floored amplitude on every bin
    (measured 1.4e-12 and 1.6e-11).  Restructured: within 1e-7 off the exact forward directions
    (measured 2.0e-8 and 7.2e-9)."""
    th, ph = _coarse()
    b, _ = _column_case()
    ref, _ = orc.diff_main(th, ph, lam, b, BeamSet.empty(), nthreads=8)
    recs, vr = lib.prepare_beams_host(b, lam)
    got = emu.far_field(0, th, ph, lam, recs, b.n, vr)
    err = ampl_rel_err(got, ref)
    assert err.max() <= 1e-9, summary(err)
    got = emu.far_field(4, th, ph, lam, recs, b.n, vr)
    err = ampl_rel_err(got, ref)
    fwd = _forward_mask(th, ph, [b], lam=lam)
    assert err[~fwd].max() <= 1e-7, summary(err[~fwd])


def test_emu_exact_forward_with_clamped_k(orc):
    """The plain body on the reference's clamped exact-forward bins (on_axis_triangles): within
    1e-9 of the oracle (measured 3e-16), where the exact form factor differs by ~1e-5."""
    from tests.test_gpu_parity import on_axis_triangles
    th = np.radians([0.0, 0.5, 1.0, 2.0])
    ph = np.radians(np.arange(0.0, 360.0, 30.0))
    bs = on_axis_triangles()
    ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=4)
    recs, vr = lib.prepare_beams_host(bs, LAM)
    err = ampl_rel_err(emu.far_field(0, th, ph, LAM, recs, bs.n, vr), ref)
    assert err.max() <= 1e-9, summary(err)
    assert ampl_rel_err(emu.far_field(4, th, ph, LAM, recs, bs.n, vr), ref)[0].max() > 1e-7


def test_emu_inside_the_ky_clamp(orc):
    """The plain body on bins inside the reference's k_y clamp (ky_clamp_case): within 1e-9 of
    the oracle (measured 1.7e-11)."""
    from tests.test_gpu_parity import ky_clamp_case
    th, ph, bs, inside = ky_clamp_case()
    assert inside > 100
    ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=4)
    recs, vr = lib.prepare_beams_host(bs, LAM)
    err = ampl_rel_err(emu.far_field(0, th, ph, LAM, recs, bs.n, vr), ref)
    assert err.max() <= 1e-9, summary(err)


def test_emu_clamped_edge_slopes(orc):
    """The plain body with edges parallel to the aperture frame's axes (axis_edge_triangles):
    within 1e-9 of the oracle (measured 1.4e-11)."""
    from tests.test_gpu_parity import axis_edge_triangles
    th, ph = _coarse()
    bs = axis_edge_triangles()
    ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=4)
    recs, vr = lib.prepare_beams_host(bs, LAM)
    err = ampl_rel_err(emu.far_field(0, th, ph, LAM, recs, bs.n, vr), ref)
    assert err.max() <= 1e-9, summary(err)


def test_markstein_quotient_equals_division():
    """div_by, the plain body's quotient by a shared divisor (q = a y, q + (a - b q) y with
    y = 1/b correctly rounded), is bit-equal to IEEE division: on 10^6 pairs over 26 decades,
    a quarter of the divisors with significands next to all ones.  The data can tell: the bare
    product a y differs from the quotient on about a quarter of them."""
    rng = np.random.default_rng(4)
    n = 1_000_000
    a = np.exp(rng.uniform(-30, 30, n)) * rng.choice([-1.0, 1.0], n)
    b = np.exp(rng.uniform(-30, 30, n)) * rng.choice([-1.0, 1.0], n)
    k = n // 4
    b[:k] = np.ldexp(2.0 - np.ldexp(rng.integers(1, 8, k).astype(float), -52), rng.integers(-20, 20, k))
    q = emu.div_by(a, b)
    assert np.array_equal(q, a / b)
    assert np.count_nonzero(a * (1.0 / b) != a / b) > n // 10


def test_emu_mixed_body_meets_criteria(orc):
    th, ph = _coarse()
    b, e = _column_case()
    rb, re = orc.diff_main(th, ph, LAM, b, e, nthreads=8)
    check_mixed(_run(1, th, ph, b), rb, _forward_mask(th, ph, [b]))
    check_mixed(_run(1, th, ph, e, O.PI_R4 / 2), re, _forward_mask(th, ph, [e]))
    th, ph = grid((0, 2, 180), (0, 6, 354))
    for bs in (stress_beams(150, seed=7), random_triangles(200, seed=4, box=40.0)):
        ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=8)
        check_mixed(_run(1, th, ph, bs), ref, _forward_mask(th, ph, [bs]))


def test_emu_mixed_quadrilaterals(orc):
    from tests.test_gpu_parity import test_fp64_quadrilateral_apertures  # noqa: F401
    th, ph = _coarse()
    rng = np.random.default_rng(8)
    sets = []
    for i in range(5):
        c = rng.uniform(-5, 5, 3)
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        a = np.cross(n, rng.normal(size=3)); a /= np.linalg.norm(a)
        bb = np.cross(n, a)
        ang = np.sort(rng.uniform(0, 2 * np.pi, 5))
        rad = rng.uniform(1, 3, 5)
        verts = c + (rad * np.cos(ang))[:, None] * a + (rad * np.sin(ang))[:, None] * bb
        prop = n + 0.4 * a; prop /= np.linalg.norm(prop)
        vk7 = np.cross(prop, a); vk7 /= np.linalg.norm(vk7)
        sets.append(BeamSet(rng.normal(size=(1, 8)), vk7, prop, [1.0], [0, 5], verts))
    from abt_amd.beams import concat
    bs = concat(sets)
    ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=8)
    got = _run(1, th, ph, bs)
    err = ampl_rel_err(got, ref)
    assert err.max() <= 2e-6, summary(err)


def test_emu_mixed_reflections_near_forward(orc):
    """External reflections: the forward direction is the mirror image of every beam's
    prop in its aperture plane, where each beam's Karczewski-Wolf amplitude vanishes
    linearly in theta.  The mixed body keeps FP32 relative precision there: amplitude
    error <= 1e-6 per bin for theta 0.05-2.5 deg (a cancelling FP32 form of diff_ampl
    gives ~6e-8 / theta: 3e-5 at 0.1 deg)."""
    from abt_amd.workloads import reflected_facets
    th = np.radians([0.05, 0.1, 0.2, 0.5, 1.0, 2.0, 2.5])
    ph = np.radians(np.arange(0.0, 360.0, 30.0))
    bs = reflected_facets(200, seed=3)
    ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=8)
    err = ampl_rel_err(_run(1, th, ph, bs), ref)
    assert err.max() <= 1e-6, summary(err)


def _fp64r_check(got, ref, fwd, model, tol=1e-7):
    """The restructured FP64 body against the oracle: per bin to tol of the bin (floored at
    1e-3 of the largest), the oracle's own split-formula rounding setting the level.  Along a
    beam's exact forward direction (|q| a < 1e-3) the oracle loses ~1e-16 / (|q| a)^2; there
    the body is checked against the frame-free FP64 model (adjbody.h; emulator precision 3),
    an independent FP64 evaluation without that cancellation, to 1e-8."""
    err = ampl_rel_err(got, ref)
    assert err[~fwd].max() <= tol, summary(err[~fwd])
    if fwd.any():
        err_m = ampl_rel_err(got, model)
        assert err_m[fwd].max() <= 1e-8, summary(err_m[fwd])


def test_emu_fp64r_body_matches_oracle(orc):
    th, ph = _coarse()
    b, e = _column_case()
    rb, re = orc.diff_main(th, ph, LAM, b, e, nthreads=8)
    _fp64r_check(_run(4, th, ph, b), rb, _forward_mask(th, ph, [b]), _run(3, th, ph, b))
    _fp64r_check(_run(4, th, ph, e, O.PI_R4 / 2), re, _forward_mask(th, ph, [e]), _run(3, th, ph, e, O.PI_R4 / 2))
    th, ph = grid((0, 2, 180), (0, 6, 354))
    for bs in (stress_beams(150, seed=7), random_triangles(200, seed=4, box=40.0)):
        ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=8)
        _fp64r_check(_run(4, th, ph, bs), ref, _forward_mask(th, ph, [bs]), _run(3, th, ph, bs))


def test_emu_fp64r_edges_normal_to_q(orc):
    """Bins where an edge is (nearly) normal to q, so its two vertex phases agree to < 1e-3
    revolution: the difference form of the edge gives way to the product form there.  The
    triangles have an edge along y in the aperture plane and the bins sweep through the
    direction whose q is along x."""
    from tests.formfactor import ff_quadrature  # noqa: F401
    th = np.radians(np.linspace(20.0, 40.0, 81))
    ph = np.radians([0.0, 0.05, 90.0, 180.0, 270.0])
    bs = random_triangles(60, seed=12, box=10.0)
    ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=8)
    _fp64r_check(_run(4, th, ph, bs), ref, _forward_mask(th, ph, [bs]), _run(3, th, ph, bs))


def test_emu_fp64r_polygons(orc):
    th, ph = _coarse()
    rng = np.random.default_rng(8)
    sets = []
    for i in range(5):
        c = rng.uniform(-5, 5, 3)
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        a = np.cross(n, rng.normal(size=3)); a /= np.linalg.norm(a)
        bb = np.cross(n, a)
        ang = np.sort(rng.uniform(0, 2 * np.pi, 5))
        rad = rng.uniform(1, 3, 5)
        verts = c + (rad * np.cos(ang))[:, None] * a + (rad * np.sin(ang))[:, None] * bb
        prop = n + 0.4 * a; prop /= np.linalg.norm(prop)
        vk7 = np.cross(prop, a); vk7 /= np.linalg.norm(vk7)
        sets.append(BeamSet(rng.normal(size=(1, 8)), vk7, prop, [1.0], [0, 5], verts))
    from abt_amd.beams import concat
    bs = concat(sets)
    ref, _ = orc.diff_main(th, ph, LAM, bs, BeamSet.empty(), nthreads=8)
    err = ampl_rel_err(_run(4, th, ph, bs), ref)   # the frame-free model takes triangles only
    assert err.max() <= 1e-7, summary(err)
