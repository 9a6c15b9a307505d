"""The reference-facing host entry points as PBT's -memory mode uses them: `diffraction`
called per beam from several host threads at once (beam_loop_mod.f90:37-51, :1935-1967),
on one set of bins, against the CPU oracle; and the grid cache behind the host entry
points (rebuilt only when the angles, wavelength or bin positions change)."""
import threading

import numpy as np
import pytest

from abt_amd import lib
from abt_amd.workloads import grid, random_triangles
from tests.test_gpu_parity import LAM, _check_fp64, _forward_mask

pytestmark = pytest.mark.gpu


def test_diffraction_from_concurrent_threads(orc):
    th, ph = grid((0, 2, 180), (0, 6, 354))
    x, y, z = orc.make_far_field_bins(th, ph)
    b = random_triangles(16, seed=31, box=12.0)
    b.fov[:] = 0.5
    refs = [orc.diffraction(th, ph, x, y, z, LAM, b, i, fast_diff=bool(i % 2)) for i in range(b.n)]
    got = [None] * b.n
    errors = []

    def work(t):
        try:
            for rep in range(3):   # every thread diffracts its beams three times
                for i in range(t, b.n, 4):
                    got[i] = lib.diffraction(th, ph, x, y, z, LAM, b, i, precision="fp64",
                                             fast_diff=bool(i % 2))
        except Exception as exc:   # surfaced below, on the test's thread
            errors.append(exc)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i in range(b.n):
        _check_fp64(got[i], refs[i], _forward_mask(th, ph, [b.subset([i])]))


def test_grid_cache_follows_inputs(orc):
    """A cached grid is reused only for identical inputs: another wavelength, or other bin
    positions for `diffraction`, give the oracle's answer for those inputs."""
    th, ph = grid((0, 5, 180), (0, 10, 350))
    b = random_triangles(4, seed=33, box=6.0)
    x, y, z = orc.make_far_field_bins(th, ph)
    for lam in (LAM, 0.633, LAM):
        ref = orc.diffraction(th, ph, x, y, z, lam, b, 1)
        got = lib.diffraction(th, ph, x, y, z, lam, b, 1, precision="fp64")
        _check_fp64(got, ref, _forward_mask(th, ph, [b.subset([1])]))
    # the same angles with bins at another distance (diffraction takes the caller's bins)
    x2, y2, z2 = 0.5 * x, 0.5 * y, 0.5 * z
    ref = orc.diffraction(th, ph, x2, y2, z2, LAM, b, 2)
    got = lib.diffraction(th, ph, x2, y2, z2, LAM, b, 2, precision="fp64")
    assert np.abs(ref - orc.diffraction(th, ph, x, y, z, LAM, b, 2)).max() > 1e-3 * np.abs(ref).max()
    _check_fp64(got, ref, _forward_mask(th, ph, [b.subset([2])]))
    # more distinct grids than the cache holds
    for k in range(10):
        phk = ph + 1e-3 * k
        xk, yk, zk = orc.make_far_field_bins(th, phk)
        ref = orc.diffraction(th, phk, xk, yk, zk, LAM, b, 0)
        got = lib.diffraction(th, phk, xk, yk, zk, LAM, b, 0, precision="fp64")
        _check_fp64(got, ref, _forward_mask(th, phk, [b.subset([0])]))


@pytest.mark.parametrize("precision", ["fp64", "fp64_fast", "mixed"])
def test_diff_main_one_tree_per_call(precision):
    """abt_diff_main with one output NULL computes only the other tree (the Fortran drop-in
    calls it once per tree, to write the reference's per-tree timing lines): each output equals
    the two-tree call's bit for bit."""
    th, ph = grid((0, 5, 180), (0, 10, 350))
    b = random_triangles(20, seed=41, box=10.0)
    e = random_triangles(6, seed=42, box=10.0)
    gb, ge = lib.diff_main(th, ph, LAM, b, e, precision=precision)
    ob, none = lib.diff_main(th, ph, LAM, b, e, precision=precision, trees=("beam",))
    assert none is None
    np.testing.assert_array_equal(ob, gb)
    none, oe = lib.diff_main(th, ph, LAM, b, e, precision=precision, trees=("ext",))
    assert none is None
    np.testing.assert_array_equal(oe, ge)


def test_diff_main_on_beams_loaded_from_pbt_export(orc):
    """The outgoing beams read from a file in the layout of PBT's -export_beam
    (tests/golden/beam_export.json) through diff_main, with and without -fast_diff (which
    reads the per-beam field of view the reader forms), against the oracle."""
    import os
    from abt_amd.beams import BeamSet, outbeams_from_pbt_json
    from tests.golden.make_beam_json import case
    path = os.path.join(os.path.dirname(__file__), "golden", "beam_export.json")
    V, F, _, _ = case()
    b = outbeams_from_pbt_json(path, V, F, LAM)
    th, ph = grid((0, 2, 180), (0, 6, 354))
    refs = []
    for fast in (False, True):
        rb, _ = orc.diff_main(th, ph, LAM, b, BeamSet.empty(), fast_diff=fast, nthreads=8)
        gb, ge = lib.diff_main(th, ph, LAM, b, BeamSet.empty(), precision="fp64", fast_diff=fast)
        _check_fp64(gb, rb, _forward_mask(th, ph, [b]))
        assert np.all(ge == 0)
        refs.append(rb)
    assert np.abs(refs[0] - refs[1]).max() > 1e-3 * np.abs(refs[0]).max()   # the fields of view cull
