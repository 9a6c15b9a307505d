"""abt_finalise_dev_static (the epilogue without allocation) through the C-ABI against
abt_finalise_dev, on two calls in a row, the second on a smaller grid, and its rejection of
grids beyond its scratch."""
import ctypes

import numpy as np
import pytest

from abt_amd import lib
from abt_amd.workloads import grid

pytestmark = pytest.mark.gpu


def _api():
    L = lib.load()
    vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    L.abt_finalise_dev_static.argtypes = [vp, vp, vp, i, d, d, vp, vp, vp, vp]
    L.abt_finalise_static_max_bins.argtypes = [ctypes.c_int64]
    L.abt_finalise_static_max_bins.restype = ctypes.c_int64
    return L


def _case(th, ph, seed):
    import torch
    dev = torch.device("cuda:0")
    g = lib.DeviceGrid(th, ph, 0.532)
    rng = np.random.default_rng(seed)
    ab = torch.from_numpy(rng.normal(size=8 * g.nbins)).to(dev)
    ae = torch.from_numpy(0.3 * rng.normal(size=8 * g.nbins)).to(dev)
    return g, ab, ae


def _run(g, ab, ae, static, scaling):
    import torch
    f64 = dict(dtype=torch.float64, device=ab.device)
    ab, ae = ab.clone(), ae.clone()
    m = torch.full((48 * g.nbins,), np.nan, **f64)
    m1 = torch.full((48 * len(g.theta),), np.nan, **f64)
    par = torch.full((16,), np.nan, **f64)
    if static:
        L = _api()
        s = torch.cuda.current_stream().cuda_stream
        rc = L.abt_finalise_dev_static(g.handle, ctypes.c_void_p(ab.data_ptr()), ctypes.c_void_p(ae.data_ptr()),
                                       int(scaling), 3.0, 1.5, ctypes.c_void_p(m.data_ptr()),
                                       ctypes.c_void_p(m1.data_ptr()), ctypes.c_void_p(par.data_ptr()),
                                       ctypes.c_void_p(s))
        assert rc == 0, L.abt_last_error()
    else:
        lib.finalise_dev(g, ab, ae, m, m1, par, scaling=scaling, beam_energy_out=3.0, ext_energy_out=1.5)
    torch.cuda.synchronize()
    return [x.cpu().numpy() for x in (ab, ae, m, m1, par[:9])]


@pytest.mark.parametrize("scaling", [False, True])
def test_static_finalise_matches_allocating_one(scaling):
    big = grid((0, 0.5, 180), (0, 2, 358))      # 361 x 180
    small = grid((0, 2, 180), (0, 6, 354))      # 91 x 60
    cases = [_case(*big, 1), _case(*small, 2)]
    for g, ab, ae in cases:                      # two calls in a row, the second smaller
        want = _run(g, ab, ae, False, scaling)
        got = _run(g, ab, ae, True, scaling)
        for w, x in zip(want, got):
            np.testing.assert_array_equal(x, w)


def test_static_finalise_rejects_grids_beyond_its_scratch():
    L = _api()
    nth = 1801
    nmax = L.abt_finalise_static_max_bins(nth)
    assert 1_000_000 < nmax < 1_100_000
    th, ph = grid((0, 0.1, 180), (0, 0.5, 359.5))   # 1801 x 720 = 1,296,720 bins
    g, ab, ae = _case(th, ph, 3)
    assert g.nbins > nmax
    import torch
    out = torch.empty(48 * g.nbins, dtype=torch.float64, device=ab.device)
    rc = L.abt_finalise_dev_static(g.handle, ctypes.c_void_p(ab.data_ptr()), ctypes.c_void_p(ae.data_ptr()), 0,
                                   1.0, 1.0, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), None)
    assert rc == 1 and b"scratch" in L.abt_last_error()
