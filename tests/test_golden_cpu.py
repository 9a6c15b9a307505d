"""The CPU oracle reproduces the committed fixtures (tests/golden)."""
import numpy as np

import oracle as O
from tests.golden import load_small


def test_oracle_reproduces_golden_small():
    d, b, e = load_small()
    th, ph, lam = d["theta"], d["phi"], float(d["lam"])
    rb, re = O.diff_main(th, ph, lam, b, e, raw=True)
    np.testing.assert_array_equal(rb, d["ampl_far_beam"])
    np.testing.assert_array_equal(re, d["ampl_far_ext_diff"])
    fin = O.finalise(th, ph, lam, rb, re)
    np.testing.assert_array_equal(fin["mueller"], d["mueller"])
    np.testing.assert_array_equal(fin["mueller_1d"], d["mueller_1d"])


def test_oracle_reproduces_golden_visibility():
    from tests.golden import load_vis
    for name, g, fin, bz, res in load_vis():
        for mode in ("int", "ext", "inc"):
            r = O.find_vis(mode, g, None if mode == "inc" else fin, bz)
            for k, want in res[mode].items():
                np.testing.assert_array_equal(r[k], want, err_msg=f"{name} {mode} {k}")
