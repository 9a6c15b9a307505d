"""Shared helpers of the GPU parity tests: error metrics against the oracle."""
import numpy as np


def ampl_rel_err(got, ref, floor=1e-3):
    """Per-bin amplitude error: |A - A_ref|_F / max(|A_ref|_F, floor * max_bins |A_ref|_F).
    got/ref: complex [r, c, theta, phi]."""
    d = np.sqrt((abs(got - ref) ** 2).sum(axis=(0, 1)))
    n = np.sqrt((abs(ref) ** 2).sum(axis=(0, 1)))
    return d / np.maximum(n, floor * n.max())


def mueller_of(orc, a):
    nth, nph = a.shape[2], a.shape[3]
    return orc.ampl_to_mueller(orc.complex_to_raw(a), nth * nph)


def mueller_errors(M, Mref):
    """Per-element errors, M: (16, nbins).
    rel_floor: |dS_ij| / max(|S_ij_ref|, 1e-3 |S11_ref|)   (per-element relative, floored)
    norm:      |dS_ij| / |S11_ref|                          (normalised-element absolute)"""
    s11 = np.abs(Mref[0])
    d = np.abs(M - Mref)
    rel_floor = d / np.maximum(np.abs(Mref), 1e-3 * s11)
    norm = d / s11
    return rel_floor, norm


def summary(x):
    x = np.asarray(x).ravel()
    return dict(max=float(x.max()), p999=float(np.quantile(x, 0.999)), median=float(np.median(x)))
