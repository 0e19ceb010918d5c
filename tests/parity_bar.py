"""The f64 parity bar shared by the GPU tests.

An f64 GPU result r matches the oracle's o when both are NaN, or both are
the same infinity, or |r - o| <= TOL * max(|o|, M), where M is the sum of
the absolute values of the terms behind o (pyoracle.magnitudes: running
error analysis for values, absolute-value reverse sweeps for gradients and
H v, and their products with |v| for jet coefficients).  TOL = 1e-12: the
GPU and the oracle evaluate the same operations in the same order, so they
differ only by the last bits of the transcendental functions, propagated
through those terms.  A wrong rule, index or sign is off by about one
term, far above TOL * M.
"""

import numpy as np

TOL = 1e-12


def f64_mismatch(got, ref, mag, tol=TOL):
    """Indices and worst ratio |got - ref| / (tol * max(|ref|, mag)) over
    the finite entries; NaN and infinities must coincide exactly."""
    got = np.asarray(got, np.float64).reshape(np.shape(ref))
    ref = np.asarray(ref, np.float64)
    mag = np.broadcast_to(np.asarray(mag, np.float64), ref.shape)
    nan = np.isnan(ref)
    bad = nan != np.isnan(got)
    inf = np.isinf(ref)
    bad |= inf & (got != ref)
    fin = ~nan & ~inf
    bad |= fin & ~np.isfinite(got)
    scale = np.maximum(np.abs(ref), np.where(np.isfinite(mag), mag, np.abs(ref)))
    ratio = np.zeros(ref.shape)
    ok = fin & np.isfinite(got)
    ratio[ok] = np.abs(got[ok] - ref[ok]) / np.maximum(tol * scale[ok], 1e-300)
    bad |= ratio > 1.0
    return np.argwhere(bad), float(ratio.max(initial=0.0))


def assert_f64(got, ref, mag, what="", tol=TOL):
    idx, worst = f64_mismatch(got, ref, mag, tol)
    assert len(idx) == 0, f"{what}: {len(idx)} entries off the bar (worst ratio {worst:.3g}), first {idx[:5].tolist()}"
    return worst
