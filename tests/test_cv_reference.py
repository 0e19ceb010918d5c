"""CPU checks that pin the learned control variate's mathematics (no GPU):
the numpy gradient against central differences, the unbiasedness of the
control variate and the meaning of the training loss, both exactly over all
Rademacher sign vectors."""

import itertools

import numpy as np
import pytest

import cv_ref


@pytest.mark.parametrize("npos", [None, 1, 0])
def test_gradient_matches_finite_differences(npos):
    n, H, r, T = 3, 4, 2, 5
    rng = np.random.default_rng(0)
    flat = rng.normal(size=H * n + H + H * H + H + n * r * H + n * r) * 0.7
    X, V, HV = rng.normal(size=(T, n)), rng.normal(size=(T, n)), rng.normal(size=(T, n))
    loss, g = cv_ref.loss_grad(flat, X, V, HV, n, H, r, bf=False, gain=1.3, npos=npos)
    h = 1e-6
    fd = np.empty_like(flat)
    for k in range(flat.size):
        e = np.zeros_like(flat)
        e[k] = h
        lp, _ = cv_ref.loss_grad(flat + e, X, V, HV, n, H, r, bf=False, gain=1.3, npos=npos)
        lm, _ = cv_ref.loss_grad(flat - e, X, V, HV, n, H, r, bf=False, gain=1.3, npos=npos)
        fd[k] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_control_variate_has_mean_zero_over_all_rademacher_vectors():
    rng = np.random.default_rng(1)
    n, r = 5, 3
    L = rng.normal(size=(n, r))
    vs = np.array(list(itertools.product([-1.0, 1.0], repeat=n)))
    for s in (np.ones(r), np.array([1.0, -1.0, -1.0])):
        cv = np.sum(s * (vs @ L) ** 2, axis=1) - np.sum(s * L * L)
        assert abs(cv.mean()) <= 1e-12


def test_loss_is_frobenius_distance_over_all_rademacher_vectors():
    rng = np.random.default_rng(2)
    n, r = 4, 2
    A = rng.normal(size=(n, n))
    Hs = A + A.T
    L = rng.normal(size=(n, r))
    vs = np.array(list(itertools.product([-1.0, 1.0], repeat=n)))
    e = vs @ Hs.T - (vs @ L) @ L.T
    lhs = np.mean(np.sum(e * e, axis=1))
    assert abs(lhs - np.sum((Hs - L @ L.T) ** 2)) <= 1e-10 * lhs
    # and the control variate's residual v^T (H - L L^T) v has mean tr(H - L L^T)
    res = np.einsum("ti,ij,tj->t", vs, Hs - L @ L.T, vs)
    assert abs(res.mean() - np.trace(Hs - L @ L.T)) <= 1e-10


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 1e-3], dtype=np.float32)
    got = cv_ref.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to the even mantissa (1.0)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -6 and got[3] == 1.0
    assert got[4] == -2.5 and abs(got[5] - 1e-3) <= 2 ** -9 * 1e-3
