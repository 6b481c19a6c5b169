"""Pins of the periodic RVE restatement (tests/rve_oracle.py) by what the mechanics fixes:
  * a homogeneous RVE under any F solves with zero fluctuation in one evaluation and gives
    P = P_material(F) (the patch test of the periodic system);
  * the solved displacement is periodic-compatible: U(Y+) - U(Y-) = (F - I)(Y+ - Y-);
  * the homogenized F equals the prescribed F (the fluctuation is periodic);
  * the condensed tangent equals central differences of P(F) (GetTangentModulusBase);
  * frame indifference: P(R F) = R P(F), so S = F^-1 P depends on U = sqrt(F^T F) only;
  * the seeded generator is reproducible and hits the volume fraction.
"""
import numpy as np
import pytest

import oracle_lib as ol
import rve_oracle as ro
from ceos_amd import rve


def _F(seed, scale=0.08):
    rng = np.random.default_rng(seed)
    return np.eye(3) + rng.uniform(-scale, scale, size=(3, 3))


def test_generator_reproducible_and_fraction():
    a = rve.phases(12, 5)
    b = rve.phases(12, 5)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, rve.phases(12, 6))
    cen, rad = rve.spheres(5)
    assert abs(4.0 / 3.0 * np.pi * np.sum(rad ** 3) - rve.VOLUME_FRACTION) < 1e-9
    for i in range(len(rad)):
        for j in range(i):
            d = cen[i] - cen[j]
            d -= np.round(d)
            assert np.linalg.norm(d) >= rad[i] + rad[j] - 1e-12
    fr = np.mean([rve.phases(16, i).mean() for i in range(8)])
    assert abs(fr - 0.2) < 0.03
    s = rve.sides(1000)
    assert set(np.unique(s)) == {8, 12, 16}


def test_homogeneous_rve_is_affine():
    n = 4
    o = ro.OracleRVE(n, np.zeros(n ** 3, dtype=np.uint8))
    F = _F(1)
    P, Fh, it, hist = o.solve(F)
    assert it == 0 and np.all(o.w == 0.0)
    s, _, _ = ol.material_point(ol.MODEL_NH, list(rve.MATRIX_PROPS), F)
    sig = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    Pm = np.linalg.det(F) * sig @ np.linalg.inv(F).T
    assert np.allclose(P.reshape(3, 3).T, Pm, rtol=0, atol=1e-12 * np.abs(Pm).max())
    assert np.allclose(Fh.reshape(3, 3).T, F, atol=1e-14)


def test_heterogeneous_rve_periodicity_and_average_F():
    n = 6
    ph = rve.phases(n, 3)
    assert 0 < ph.sum() < n ** 3
    o = ro.OracleRVE(n, ph)
    F = _F(2)
    P, Fh, it, hist = o.solve(F)
    assert hist[-1] <= 1e-10 * hist[0]
    # Newton converges quadratically at the end
    assert hist[-1] < 1e-3 * hist[-2] or hist[-1] < 1e-13
    U = ((F - np.eye(3)) @ o.X.T).T.ravel() + o.T @ o.w
    U = U.reshape(-1, 3)
    X = o.X
    for ax in range(3):
        lo = np.nonzero(np.abs(X[:, ax]) < 1e-12)[0]
        for a in lo:
            Xb = X[a].copy()
            Xb[ax] = 1.0
            b = np.nonzero(np.all(np.abs(X - Xb) < 1e-12, axis=1))[0][0]
            assert np.allclose(U[b] - U[a], (F - np.eye(3))[:, ax], atol=1e-13)
    assert np.allclose(Fh.reshape(3, 3).T, F, atol=1e-13)
    # a genuinely heterogeneous answer
    s, _, _ = ol.material_point(ol.MODEL_NH, list(rve.MATRIX_PROPS), F)
    assert np.linalg.norm(o.w) > 1e-4
    # the same state reached in two macro increments
    o2 = ro.OracleRVE(n, ph)
    o2.solve(np.eye(3) + 0.5 * (F - np.eye(3)))
    P2, _, _, _ = o2.solve(F)
    assert np.linalg.norm(P2 - P) < 1e-9 * np.linalg.norm(P)


def test_large_stretch_needs_cut_back():
    """Principal stretches at the edge of the surrogate's range: Newton from the identity
    fails and the golden-ratio substepping reaches the state."""
    n = 4
    o = ro.OracleRVE(n, rve.phases(n, 7))
    F = np.diag([1.25, 0.8, 1.1])
    P, Fh, it, hist = o.solve(F)
    assert o.cutbacks >= 1
    assert np.allclose(Fh.reshape(3, 3).T, F, atol=1e-13)


def test_condensed_tangent_matches_finite_differences():
    n = 4
    o = ro.OracleRVE(n, rve.phases(n, 11))
    F = _F(3, 0.05)
    o.solve(F, rtol=1e-13)
    A = o.tangent()
    Afd = o.tangent_fd(F, h=1e-6)
    assert np.linalg.norm(A - Afd) / np.linalg.norm(A) < 1e-7


def test_rve_response_is_objective():
    n = 4
    ph = rve.phases(n, 2)
    F = _F(4)
    # a 0.4 rad rotation about a seeded axis: the load path I + alpha (R F - I) stays invertible
    rng = np.random.default_rng(9)
    a = rng.normal(size=3)
    a /= np.linalg.norm(a)
    Kx = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + np.sin(0.4) * Kx + (1 - np.cos(0.4)) * Kx @ Kx
    o1, o2 = ro.OracleRVE(n, ph), ro.OracleRVE(n, ph)
    P1, F1, _, _ = o1.solve(F, rtol=1e-13)
    P2, F2, _, _ = o2.solve(R @ F, rtol=1e-13)
    assert np.allclose(P2.reshape(3, 3).T, R @ P1.reshape(3, 3).T, atol=1e-10)
    assert np.allclose(ro.second_piola(P1, F1), ro.second_piola(P2, F2), atol=1e-10)
