"""The FE^2 oracle pieces against what the mathematics fixes (CPU only).

  * pull_back inverts the push forward: a Neo-Hookean point's (sigma, D) pulled back equals its
    closed-form S and dS/dE, and central differences of S(E);
  * rve_to_ext: P and A = dP/dF of a homogeneous hyperelastic point (central differences of the
    oracle's P(F)) give back the same S and dS/dE;
  * the surrogate's analytic tangent equals central differences of its S(E) (FP64 network);
  * the surrogate is objective (S depends on U only) and exact to first order at U = I.
"""
import numpy as np
import pytest

import fe2_oracle as fo
import mlp_oracle as mo
import oracle_lib as ol

NH = [0.5, 50.0]


def _F(seed, scale=0.15):
    return np.eye(3) + np.random.default_rng(seed).uniform(-scale, scale, size=(3, 3))


def _nh_material(F):
    s, D, _ = ol.material_point(ol.MODEL_NH, NH, F)
    return s, D


def _S_of_E(E):
    """Neo-Hookean S at the symmetric F = (2E + I)^(1/2)."""
    lam2, Q = np.linalg.eigh(2.0 * E + np.eye(3))
    F = Q @ np.diag(np.sqrt(lam2)) @ Q.T
    s, D = _nh_material(F)
    return fo.pull_back(s, D, F)[0]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pull_back_of_neo_hookean(seed):
    F = _F(seed)
    s, D = _nh_material(F)
    S, C = fo.pull_back(s, D, F)
    # closed form of model 3: S = 2 C10 J^(-2/3) (I - tr C / 3 C^-1) + J p C^-1
    Cg = F.T @ F
    J = np.linalg.det(F)
    Ci = np.linalg.inv(Cg)
    p = 3.0 * NH[1] * J ** (-2.0 / 3.0) * (J ** (1.0 / 3.0) - 1.0)
    S_cf = 2 * NH[0] * J ** (-2.0 / 3.0) * (np.eye(3) - np.trace(Cg) / 3.0 * Ci) + J * p * Ci
    assert np.allclose(S, fo.to_voigt(S_cf), rtol=1e-12, atol=1e-12)
    # dS/dE by central differences (objectivity: S depends on C = 2E + I only)
    E0 = 0.5 * (Cg - np.eye(3))
    h = 1e-6
    Cfd = np.zeros((6, 6))
    for b, (K, L) in enumerate(fo.VOIGT):
        dE = np.zeros((3, 3))
        dE[K, L] += 0.5 * h
        dE[L, K] += 0.5 * h
        Cfd[:, b] = (_S_of_E(E0 + dE) - _S_of_E(E0 - dE)) / (2 * h)
    assert np.allclose(C, Cfd, rtol=0, atol=1e-6 * np.abs(C).max())


@pytest.mark.parametrize("seed", [3, 4])
def test_rve_to_ext_of_a_homogeneous_point(seed):
    F = _F(seed)
    s, D = _nh_material(F)
    P = fo.first_piola(s, F)
    h = 1e-6
    A = np.zeros((9, 9))
    for k in range(3):
        for L in range(3):
            dF = np.zeros((3, 3))
            dF[k, L] = h
            dP = fo.first_piola(_nh_material(F + dF)[0], F + dF) - fo.first_piola(_nh_material(F - dF)[0], F - dF)
            for i in range(3):
                for J in range(3):
                    A[3 * J + i, 3 * L + k] = dP[i, J] / (2 * h)
    S, C = fo.rve_to_ext(F, P, A)
    S0, C0 = fo.pull_back(s, D, F)
    assert np.allclose(S, S0, rtol=1e-12, atol=1e-12)
    assert np.allclose(C, C0, rtol=0, atol=1e-6 * np.abs(C0).max())
    # a dropped geometric term (delta_ik S_JL) would fail this by |S| ~ 1e-1
    A4 = np.einsum("Ii,Kk,iJkL->IJKL", np.linalg.inv(F), np.linalg.inv(F),
                   A.reshape(3, 3, 3, 3).transpose(1, 0, 3, 2))
    assert np.abs(fo.tensor4_to_voigt(A4) - C0).max() > 1e-3


@pytest.mark.parametrize("seed", [12, 13])
def test_nh_material_is_the_pulled_back_model(seed):
    F = _F(seed)
    S, C = fo.nh_material(NH[0], NH[1], F)
    s, D = _nh_material(F)
    S0, C0 = fo.pull_back(s, D, F)
    assert np.allclose(S, S0, rtol=1e-12, atol=1e-12 * np.abs(S0).max())
    assert np.allclose(C, C0, rtol=1e-10, atol=1e-10 * np.abs(C0).max())


def test_base_linearizes_to_c_lin():
    rng = np.random.default_rng(14)
    A = rng.normal(size=(6, 6))
    Cl = fo.iso_stiffness(40.0, 2.0) + 0.3 * (A + A.T)
    Clin = fo.up21(Cl)
    S0, C0 = fo.base(Clin, np.eye(3))
    assert np.allclose(S0, 0.0, atol=1e-12) and np.allclose(C0, Cl, atol=1e-10)
    # Voigt averages of an isotropic stiffness are its moduli
    assert np.allclose(fo.iso_moduli(fo.up21(fo.iso_stiffness(40.0, 2.0))), (40.0, 2.0))
    # the base of an isotropic C_lin is the Neo-Hookean with those moduli
    F = _F(15)
    Sb, _ = fo.base(fo.up21(fo.iso_stiffness(NH[1], 2 * NH[0])), F)
    assert np.allclose(Sb, fo.nh_material(NH[0], NH[1], F)[0], rtol=1e-12, atol=1e-12)


def _random_surrogate(seed, width=32):
    rng = np.random.default_rng(seed)
    _, n = mo.layout(27, width, 6)
    p = rng.normal(scale=0.3, size=n)
    cfg = {"u_scale": 0.25, "lam_lo": 0.8, "lam_hi": 1.25, "e_ref": 0.1,
           "d_mean": rng.normal(size=21), "d_std": rng.uniform(0.5, 2, 21),
           "d_lo": -np.full(21, 3.0), "d_hi": np.full(21, 3.0),
           "t_mean": rng.normal(size=6), "t_std": rng.uniform(0.5, 2, 6)}
    Clin = fo.up21(np.diag(rng.uniform(1, 2, 6)) + 0.05 * np.ones((6, 6)))
    return p, cfg, Clin, width


@pytest.mark.parametrize("seed", [5, 6, 7])
def test_surrogate_tangent_is_the_derivative_of_its_stress(seed):
    p, cfg, Clin, w = _random_surrogate(seed)
    F = _F(seed + 100, 0.1)
    S, C, _ = fo.surrogate(p, cfg, F, Clin, w)
    Cg = F.T @ F
    E0 = 0.5 * (Cg - np.eye(3))

    def S_of_E(E):
        lam2, Q = np.linalg.eigh(2.0 * E + np.eye(3))
        return fo.surrogate_stress(p, cfg, Q @ np.diag(np.sqrt(lam2)) @ Q.T, Clin, w)[0]

    assert np.allclose(S_of_E(E0), S, rtol=1e-12, atol=1e-12)  # S depends on U only
    h = 1e-6
    Cfd = np.zeros((6, 6))
    for b, (K, L) in enumerate(fo.VOIGT):
        dE = np.zeros((3, 3))
        dE[K, L] += 0.5 * h
        dE[L, K] += 0.5 * h
        Cfd[:, b] = (S_of_E(E0 + dE) - S_of_E(E0 - dE)) / (2 * h)
    assert np.allclose(C, 0.5 * (Cfd + Cfd.T), rtol=0, atol=1e-6 * np.abs(C).max())


def test_surrogate_is_objective_and_linear_at_identity():
    p, cfg, Clin, w = _random_surrogate(8)
    F = _F(9, 0.1)
    q, _ = np.linalg.qr(np.random.default_rng(10).normal(size=(3, 3)))
    q *= np.sign(np.linalg.det(q))
    S1, C1, _ = fo.surrogate(p, cfg, F, Clin, w)
    S2, C2, _ = fo.surrogate(p, cfg, q @ F, Clin, w)
    assert np.allclose(S1, S2, rtol=1e-12, atol=1e-12) and np.allclose(C1, C2, rtol=1e-10, atol=1e-10)
    # at U = I the tangent is C_lin and S = 0; near it S = C_lin : E + O(|E|^2)
    S0, C0, inside = fo.surrogate(p, cfg, np.eye(3), Clin, w)
    assert inside and np.allclose(S0, 0.0) and np.allclose(C0, fo.from_up21(Clin), atol=1e-10)
    eps = 1e-4
    F = np.eye(3) + eps * np.diag([1.0, -0.5, 0.2])
    S, _, _ = fo.surrogate(p, cfg, F, Clin, w)
    E = 0.5 * (F.T @ F - np.eye(3))
    lin = fo.from_up21(Clin) @ (fo.to_voigt(E) * fo.ENG)
    assert np.abs(S - lin).max() < 1e3 * eps ** 2 * (1 + np.abs(cfg["t_mean"] + 3 * cfg["t_std"]).max() / cfg["e_ref"] ** 2)


def test_hull_flag():
    p, cfg, _, w = _random_surrogate(11)
    Clin = np.asarray(cfg["d_mean"]) + 0.5 * np.asarray(cfg["d_std"])  # normalized descriptor 0.5
    assert fo.features(cfg, np.diag([1.2, 0.85, 1.0]), Clin)[1]
    assert not fo.features(cfg, np.diag([1.3, 0.85, 1.0]), Clin)[1]
    assert not fo.features(cfg, np.diag([1.0, 0.75, 1.0]), Clin)[1]
    far = np.asarray(cfg["d_mean"]) + 4.0 * np.asarray(cfg["d_std"])
    assert not fo.features(cfg, np.eye(3), far)[1]

