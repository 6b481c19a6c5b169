"""Pins of the CPU oracle's material layer: closed forms at special states, known answers of
the textbook J2 radial return (Simo & Hughes, Computational Inelasticity, Box 3.1, linear
isotropic hardening) and finite-difference checks of the consistent tangent."""
import numpy as np
import pytest

import oracle_lib as ol

ISYM = np.diag([1, 1, 1, 0.5, 0.5, 0.5])
I1 = np.zeros((6, 6))
I1[:3, :3] = 1.0


def test_cnh_at_identity():
    mu, lam = 1.0, 99.3333333333
    s, D, _ = ol.material_point(ol.MODEL_CNH, [mu, lam], np.eye(3))
    assert np.all(s == 0.0)
    assert np.allclose(D, lam * I1 + 2 * mu * ISYM, rtol=0, atol=1e-13)


def test_cnh_closed_form_uniaxial():
    mu, lam, l = 1.0, 99.0, 1.2
    F = np.diag([l, 1.0, 1.0])
    s, D, _ = ol.material_point(ol.MODEL_CNH, [mu, lam], F)
    J = l
    assert abs(s[0] - ((mu / J) * (l * l - 1) + lam / J * np.log(J))) < 1e-12
    assert abs(s[1] - lam / J * np.log(J)) < 1e-12


def test_nh_equals_cnh_tangent_at_identity():
    """neo_hookean(C10=mu/2, K) and compressible_neo_hookean(mu, K-2mu/3) have the same
    initial tangent (linear elasticity with shear modulus mu and bulk modulus K)."""
    mu, K = 1.0, 100.0
    _, D_nh, _ = ol.material_point(ol.MODEL_NH, [mu / 2, K], np.eye(3))
    _, D_cnh, _ = ol.material_point(ol.MODEL_CNH, [mu, K - 2 * mu / 3], np.eye(3))
    assert np.allclose(np.triu(D_nh), np.triu(D_cnh), atol=1e-12)


def test_svk_small_strain_limit():
    E, nu = 200.0, 0.25
    _, D, _ = ol.material_point(ol.MODEL_SVK, [E, nu], np.eye(3))
    _, De, _ = ol.material_point(ol.MODEL_GHL, [E, nu], np.eye(3))
    assert np.allclose(np.triu(D), np.triu(De), atol=1e-12)


def test_hyperelastic_rigid_rotation():
    """sigma(R F) = R sigma(F) R^T for every objective hyperelastic model."""
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    R = q * np.sign(np.linalg.det(q))
    F = np.eye(3) + rng.normal(scale=0.1, size=(3, 3))
    for model, p in [(ol.MODEL_CNH, [1.0, 50.0]), (ol.MODEL_NH, [0.5, 50.0]), (ol.MODEL_SVK, [10.0, 0.3])]:
        s1, _, _ = ol.material_point(model, p, F)
        s2, _, _ = ol.material_point(model, p, R @ F)
        T1 = np.array([[s1[0], s1[3], s1[5]], [s1[3], s1[1], s1[4]], [s1[5], s1[4], s1[2]]])
        T2 = np.array([[s2[0], s2[3], s2[5]], [s2[3], s2[1], s2[4]], [s2[5], s2[4], s2[2]]])
        assert np.allclose(R @ T1 @ R.T, T2, atol=1e-12 * np.abs(T1).max())


J2P = [210000.0, 0.3, 250.0, 1000.0]


def _j2_KG(p):
    E, nu = p[0], p[1]
    return E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))


def test_j2_uniaxial_strain_known_answer():
    K, G = _j2_KG(J2P)
    sy, H = J2P[2], J2P[3]
    ey = sy / (2 * G)
    # elastic
    e = 0.5 * ey
    s, _, st = ol.material_point(ol.MODEL_J2, J2P, np.diag([1 + e, 1, 1]), state_old=np.zeros(17))
    assert abs(s[0] - (K + 4 * G / 3) * e) < 1e-9 and st[15] == 0.0
    # plastic, one step from the virgin state
    e = 0.02
    s, _, st = ol.material_point(ol.MODEL_J2, J2P, np.diag([1 + e, 1, 1]), state_old=np.zeros(17))
    qt = 2 * G * e
    dg = (qt - sy) / (3 * G + H)
    s11_dev = (1 - 3 * G * dg / qt) * (4 * G * e / 3)
    assert st[15] == 1.0
    assert abs(st[14] - dg) < 1e-15
    assert abs(s[0] - (s11_dev + K * e)) < 1e-9
    assert abs(st[6] - dg) < 1e-15
    # the returned stress lies on the updated yield surface
    sd = s[:3] - s[:3].mean()
    q = np.sqrt(1.5 * (np.sum(sd ** 2) + 2 * np.sum(s[3:] ** 2)))
    assert abs(q - (sy + H * dg)) < 1e-9
    # plastic strain is deviatoric
    assert abs(st[0] + st[1] + st[2]) < 1e-15


def test_j2_elastic_unloading_after_commit():
    K, G = _j2_KG(J2P)
    _, _, st = ol.material_point(ol.MODEL_J2, J2P, np.diag([1.02, 1, 1]), state_old=np.zeros(17))
    old = np.zeros(17)
    old[:7] = st[:7]
    s, _, st2 = ol.material_point(ol.MODEL_J2, J2P, np.diag([1.019, 1, 1]), state_old=old)
    assert st2[15] == 0.0 and np.array_equal(st2[:7], old[:7])


@pytest.mark.parametrize("e11,shear", [(0.0005, 0.0), (0.02, 0.0), (0.015, 0.01), (-0.03, 0.004)])
def test_j2_consistent_tangent_fd(e11, shear):
    old = np.zeros(17)
    old[:6] = [0.001, -0.0004, -0.0006, 0.0002, 0.0, -0.0003]
    old[6] = 0.0015
    F0 = np.eye(3) + np.array([[e11, shear, 0.0], [0.0, -0.3 * e11, 0.5 * shear], [0.2 * shear, 0.0, 0.001]])
    s0, D, st = ol.material_point(ol.MODEL_J2, J2P, F0, state_old=old)
    pert = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    h = 1e-8
    Dfd = np.zeros((6, 6))
    for J, (a, b) in enumerate(pert):
        Fp, Fm = F0.copy(), F0.copy()
        Fp[a, b] += h
        Fm[a, b] -= h
        sp, _, stp = ol.material_point(ol.MODEL_J2, J2P, Fp, state_old=old)
        sm, _, stm = ol.material_point(ol.MODEL_J2, J2P, Fm, state_old=old)
        assert stp[15] == stm[15] == st[15]
        Dfd[:, J] = (sp - sm) / (2 * h)
    assert np.linalg.norm(D - Dfd) / np.linalg.norm(D) < 1e-6


VISCO = [50.0, 1.0, 10.0, 20.0, 2.0, 5.0]  # K_inf, Mu_inf, Lambda_inf, K_e, Mu_e, Ni_v


def _virgin_visco_state():
    st = np.zeros(20)
    st[[0, 4, 8]] = 1.0  # Fv = I, dv = 0, gamma = 0, time 0
    return st


def test_visco_without_viscous_branch_is_compressible_neo_hookean():
    """Mu_e = 0: S = Mu_inf (I - C^-1) + Lambda_inf lnJ C^-1, the compressible Neo-Hookean law."""
    p = list(VISCO)
    p[4] = 0.0
    F = np.eye(3) + np.array([[0.1, 0.03, 0.0], [0.02, -0.05, 0.01], [0.0, 0.04, 0.07]])
    s, D, _ = ol.material_point(ol.MODEL_VISCO, p, F, time=0.3, state_old=_virgin_visco_state())
    sc, Dc, _ = ol.material_point(ol.MODEL_CNH, [p[1], p[2]], F)
    assert np.allclose(s, sc, rtol=1e-13, atol=1e-13)
    assert np.allclose(np.triu(D), np.triu(Dc), rtol=1e-11, atol=1e-11)


def test_visco_instantaneous_response():
    """dt = 0 from the virgin state: no viscous flow, the response is compressible Neo-Hookean
    with shear modulus Mu_inf + Mu_e."""
    F = np.eye(3) + np.array([[0.08, 0.02, 0.0], [0.01, -0.04, 0.02], [0.03, 0.0, 0.05]])
    s, _, st = ol.material_point(ol.MODEL_VISCO, VISCO, F, time=0.0, state_old=_virgin_visco_state())
    sc, _, _ = ol.material_point(ol.MODEL_CNH, [VISCO[1] + VISCO[4], VISCO[2]], F)
    assert np.allclose(s, sc, rtol=1e-12, atol=1e-12)
    assert np.allclose(st[:9], [1, 0, 0, 0, 1, 0, 0, 0, 1], atol=1e-14)


def test_visco_relaxes_towards_equilibrium():
    """Held isochoric deformation, growing time step: the stress decays from the instantaneous
    value towards the long-term branch (Mu_inf) and stays between the two bounds (viscous flow
    is isochoric, Jv = 1, so a volume-preserving F relaxes completely)."""
    F = np.diag([1.1, 1.1 ** -0.5, 1.1 ** -0.5])
    s_inst, _, _ = ol.material_point(ol.MODEL_CNH, [VISCO[1] + VISCO[4], VISCO[2]], F)
    s_inf, _, _ = ol.material_point(ol.MODEL_CNH, [VISCO[1], VISCO[2]], F)
    prev = s_inst[0]
    for dt in (0.01, 0.1, 1.0, 10.0, 100.0):
        s, _, _ = ol.material_point(ol.MODEL_VISCO, VISCO, F, time=dt, state_old=_virgin_visco_state())
        assert s_inf[0] - 1e-9 <= s[0] <= prev + 1e-12
        prev = s[0]
    assert abs(prev - s_inf[0]) < 0.05 * abs(s_inst[0] - s_inf[0])


@pytest.mark.parametrize("dt", [0.1, 0.5, 2.0])
def test_visco_small_strain_relaxation_rate(dt):
    """Small-strain limit of the local problem (ModViscoelasticMatrix.f90:452-623): with
    Fv = (I - dt dv)^-1 Fv_old, Me = Ce Se = Mu_e (Ce - I) and the residual
    -dt sym(FvFo^T Me) + dt Ni_v dv + gamma dt Jv FvFo = 0, Jv = 1, a small held strain gives
    the backward-Euler Maxwell branch Ni_v dv = 2 Mu_e dev(eps - eps_v): under a held shear the
    overstress (stress minus the long-term branch) shrinks by exactly
    r = 1 / (1 + 2 Mu_e dt / Ni_v) per step.  The deviation is second order in the strain
    (1e-5 at 1e-2, 1e-7 at 1e-3).  A wrong role of Ni_v, dt or the factor 2 moves r by 10-60 %."""
    g = 1e-3
    F = np.eye(3)
    F[0, 1] = F[1, 0] = g
    s_inf = ol.material_point(ol.MODEL_CNH, [VISCO[1], VISCO[2]], F)[0][3]
    tau0 = ol.material_point(ol.MODEL_CNH, [VISCO[1] + VISCO[4], VISCO[2]], F)[0][3] - s_inf
    r = 1.0 / (1.0 + 2.0 * VISCO[4] * dt / VISCO[5])
    st = _virgin_visco_state()
    for n in range(1, 6):
        s, _, st = ol.material_point(ol.MODEL_VISCO, VISCO, F, time=n * dt, state_old=st)
        st = st.copy()
        assert abs((s[3] - s_inf) / (tau0 * r ** n) - 1.0) < 1e-5


@pytest.mark.parametrize("lam", [1.05, 0.9])
def test_visco_pure_dilation_has_no_viscous_flow(lam):
    """F = lam I: Ce = lam^2 I, so the residual vanishes at dv = 0 with the multiplier
    gamma = Mu_e (lam^2 - 1) (the pressure that keeps the viscous flow isochoric): Fv stays I
    and the stress stays the instantaneous one (Neo-Hookean with Mu_inf + Mu_e) at any time step,
    at finite strain."""
    F = lam * np.eye(3)
    s0 = ol.material_point(ol.MODEL_CNH, [VISCO[1] + VISCO[4], VISCO[2]], F)[0]
    st = _virgin_visco_state()
    for n in range(1, 4):
        s, _, st = ol.material_point(ol.MODEL_VISCO, VISCO, F, time=0.5 * n, state_old=st)
        st = st.copy()
        assert np.allclose(s, s0, rtol=1e-12, atol=1e-12)
        assert np.allclose(st[:9], [1, 0, 0, 0, 1, 0, 0, 0, 1], rtol=0, atol=1e-14)
        assert st[18] == pytest.approx(VISCO[4] * (lam * lam - 1.0), rel=1e-12)


def test_visco_element_tangent_is_consistent():
    """Ke = dFe/du (finite differences) after a committed viscous step; the algorithmic tangent
    assumes an exactly converged local Newton (tolerance 1e-5), hence the 1e-6 bar."""
    from ceos_amd import meshgen
    X, conn = meshgen.cube(1)
    rng = np.random.default_rng(11)
    X = X + rng.uniform(-0.1, 0.1, size=X.shape)
    m = ol.OracleMesh(X, conn)
    m.set_material(1, ol.MODEL_VISCO, VISCO)
    gdof = (3 * (conn[0][:, None] - 1) + np.arange(3)[None, :]).reshape(24)
    U1 = rng.normal(scale=0.03, size=24)
    m.update_mesh(U1)
    assert m.solve_constitutive(0.1, U1) == 0
    m.switch_converged_state()
    U0 = U1 + rng.normal(scale=0.03, size=24)

    def fe(U):
        m.update_mesh(U)
        assert m.solve_constitutive(0.25, U) == 0
        return m.element_internal_force(0)

    fe(U0)
    Ke = m.element_stiffness(0)
    h = 1e-6
    Kfd = np.zeros((24, 24))
    for j in range(24):
        dp, dm = U0.copy(), U0.copy()
        dp[gdof[j]] += h
        dm[gdof[j]] -= h
        Kfd[:, j] = (fe(dp) - fe(dm)) / (2 * h)
    assert np.linalg.norm(Ke - Kfd) / np.linalg.norm(Ke) < 1e-6
    st = m.state(1)
    assert np.all(st[0, :, 19] == 0.25) and np.any(np.abs(st[0, :, 18]) > 0)
