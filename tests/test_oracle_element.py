"""Pins of the CPU oracle's element layer against what the mathematics fixes.

hex8 shape functions (partition of unity, Kronecker property), B/G/F under affine fields,
element volumes, the patch test, and Ke = dFe/du by central differences for every
hyperelastic model (this also pins the spatial tangents and the geometric stiffness).
"""
import numpy as np
import pytest

import oracle_lib as ol
from ceos_amd import meshgen

SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def _gauss():
    nat, w = np.zeros(24), np.zeros(8)
    ol.lib().or_hex8_gauss(nat, w)
    return nat.reshape(8, 3), w


def test_gauss_rule():
    nat, w = _gauss()
    assert np.allclose(np.abs(nat), 1.0 / np.sqrt(3.0), rtol=0, atol=1e-16)
    assert np.array_equal(np.sign(nat), SIGNS)
    assert w.sum() == 8.0


def test_shape_functions_partition_and_kronecker():
    L = ol.lib()
    rng = np.random.default_rng(1)
    for p in rng.uniform(-1, 1, size=(20, 3)):
        N, dN = np.zeros(8), np.zeros(24)
        L.or_hex8_shape(np.ascontiguousarray(p), N)
        L.or_hex8_dshape(np.ascontiguousarray(p), dN)
        assert abs(N.sum() - 1.0) < 4e-16
        assert np.allclose(dN.reshape(8, 3).sum(axis=0), 0.0, atol=1e-15)
        # derivative by central difference
        h = 1e-6
        for k in range(3):
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            Np, Nm = np.zeros(8), np.zeros(8)
            L.or_hex8_shape(pp, Np)
            L.or_hex8_shape(pm, Nm)
            assert np.allclose((Np - Nm) / (2 * h), dN.reshape(8, 3)[:, k], atol=1e-9)
    for a in range(8):
        N = np.zeros(8)
        L.or_hex8_shape(np.ascontiguousarray(SIGNS[a]), N)
        assert np.array_equal(N, np.eye(8)[a])


def _random_hex(rng, scale=0.15):
    X = 0.5 * (SIGNS + 1.0) + rng.uniform(-scale, scale, size=(8, 3))
    return np.ascontiguousarray(X)


def test_bg_and_F_exact_for_affine_fields():
    L = ol.lib()
    rng = np.random.default_rng(2)
    nat, _ = _gauss()
    for _ in range(5):
        X = _random_hex(rng)
        A = rng.normal(scale=0.1, size=(3, 3))
        u = X @ A.T  # u = A X  -> grad u = A, F = I + A
        Ue = np.ascontiguousarray(u.reshape(24))
        for g in range(8):
            F = np.zeros(9)
            assert L.or_deformation_gradient(X.reshape(24), np.ascontiguousarray(nat[g]), Ue, F) == 0
            assert np.allclose(F.reshape(3, 3), np.eye(3) + A, atol=1e-13)
            B, G = np.zeros(144), np.zeros(216)
            import ctypes
            dj = ctypes.c_double()
            assert L.or_matrix_bg(X.reshape(24), np.ascontiguousarray(nat[g]), B, G, ctypes.byref(dj)) == 0
            B, G = B.reshape(6, 24), G.reshape(9, 24)
            grad = (G @ Ue).reshape(3, 3)
            assert np.allclose(grad, A, atol=1e-13)
            eps = B @ Ue
            ref = [A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1], A[0, 2] + A[2, 0]]
            assert np.allclose(eps, ref, atol=1e-13)


def test_detj_guard():
    import ctypes
    L = ol.lib()
    X = 0.5 * (SIGNS + 1.0)
    X[[0, 1]] = X[[1, 0]]  # swap two nodes: inverted element
    B, G = np.zeros(144), np.zeros(216)
    dj = ctypes.c_double()
    nat, _ = _gauss()
    st = L.or_matrix_bg(np.ascontiguousarray(X.reshape(24)), np.ascontiguousarray(nat[0]), B, G, ctypes.byref(dj))
    assert st == -1


def test_perturbed_cube_volume_is_exact():
    X, conn = meshgen.cube(4, perturb=0.1, seed=7)
    m = ol.OracleMesh(X, conn)
    m.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
    assert m.solve_constitutive(0.0, np.zeros(3 * len(X))) == 0
    vx = sum(m.element_volume(e)[1] for e in range(len(conn)))
    assert abs(vx - 1.0) < 1e-13


def test_patch_test_uniform_stress_zero_interior_residual():
    """Affine displacement on a distorted mesh -> uniform F and sigma -> Fint = 0 at interior
    nodes (every interior node's contributions cancel)."""
    X, conn = meshgen.cube(3, perturb=0.15, seed=3)
    m = ol.OracleMesh(X, conn)
    m.set_material(1, ol.MODEL_CNH, [1.0, 10.0])
    A = np.array([[0.05, 0.02, 0.0], [0.01, -0.03, 0.02], [0.0, 0.01, 0.04]])
    U = (X @ A.T).reshape(-1)
    m.update_mesh(U)
    assert m.solve_constitutive(0.0, U) == 0
    Fint, st = m.internal_force()
    assert st == 0
    s = m.gp_stress()
    assert np.allclose(s, s[0, 0], atol=1e-12)
    n = 4
    interior = [(i + n * (j + n * k)) for k in range(1, 3) for j in range(1, 3) for i in range(1, 3)]
    f = Fint.reshape(-1, 3)
    assert np.max(np.abs(f[interior])) < 1e-13 * np.max(np.abs(f))


@pytest.mark.parametrize("model,props", [
    (ol.MODEL_CNH, [1.0, 99.3333333333]),
    (ol.MODEL_NH, [0.5, 100.0]),
    (ol.MODEL_SVK, [10.0, 0.3]),
    (ol.MODEL_NH_Q1P0, [0.5, 100.0]),
])
def test_ke_is_derivative_of_fe(model, props):
    """Ke = dFe/du (updated Lagrangian: x = X + u, F from X and u) by central differences.
    For the mean-dilatation model this checks the Bdiv term: Fe carries the pressure of
    Jbar = V/V0, dV/du = Bdiv, so dFe/du gains d2Psi/dJbar2 / V0 Bdiv^T Bdiv, and a dropped or
    mis-scaled term fails."""
    X, conn = meshgen.cube(1, perturb=0.0)
    rng = np.random.default_rng(11)
    X = X + rng.uniform(-0.1, 0.1, size=X.shape)
    m = ol.OracleMesh(X, conn)
    m.set_material(1, model, props)
    U0 = rng.normal(scale=0.05, size=24)

    def fe(U):
        m.update_mesh(U)
        assert m.solve_constitutive(0.0, U) == 0
        return m.element_internal_force(0)

    fe(U0)
    Ke = m.element_stiffness(0)
    h = 1e-6
    Kfd = np.zeros((24, 24))
    gdof = (3 * (conn[0][:, None] - 1) + np.arange(3)[None, :]).reshape(24)  # local -> global
    for j in range(24):
        dp, dm = U0.copy(), U0.copy()
        dp[gdof[j]] += h
        dm[gdof[j]] -= h
        Kfd[:, j] = (fe(dp) - fe(dm)) / (2 * h)
    assert np.linalg.norm(Ke - Kfd) / np.linalg.norm(Ke) < 1e-7
    assert np.linalg.norm(Ke - Ke.T) / np.linalg.norm(Ke) < 1e-13
    # rigid translations are in the null space
    for k in range(3):
        t = np.zeros(24)
        t[k::3] = 1.0
        assert np.linalg.norm(Ke @ t) < 1e-10 * np.linalg.norm(Ke)


def test_linear_ke_matches_derivative_for_j2_elastic():
    X, conn = meshgen.cube(1, perturb=0.0)
    m = ol.OracleMesh(X, conn, nl_analysis=False)
    m.set_material(1, ol.MODEL_J2, [210000.0, 0.3, 1e30, 1000.0])
    rng = np.random.default_rng(5)
    U0 = rng.normal(scale=1e-3, size=24)

    def fe(U):
        assert m.solve_constitutive(0.0, U) == 0
        return m.element_internal_force(0)

    fe(U0)
    Ke = m.element_stiffness(0)
    h = 1e-7
    gdof = (3 * (conn[0][:, None] - 1) + np.arange(3)[None, :]).reshape(24)
    Kfd = np.column_stack([(fe(U0 + h * np.eye(24)[gdof[j]]) - fe(U0 - h * np.eye(24)[gdof[j]])) / (2 * h)
                           for j in range(24)])
    assert np.linalg.norm(Ke - Kfd) / np.linalg.norm(Ke) < 1e-8


@pytest.mark.parametrize("model,props,nl", [
    (ol.MODEL_CNH, [1.0, 99.3333333333], True),
    (ol.MODEL_NH, [0.5, 100.0], True),
    (ol.MODEL_SVK, [10.0, 0.3], True),
])
def test_numerical_tangent_matches_analytic(model, props, nl):
    """GetTangentModulusBase (central differences of P over F, pulled back to C and pushed
    forward) reproduces the analytic spatial tangent of every hyperelastic model: the element
    stiffnesses agree.  (It is a finite-strain construction: for the small-strain J2 model it
    differs by O(strain), and J2 supplies its own tangent in the reference.)"""
    X, conn = meshgen.cube(2, perturb=0.1, seed=5)
    rng = np.random.default_rng(6)
    U = rng.normal(scale=0.03 if nl else 0.001, size=3 * len(X))
    Ks = []
    for numerical in (False, True):
        m = ol.OracleMesh(X, conn, nl_analysis=nl)
        m.set_material(1, model, props)
        m.set_tangent_mode(numerical)
        if nl:
            m.update_mesh(U)
        assert m.solve_constitutive(0.0, U) == 0
        Ks.append(m.tangent())
    assert np.linalg.norm(Ks[1] - Ks[0]) / np.linalg.norm(Ks[0]) < 1e-6
