"""FP64 numpy restatement of the FE^2 macro-point pieces (test infrastructure, checker only).

  pull_back(sigma, D, F)     spatial (Cauchy, Voigt tangent) -> material (S, dS/dE): the inverse
                             of the push forward the element kernel applies to external points
                             (Push_Forward_Voigt, Tools/MathTools/ModMathRoutines.f90:254-311)
  rve_to_ext(F, P, A)        S = F^-1 P, dS/dE from A = dP/dF
                             (First_Elasticity_Modulus_To_Second_Voigt, ModMathRoutines.f90:315)
  nh_material(C10, K, F)     model 3 (Neo-Hookean) S and dS/dE in the reference configuration
  base(Clin, F)              the surrogate's base: Neo-Hookean with the Voigt-average moduli of
                             C_lin plus (C_lin - C_iso) : E
  surrogate(p, cfg, F, Clin) the surrogate's S, dS/dE and hull flag:
                             S = base + |E|^2/e_ref^2 (t_mean + t_std MLP(x)),
                             x = [(U - I)/u_scale, (C_lin - d_mean)/d_std]
The pieces are pinned in tests/test_fe2_oracle.py by what the mathematics fixes: the pull back of
a pushed-forward tangent, central differences of P(F) and of S(E).
"""
import numpy as np

import mlp_oracle as mo

VOIGT = ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2))
ENG = np.array([1, 1, 1, 2, 2, 2.0])


def up21(M):
    return np.array([M[..., a, b] for a in range(6) for b in range(a, 6)]).T


def from_up21(c):
    c = np.asarray(c)
    out = np.zeros(c.shape[:-1] + (6, 6))
    k = 0
    for a in range(6):
        for b in range(a, 6):
            out[..., a, b] = out[..., b, a] = c[..., k]
            k += 1
    return out


def voigt_to_tensor4(D):
    """6x6 Voigt (engineering shear) -> c_ijkl with minor symmetries."""
    c = np.zeros((3, 3, 3, 3))
    for a, (i, j) in enumerate(VOIGT):
        for b, (k, l) in enumerate(VOIGT):
            for (p, q) in {(i, j), (j, i)}:
                for (r, s) in {(k, l), (l, k)}:
                    c[p, q, r, s] = D[a, b]
    return c


def tensor4_to_voigt(c):
    D = np.zeros((6, 6))
    for a, (i, j) in enumerate(VOIGT):
        for b, (k, l) in enumerate(VOIGT):
            D[a, b] = 0.25 * (c[i, j, k, l] + c[j, i, k, l] + c[i, j, l, k] + c[j, i, l, k])
    return D


def sym(T):
    return 0.5 * (T + T.T)


def to_voigt(T):
    return np.array([T[i, j] for i, j in VOIGT])


def from_voigt(v):
    return np.array([[v[0], v[3], v[5]], [v[3], v[1], v[4]], [v[5], v[4], v[2]]])


def pull_back(sigma6, D, F):
    """Cauchy stress + spatial tangent (Voigt) -> second Piola stress + material tangent."""
    F = np.asarray(F)
    J = np.linalg.det(F)
    Fi = np.linalg.inv(F)
    S = J * Fi @ from_voigt(sigma6) @ Fi.T
    C = J * np.einsum("Ii,Jj,Kk,Ll,ijkl->IJKL", Fi, Fi, Fi, Fi, voigt_to_tensor4(D))
    return to_voigt(S), tensor4_to_voigt(C)


def first_piola(sigma6, F):
    return np.linalg.det(F) * from_voigt(sigma6) @ np.linalg.inv(F).T


def rve_to_ext(F, P, A):
    """F, P (3, 3); A (9, 9) with A[3J + i, 3L + k] = dP_iJ/dF_kL.  Returns S (6), dS/dE (6, 6)."""
    Fi = np.linalg.inv(F)
    S = sym(Fi @ P)
    A4 = np.zeros((3, 3, 3, 3))  # A4[i, J, k, L]
    for i in range(3):
        for J in range(3):
            for k in range(3):
                for L in range(3):
                    A4[i, J, k, L] = A[3 * J + i, 3 * L + k]
    C = np.einsum("Ii,Kk,iJkL->IJKL", Fi, Fi, A4) - np.einsum("Ii,Ki,JL->IJKL", Fi, Fi, S)
    Cv = tensor4_to_voigt(C)
    return to_voigt(S), 0.5 * (Cv + Cv.T)


def iso_moduli(Clin):
    """Voigt-average bulk and shear moduli of a 6x6 stiffness (upper 21 given)."""
    C = from_up21(Clin)
    K = (C[0, 0] + C[1, 1] + C[2, 2] + 2 * (C[0, 1] + C[0, 2] + C[1, 2])) / 9.0
    mu = ((C[0, 0] + C[1, 1] + C[2, 2]) - (C[0, 1] + C[0, 2] + C[1, 2]) + 3 * (C[3, 3] + C[4, 4] + C[5, 5])) / 15.0
    return K, mu


def iso_stiffness(K, mu):
    lam = K - 2.0 * mu / 3.0
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def nh_material(C10, K, F):
    """Model 3 Neo-Hookean second Piola stress and tangent dS/dE (Voigt 6x6) at F, from the
    closed form S = 2 C10 J^-2/3 (I - tr C/3 C^-1) + J p C^-1, p = 3K J^-2/3 (J^1/3 - 1) and its
    derivative d/dE = 2 d/dC (dC^-1/dC = -C^-1 (x) C^-1 sym, dJ/dC = J C^-1 / 2)."""
    Cg = F.T @ F
    J = np.linalg.det(F)
    Ci = np.linalg.inv(Cg)
    I = np.eye(3)
    trC = np.trace(Cg)
    p = 3.0 * K * J ** (-2.0 / 3.0) * (J ** (1.0 / 3.0) - 1.0)
    S = 2 * C10 * J ** (-2.0 / 3.0) * (I - trC / 3.0 * Ci) + J * p * Ci
    # dS/dC (4th order), then dS/dE = 2 dS/dC
    dCi = -0.5 * (np.einsum("ik,jl->ijkl", Ci, Ci) + np.einsum("il,jk->ijkl", Ci, Ci))
    dJ = 0.5 * J * Ci                                   # dJ/dC_kl
    dJm23 = -2.0 / 3.0 * J ** (-5.0 / 3.0) * dJ
    dp = 3 * K * (-2.0 / 3.0 * J ** (-5.0 / 3.0) * (J ** (1.0 / 3.0) - 1.0) + J ** (-2.0 / 3.0) / 3.0 * J ** (-2.0 / 3.0)) * dJ
    iso = I - trC / 3.0 * Ci
    diso = -np.einsum("ij,kl->ijkl", Ci, I) / 3.0 - trC / 3.0 * dCi
    dS = 2 * C10 * (np.einsum("ij,kl->ijkl", iso, dJm23) + J ** (-2.0 / 3.0) * diso) \
        + np.einsum("ij,kl->ijkl", Ci, p * dJ + J * dp) + J * p * dCi
    return to_voigt(S), tensor4_to_voigt(2.0 * dS)


def base(Clin, F):
    """The surrogate's base stress and tangent (Voigt): NH(K_V, mu_V) + (C_lin - C_iso) : E."""
    K, mu = iso_moduli(Clin)
    dC = from_up21(Clin) - iso_stiffness(K, mu)
    E = 0.5 * (F.T @ F - np.eye(3))
    S, C = nh_material(0.5 * mu, K, F)
    return S + dC @ (to_voigt(E) * ENG), C + dC


def stretch(F):
    """U = (F^T F)^(1/2), its eigenvalues and eigenvectors."""
    lam2, Q = np.linalg.eigh(F.T @ F)
    lam = np.sqrt(lam2)
    return Q @ np.diag(lam) @ Q.T, lam, Q


def features(cfg, F, Clin):
    U, lam, _ = stretch(F)
    x = np.concatenate([to_voigt(U - np.eye(3)) / cfg["u_scale"],
                        (np.asarray(Clin) - cfg["d_mean"]) / cfg["d_std"]])
    inside = bool(np.all(lam >= cfg["lam_lo"]) and np.all(lam <= cfg["lam_hi"]) and
                  np.all(x[6:] >= cfg["d_lo"]) and np.all(x[6:] <= cfg["d_hi"]))
    return x, inside


def surrogate_stress(p, cfg, F, Clin, width):
    """S (Voigt) of the surrogate at F."""
    x, inside = features(cfg, F, Clin)
    y, _, _ = mo.forward(p, x[None, :], 27, width, 6)
    E = 0.5 * (F.T @ F - np.eye(3))
    phi = np.sum(E * E) / cfg["e_ref"] ** 2
    t = cfg["t_mean"] + cfg["t_std"] * y[0]
    return base(Clin, F)[0] + phi * t, inside


def surrogate(p, cfg, F, Clin, width):
    """S (6), symmetrized dS/dE (6, 6) and hull flag of the surrogate at F (analytic tangent:
    chain rule through the network Jacobian and dU/dE from U dU + dU U = 2 dE)."""
    x, inside = features(cfg, F, Clin)
    y, _, _ = mo.forward(p, x[None, :], 27, width, 6)
    Jx = mo.jacobian(p, x[None, :], 27, width, 6)[0][:, :6]
    E = 0.5 * (F.T @ F - np.eye(3))
    phi = np.sum(E * E) / cfg["e_ref"] ** 2
    t = cfg["t_mean"] + cfg["t_std"] * y[0]
    Sb, Cb = base(Clin, F)
    S = Sb + phi * t
    _, lam, Q = stretch(F)
    Ct = np.zeros((6, 6))
    for b, (K, L) in enumerate(VOIGT):
        dE = np.zeros((3, 3))
        dE[K, L] += 0.5
        dE[L, K] += 0.5
        dEp = Q.T @ dE @ Q
        dUp = 2.0 * dEp / (lam[:, None] + lam[None, :])
        dU = Q @ dUp @ Q.T
        dx = to_voigt(dU) / cfg["u_scale"]
        dphi = 2.0 * E[K, L] / cfg["e_ref"] ** 2
        Ct[:, b] = Cb[:, b] + t * dphi + phi * cfg["t_std"] * (Jx @ dx)
    return S, 0.5 * (Ct + Ct.T), inside

