"""FP64 numpy restatement of the FE^2 macro-point pieces (test infrastructure, checker only).

  pull_back(sigma, D, F)     spatial (Cauchy, Voigt tangent) -> material (S, dS/dE): the inverse
                             of the push forward the element kernel applies to external points
                             (Push_Forward_Voigt, Tools/MathTools/ModMathRoutines.f90:254-311)
  rve_to_ext(F, P, A)        S = F^-1 P, dS/dE from A = dP/dF
                             (First_Elasticity_Modulus_To_Second_Voigt, ModMathRoutines.f90:315)
  surrogate(p, cfg, F, Clin) the surrogate's S, dS/dE and hull flag:
                             S = C_lin:E + |E|^2/e_ref^2 (t_mean + t_std MLP(x)),
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
    return from_up21(Clin) @ (to_voigt(E) * ENG) + phi * t, inside


def surrogate(p, cfg, F, Clin, width):
    """S (6), symmetrized dS/dE (6, 6) and hull flag of the surrogate at F (analytic tangent:
    chain rule through the network Jacobian and dU/dE from U dU + dU U = 2 dE)."""
    x, inside = features(cfg, F, Clin)
    y, _, _ = mo.forward(p, x[None, :], 27, width, 6)
    Jx = mo.jacobian(p, x[None, :], 27, width, 6)[0][:, :6]
    E = 0.5 * (F.T @ F - np.eye(3))
    phi = np.sum(E * E) / cfg["e_ref"] ** 2
    t = cfg["t_mean"] + cfg["t_std"] * y[0]
    Cl = from_up21(Clin)
    S = Cl @ (to_voigt(E) * ENG) + phi * t
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
        Ct[:, b] = Cl[:, b] + t * dphi + phi * cfg["t_std"] * (Jx @ dx)
    return S, 0.5 * (Ct + Ct.T), inside
