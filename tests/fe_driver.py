"""Quasi-static Newton driver used by the tests to run whole load histories through a backend.

It restates the reference's caller of the element hot path, which stays outside the GPU
library (the reference keeps its own Newton loop and PARDISO):
  load case / step / cut-back loop   FEM/ModFEMAnalysis.f90:1136-1374
  external forces, prescribed disp.  FEM/BoundaryConditions/ModBoundaryConditions.f90:540-640
  R = Fint - Fext; K + Dirichlet BC  FEM/Modules/ModFEMSystemOfEquations.f90:40-93,
                                     FEM/BoundaryConditions/ModBoundaryConditions.f90:246-340
  Newton, max-abs norm, relative test Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:44-168
  load tables (linear interpolation) FEM/BoundaryConditions/ModLoadHistoryData.f90:122-223
The sparse direct solve (PARDISO in the reference) is scipy's SuperLU on the symmetric matrix
expanded from the upper CSR.

A backend provides: ndof, csr() -> (rowmap, col) 1-based; solve_constitutive(time, U) -> status;
internal_force() -> (Fint, status); tangent() -> Kval; update_mesh(U); switch_converged_state().
A backend may also provide the Newton system itself: newton_system(it, fext, bc, Ubar, X) ->
(R with the boundary conditions, norm) and newton_solve(b) -> K^-1 b with K after the boundary
conditions (the library's device BC, norm and PCG); the loop then calls those instead.
"""
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

GOLDEN_RATIO = (1.0 + math.sqrt(5.0)) / 2.0


class NewtonFailed(Exception):
    pass


def interp_table(table, t):
    ts = [p[0] for p in table]
    vs = [p[1] for p in table]
    return float(np.interp(t, ts, vs))


def sym_from_upper(rowmap, col, val, n):
    rows = np.repeat(np.arange(n), np.diff(rowmap))
    cols = col.astype(np.int64) - 1
    U = sp.csr_matrix((val, (rows, cols)), shape=(n, n))
    return (U + U.T - sp.diags(U.diagonal())).tocsc()


def sym_upper_matvec(rowmap, col, val, x):
    """y = K x with K symmetric stored as its upper triangle (mkl_dcsrsymv 'U')."""
    n = len(rowmap) - 1
    U = sp.csr_matrix((val, col.astype(np.int64) - 1, rowmap.astype(np.int64) - 1), shape=(n, n))
    return U @ x + U.T @ x - U.diagonal() * x


class BoundaryConditions:
    """Fixed supports (zero) and prescribed nodal displacements / nodal forces from tables.

    fixed: 0-based DOF array.  presc: list of (dof0, callable t -> value).
    loads: list of (dof0, callable t -> value)."""

    def __init__(self, ndof, fixed, presc, loads=()):
        self.ndof = ndof
        self.fixed = np.asarray(sorted(set(int(d) for d in fixed)), dtype=np.int64)
        self.presc = list(presc)
        self.loads = list(loads)
        self.presc_dof = np.asarray([d for d, _ in self.presc], dtype=np.int64)

    def fext(self, t):
        f = np.zeros(self.ndof)
        for d, fn in self.loads:
            f[d] += fn(t)
        return f

    def presc_values(self, t):
        return np.asarray([fn(t) for _, fn in self.presc])


def apply_dirichlet(rowmap, col, K, R, bc, Ubar, X, it):
    """ApplyBoundaryConditionsNEW on the residual sign used by EvaluateKt (R -> -R -> BC -> -R).
    Returns (Kmod, Rmod)."""
    K = K.copy()
    R = -R.copy()
    n = len(rowmap) - 1
    rows = np.repeat(np.arange(n), np.diff(rowmap))
    cols = col.astype(np.int64) - 1
    Ud = np.zeros(n)
    pd = bc.presc_dof
    if len(pd):
        if it == 0:
            Ud[pd] = Ubar[pd] - X[pd]
            R = R - sym_upper_matvec(rowmap, col, K, Ud)
        m = np.zeros(n, dtype=bool)
        m[pd] = True
        K[m[rows] | m[cols]] = 0.0
        K[m[rows] & (rows == cols)] = 1.0
        R[pd] = Ud[pd]
    if len(bc.fixed):
        m = np.zeros(n, dtype=bool)
        m[bc.fixed] = True
        K[m[rows] | m[cols]] = 0.0
        K[m[rows] & (rows == cols)] = 1.0
        R[bc.fixed] = 0.0
    return K, -R


def newton(backend, time, fext, bc, Ubar, Xguess, nl, tol, maxiter, norms_out=None,
           absolute=False):
    """NewtonRaphsonFull_Solve. Returns (X, iterations). Raises NewtonFailed.

    absolute=False: the current reference test ||R|| < ||R0|| tol for it > 0
    (ModNewtonRaphsonFull.f90:122).  absolute=True: ||R|| < tol, the test of the binary that
    wrote the regression logs (Checker/Cases/*/Log_CEOS.txt stop at the first norm < 1e-3)."""
    rowmap, col = backend.csr()
    n = backend.ndof
    X = Xguess.copy()
    it = 0
    normR0 = None
    system = getattr(backend, "newton_system", None)
    while True:
        st = backend.solve_constitutive(time, X)
        if st:
            raise NewtonFailed("evaluate system status %d" % st)
        if system is not None:
            R, normR = system(it, fext, bc, Ubar, X)
        else:
            Fint, st = backend.internal_force()
            if st:
                raise NewtonFailed("internal force status %d" % st)
            R = Fint - fext
            K = backend.tangent()
            Kb, R = apply_dirichlet(rowmap, col, K, R, bc, Ubar, X, it)
            normR = float(np.max(np.abs(R)))
        if norms_out is not None:
            norms_out.append(normR)
        if it == 0:
            normR0 = normR
        if it > 0 and normR < (tol if absolute else normR0 * tol):
            return X, it
        if it >= maxiter:
            raise NewtonFailed("maximum number of iterations")
        it += 1
        if system is not None:
            DX = backend.newton_solve(-R)
        else:
            A = sym_from_upper(rowmap, col, Kb, n)
            DX = spla.spsolve(A, -R)
        X = X + DX
        if nl:
            backend.update_mesh(X)


def quasi_static(backend, bc, load_cases, nl=True, tol=1e-3, maxiter=10, max_cutback=10,
                 on_step=None, norms=None, absolute=False):
    """QuasiStaticAnalysisFEM. load_cases: list of (t0, t1, nsteps).
    on_step(lc, st, time, U, iterations) is called at every converged end of step.
    norms: optional list receiving, per substep solve, the list of residual norms."""
    n = backend.ndof
    U = np.zeros(n)
    history = []
    for lc, (t0, t1, nsteps) in enumerate(load_cases, start=1):
        dt = (t1 - t0) / nsteps
        for stp in range(1, nsteps + 1):
            ta, tb = t0 + (stp - 1) * dt, t0 + stp * dt
            fext0, dfext = bc.fext(ta), bc.fext(tb) - bc.fext(ta)
            v0, v1 = bc.presc_values(ta), bc.presc_values(tb)
            dpresc = np.zeros(n)
            if len(bc.presc_dof):
                U[bc.presc_dof] = v0
                dpresc[bc.presc_dof] = v1 - v0
            Ubar0 = U.copy()
            Uconv = U.copy()
            amax, amin = 1.0, 0.0
            alpha = amax
            cutback = 0
            while True:
                time = ta + alpha * dt
                fext = fext0 + alpha * dfext
                Ubar = Ubar0 + alpha * dpresc
                nrm = []
                try:
                    U, its = newton(backend, time, fext, bc, Ubar, Uconv, nl, tol, maxiter, nrm,
                                    absolute)
                    ok = True
                except NewtonFailed:
                    ok = False
                if norms is not None:
                    norms.append(nrm)
                if not ok:
                    alpha = amin + (1.0 - 1.0 / GOLDEN_RATIO) * (alpha - amin)
                    U = Uconv.copy()
                    if nl:
                        backend.update_mesh(U)
                    cutback += 1
                    if cutback > max_cutback:
                        raise NewtonFailed("step %d did not converge with %d cut backs" % (stp, max_cutback))
                elif alpha == 1.0:
                    history.append((lc, stp, time, U.copy(), its))
                    if on_step:
                        on_step(lc, stp, time, U, its)
                    break
                else:
                    aux = amin
                    amin = alpha
                    alpha = min(alpha + GOLDEN_RATIO * (alpha - aux), 1.0)
                    Uconv = U.copy()
            backend.switch_converged_state()
    return history


def bc_from_case(case):
    """BoundaryConditions from a tests/golden/*_case.json fixture (1-based node ids)."""
    ndof = 3 * len(case["coords"])
    fixed = []
    for row in case["fixed"]:
        node = row[0]
        for k in range(3):
            if row[1 + k]:
                fixed.append(3 * (node - 1) + k)
    tabs = case["tables"]

    def mk(tab):
        return lambda t, tab=tab: interp_table(tabs[tab], t)

    presc = [(3 * (node - 1) + dof - 1, mk(tab)) for node, dof, tab in case["nodal_disp"]]
    loads = [(3 * (node - 1) + dof - 1, mk(tab)) for node, dof, tab in case["nodal_load"]]
    return BoundaryConditions(ndof, fixed, presc, loads)


def load_cases_from_case(case):
    return [(c["t0"], c["t1"], c["steps"]) for c in case["time"]]
