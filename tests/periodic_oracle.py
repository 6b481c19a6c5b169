"""CPU restatement of the reference's periodic RVE analysis on a general hex8 mesh
(FEM/Modules/ModMultiscalePeriodicFEMSoE.f90:77-222 and BuildT :226-895; the corners fixed,
FEM/ModFEMAnalysis.f90:1529-1530; U = (F - I) X, MultiScale/ModMultiscaleBoundaryConditions.f90:
200-289; HomogenizeStress, MultiScale/ModMultiscaleAnalysis.f90:140-329), used as the checker of
ceos_amd.periodic.  Test infrastructure only.

It pairs the nodes without the face / edge / corner classification of BuildT: two nodes are
periodic images when their coordinates differ by a lattice vector of the box (within tol), and
the lowest node id of every class of images is its representative; the corner class is fixed.
The element loops are the oracle's (tests/oracle_lib.py); T^T K T, T^T R and the solves are scipy
sparse algebra; the homogenized stress is the oracle's volume average of P over the Gauss points.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_lib as ol


def image_classes(X, tol=1e-6):
    """Representative (lowest id) of every node's class of periodic images, and the box."""
    X = np.asarray(X, dtype=np.float64)
    lo, hi = X.min(axis=0), X.max(axis=0)
    L = hi - lo
    # wrap every coordinate onto [lo, hi): a node on a max face lands on the min face
    W = X - lo
    for ax in range(3):
        W[np.abs(W[:, ax] - L[ax]) < tol, ax] = 0.0
    rep = np.arange(len(X))
    for A in range(len(X)):
        d = np.max(np.abs(W[:A] - W[A]), axis=1) if A else np.zeros(0)
        hit = np.nonzero(d < tol)[0]
        if len(hit):
            rep[A] = rep[hit[0]]
    return rep, lo, hi


class OraclePeriodicRve:
    def __init__(self, X, conn, mat_id, materials, nl=True, tol=1e-6):
        self.X = np.asarray(X, dtype=np.float64)
        self.nn = len(self.X)
        self.m = ol.OracleMesh(self.X, conn, mat_id=mat_id, nl_analysis=nl)
        for mid, (model, props) in materials.items():
            self.m.set_material(mid, model, props)
        self.nl = nl
        rep, lo, hi = image_classes(self.X, tol)
        reps = np.unique(rep)
        col = np.searchsorted(reps, rep)
        rows = np.concatenate([3 * np.arange(self.nn) + k for k in range(3)])
        cols = np.concatenate([3 * col + k for k in range(3)])
        self.T = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(3 * self.nn, 3 * len(reps)))
        corner = np.nonzero(np.all(np.abs(self.X - lo) < tol, axis=1))[0]
        c = col[corner[0]]
        self.free = np.setdiff1d(np.arange(3 * len(reps)), [3 * c, 3 * c + 1, 3 * c + 2])
        self.U = np.zeros(3 * self.nn)
        self.Fconv = np.eye(3)

    def _system(self, U, time):
        if self.nl:
            self.m.update_mesh(U)
        st = self.m.solve_constitutive(time, U)
        if st:
            raise RuntimeError("oracle evaluation status %d" % st)
        Fint, _ = self.m.internal_force()
        rm, col = self.m.csr()
        Kv = self.m.tangent()
        n = 3 * self.nn
        r = np.repeat(np.arange(n), np.diff(rm))
        Ku = sp.csr_matrix((Kv, (r, col - 1)), shape=(n, n))
        K = Ku + sp.triu(Ku, 1).T
        return Fint, K.tocsr()

    def solve(self, F, time=0.0, rtol=1e-12, maxit=20):
        F = np.asarray(F, dtype=np.float64)
        U = self.U + (self.X @ (F - self.Fconv).T).reshape(-1)
        nrm0 = None
        for it in range(maxit + 1):
            Fint, K = self._system(U, time)
            R = (self.T.T @ Fint)[self.free]
            nrm = np.max(np.abs(R))
            nrm0 = nrm if nrm0 is None else nrm0
            if nrm <= max(rtol * nrm0, 1e-14):
                break
            Kr = (self.T.T @ K @ self.T).tocsc()[self.free][:, self.free]
            dw = np.zeros(self.T.shape[1])
            dw[self.free] = spla.spsolve(Kr.tocsc(), R)
            U = U - self.T @ dw
        else:
            raise RuntimeError("oracle periodic Newton did not converge")
        self.Unew, self.Fnew = U, F
        P, Fh = self.m.homogenize()
        return P, U, it

    def commit(self):
        self.m.switch_converged_state()
        self.U, self.Fconv = self.Unew.copy(), self.Fnew.copy()


def periodic_perturbation(X, h, amp=0.2, tol=1e-9):
    """In-plane periodic perturbation of a box mesh: the normal displacement vanishes on every
    face and periodic images move alike, so the mesh stays periodic with non-grid face nodes."""
    X = np.asarray(X, dtype=np.float64)
    lo, hi = X.min(axis=0), X.max(axis=0)
    s = 2.0 * np.pi * (X - lo) / (hi - lo)
    d = np.zeros_like(X)
    d[:, 0] = np.sin(s[:, 0]) * (1.0 + 0.5 * np.cos(s[:, 1] + 1.0) * np.cos(s[:, 2] + 2.0))
    d[:, 1] = np.sin(s[:, 1]) * (1.0 + 0.5 * np.cos(s[:, 0] + 0.5) * np.cos(s[:, 2] + 1.5))
    d[:, 2] = np.sin(s[:, 2]) * (1.0 + 0.5 * np.cos(s[:, 0] + 2.5) * np.cos(s[:, 1] + 0.7))
    # on the faces sin() is zero only up to rounding: pin those components exactly
    for ax in range(3):
        on = (np.abs(X[:, ax] - lo[ax]) < tol) | (np.abs(X[:, ax] - hi[ax]) < tol)
        d[on, ax] = 0.0
    return X + amp * h * d
