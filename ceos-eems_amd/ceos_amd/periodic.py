"""Periodic RVEs on general hex8 meshes: the reference's periodic multiscale analysis of one RVE
(FEM/ModFEMAnalysis.f90:1440-1640 with ClassMultiscalePeriodicFEMSoE) on the GPU.

  periodicity matrix T      BuildT, FEM/Modules/ModMultiscalePeriodicFEMSoE.f90:226-895:
                            the box of the RVE from the extreme coordinates; nodes on the faces,
                            edges and corners (tol 1e-6); every node of the Xp / Yp / Zp face is
                            the slave of the node of the opposite face with the same in-plane
                            coordinates, the three parallel copies of each edge are slaves of the
                            (min, min) edge, the seven other corners of the (xmin, ymin, zmin)
                            corner; the reduced DOFs are the remaining ones in their order
  system                    EvaluateR / EvaluateKt / ExpandFluctuation (:77-222): K_red = T^T K T,
                            R_red = T^T Fint, U += T DeltaU_red; the corners fixed
                            (verticesDOF as fixed supports, FEM/ModFEMAnalysis.f90:1529-1530)
  macro displacement        GetBoundaryConditionsMultiscalePeriodic,
                            MultiScale/ModMultiscaleBoundaryConditions.f90:200-289: U = (F - I) X
                            at the new macro F; the step starts from the last converged U plus
                            the affine increment (the Taylor step of ExpandFluctuation)
  Newton                    Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:44-168,
                            max-abs residual, relative test
  homogenization            HomogenizeStress, MultiScale/ModMultiscaleAnalysis.f90:140-329

Everything per iteration runs in libceos_gpu.so on the device (element evaluation of the RVE
mesh, the T products, the boundary conditions and norm, PCG on the reduced system); the host
keeps the loop and reads one norm per iteration.  The homogenized tangent dP/dF comes from nine
solves with the converged reduced tangent (static condensation): for every macro direction
E_kL the affine field dU = e_k X_L, the fluctuation from K_red dw = T^T K dU, and
dP = (1/V0) sum_A (K (dU - T dw))_A (x) X_A.

Differences from the reference, on purpose: the box is taken from the extreme coordinates of the
nodes (the reference starts its min/max search at 0, so it needs the origin inside the box); a
node of a slave face, edge or corner without a partner is an error (the reference leaves it
unconstrained); the Taylor step enters as the start of the Newton iteration instead of through
the first correction, so iteration counts can differ by one while the converged state is the same.
"""
import numpy as np

from . import gpu, hiprt


class PeriodicMeshError(ValueError):
    pass


class PeriodicMap:
    """node_map (reduced node of every node, 0-based), masters (the full node of every reduced
    node, ascending), corners (the 8 corner nodes in the reference's order XmYmZm, XpYmZm,
    XmYpZm, XmYmZp, XpYpZp, XmYpZp, XpYmZp, XpYpZm), slave_of (master of every node, -1 for a
    node that is its own master) and the box (lo, hi)."""

    def __init__(self, node_map, masters, corners, slave_of, lo, hi):
        self.node_map, self.masters, self.corners = node_map, masters, corners
        self.slave_of, self.lo, self.hi = slave_of, lo, hi


def periodic_map(X, tol=1e-6, boundary=None):
    """BuildT's pairing of the nodes of a box-shaped RVE mesh.  boundary: node ids (0-based) to
    classify (the reference's boundary node set); default every node."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    nn = len(X)
    cand = np.arange(nn) if boundary is None else np.unique(np.asarray(boundary, dtype=np.int64))
    lo, hi = X[cand].min(axis=0), X[cand].max(axis=0)
    on_lo = np.abs(X[cand] - lo) < tol  # (n, 3)
    on_hi = np.abs(X[cand] - hi) < tol
    side = np.where(on_lo, -1, np.where(on_hi, 1, 0))  # per axis: -1 min face, +1 max face, 0 inside
    slave_of = np.full(nn, -1, dtype=np.int64)

    def group(sig):
        m = np.all(side == np.asarray(sig), axis=1)
        return cand[m]

    def pair(masters, slaves, axes, what):
        if len(masters) != len(slaves):
            raise PeriodicMeshError("%s: %d master and %d slave nodes" % (what, len(masters), len(slaves)))
        if not len(slaves):
            return
        d = np.zeros((len(slaves), len(masters)), dtype=bool)
        d[:] = True
        for ax in axes:
            d &= np.abs(X[slaves, ax][:, None] - X[masters, ax][None, :]) < tol
        cnt = d.sum(axis=1)
        if np.any(cnt != 1) or np.any(d.sum(axis=0) != 1):
            raise PeriodicMeshError("%s: the nodes of opposite sides do not match one to one within %g" % (what, tol))
        slave_of[slaves] = masters[np.argmax(d, axis=1)]

    # corners: the seven others are slaves of XmYmZm
    order = [(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1), (1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1)]
    corners = []
    for sig in order:
        g = group(sig)
        if len(g) != 1:
            raise PeriodicMeshError("corner %s: %d nodes" % (str(sig), len(g)))
        corners.append(int(g[0]))
    for c in corners[1:]:
        slave_of[c] = corners[0]
    # edges: the parallel copies of each edge are slaves of its (min, min) copy, matched along it
    for ax in range(3):
        a, b = [k for k in range(3) if k != ax]
        sig = [0, 0, 0]
        sig[a], sig[b] = -1, -1
        master = group(sig)
        for sa, sb in [(-1, 1), (1, 1), (1, -1)]:
            s2 = [0, 0, 0]
            s2[a], s2[b] = sa, sb
            pair(master, group(s2), [ax], "edge %s" % str(s2))
    # faces: the max face is the slave of the min face, matched in its plane
    for ax in range(3):
        sig_m, sig_p = [0, 0, 0], [0, 0, 0]
        sig_m[ax], sig_p[ax] = -1, 1
        pair(group(sig_m), group(sig_p), [k for k in range(3) if k != ax], "face %s" % str(sig_p))
    is_master = slave_of < 0
    masters = np.nonzero(is_master)[0]
    red = np.full(nn, -1, dtype=np.int64)
    red[masters] = np.arange(len(masters))
    node_map = np.where(is_master, red, red[np.where(is_master, 0, slave_of)])
    return PeriodicMap(node_map.astype(np.int32), masters, np.array(corners), slave_of, lo, hi)


class PeriodicRve:
    """One periodic RVE on a general hex8 mesh: X (nn, 3), conn (ne, 8) 1-based, mat_id (ne,),
    materials {material id: (model, props)}.  solve(F) brings it to equilibrium under the macro
    deformation gradient F from its last committed state; commit() makes that state the
    converged one (SwitchConvergedState).

    The reduced system is solved by PCG (in place of the reference's PARDISO solve of
    K_red, ModMultiscalePeriodicFEMSoE.f90:77-175) with the library's preconditioner
    `preconditioner`: by default the two-level one (block Jacobi plus the rigid-body modes of node
    aggregates of the reduced mesh), whose coarse inverse serves `coarse_reuse` further solves, so
    that the Newton steps of a solve and the nine condensation solves share one or two (a
    preconditioner from a nearby tangent is still symmetric positive definite).  Every solve()
    starts with a fresh coarse inverse, so a solve depends only on the state it starts from (a
    restarted analysis repeats the uninterrupted one bitwise)."""

    def __init__(self, X, conn, mat_id=None, materials=None, nl=True, tol=1e-6, boundary=None, newton_rtol=1e-10,
                 newton_atol=1e-13, newton_maxit=12, pcg_rtol=1e-10, pcg_maxit=20000,
                 preconditioner=gpu.PREC_TWO_LEVEL, coarse_reuse=8):
        self.X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, 8)
        self.nn, self.ne = len(self.X), len(conn)
        self.map = periodic_map(self.X, tol, boundary)
        nm = self.map.node_map
        conn_red = nm[conn - 1] + 1
        if any(len(set(r)) < 8 for r in conn_red):
            raise PeriodicMeshError("an element holds two periodic images of one node: the RVE needs at least two "
                                    "elements across")
        self.nl = nl
        materials = materials or {}
        self.full = gpu.CeosMesh(self.X, conn, mat_id=mat_id, nl_analysis=nl)
        self.red = gpu.CeosMesh(self.X[self.map.masters], conn_red, mat_id=mat_id, nl_analysis=nl)
        for mid, (model, props) in materials.items():
            self.full.set_material(mid, model, props)
            self.red.set_material(mid, model, props)
        self.full.set_bc([], [])  # neighbour slot tables for the products with K
        c = int(nm[self.map.corners[0]])
        self.fixed = np.array([3 * c, 3 * c + 1, 3 * c + 2])
        self.red.set_bc([], self.fixed)
        self.prec = (preconditioner, 0.0, coarse_reuse if preconditioner == gpu.PREC_TWO_LEVEL else 0)
        self.red.set_preconditioner(*self.prec)
        self.per = gpu.Periodic(self.full, self.red, nm)
        self.nn_red, _, self.volume = self.per.info()
        self.newton = (newton_rtol, newton_atol, newton_maxit)
        self.pcg = (pcg_rtol, pcg_maxit)
        nf, nr = 3 * self.nn, 3 * self.nn_red
        self.nnz = self.full.info()[1]
        self.nnz_red = self.red.info()[1]
        D = hiprt.DeviceArray
        self.dU, self.dUc, self.dFint, self.dK = D(nf), D(nf), D(nf), D(self.nnz)
        self.dKred, self.dR, self.dRbc, self.dw = D(self.nnz_red), D(nr), D(nr), D(nr)
        self.dzero_r, self.dnorm, self.dP = D(nr), D(3), D(9)
        self.dtmp, self.dtmp2, self.dy = D(nf), D(nf), D(nf)
        self.dUc.upload(np.zeros(nf))
        self.dzero_r.upload(np.zeros(nr))
        self.stream = hiprt.Stream()
        self.Fconv = np.eye(3)
        self.F = np.eye(3)
        self.time = 0.0

    def close(self):
        for k in ("per", "red", "full"):
            o = getattr(self, k, None)
            if o is not None:
                o.close()

    def _affine(self, dF):
        """(dF) X as a nodal displacement vector."""
        return (self.X @ np.asarray(dF, dtype=np.float64).T).reshape(-1)

    def solve(self, F, time=None, tangent=False):
        """Equilibrium at the macro F (3x3).  Returns (P (9, column-major), A (81, A[9(3J+i) + 3L+k]
        = dP_iJ/dF_kL) or None, info {iterations, pcg_iterations, norms})."""
        F = np.asarray(F, dtype=np.float64).reshape(3, 3)
        if time is not None:
            self.time = float(time)
        s = self.stream.ptr
        rtol, atol, maxit = self.newton
        prt, pmax = self.pcg
        self.red.set_preconditioner(*self.prec)  # a fresh coarse inverse at the first solve
        self.stream.synchronize()
        # the step starts from the committed U plus the affine increment
        self.dtmp.upload(self._affine(F - self.Fconv))
        hiprt.check(hiprt.hip().hipMemcpyAsync(self.dU.ptr, self.dUc.ptr, self.dU.nbytes, 3, s), "hipMemcpyAsync")
        self.full.newton_update_device(self.dU.ptr, self.dtmp.ptr, 1.0, s)
        norms, pcg_its = [], []
        converged = False
        for it in range(maxit + 1):
            if self.nl:
                self.full.update_mesh_device(self.dU.ptr, s)
            self.full.evaluate_device(self.time, self.dU.ptr, self.dFint.ptr, self.dK.ptr, s)
            self.per.reduce_device(self.dK.ptr, self.dFint.ptr, self.dKred.ptr, self.dR.ptr, False, s)
            self.red.residual_bc_device(it, self.dzero_r.ptr, self.dR.ptr, self.dKred.ptr, None, self.dzero_r.ptr,
                                        self.dRbc.ptr, self.dnorm.ptr, 2, s)
            self.stream.synchronize()
            code, st = self.full.device_status()
            if code != 0:
                raise gpu.CeosError(code, "RVE element evaluation")
            nrm = float(self.dnorm.numpy()[2])
            norms.append(nrm)
            if nrm <= max(rtol * norms[0], atol) and (it > 0 or nrm <= atol):
                converged = True
                break
            if it == maxit:
                break
            its, rr, fl = self.red.linear_solve_device(self.dKred.ptr, self.dRbc.ptr, self.dw.ptr, prt, pmax, s)
            pcg_its.append(its)
            if fl not in (1, 4):
                raise RuntimeError("RVE linear solve flag %d at Newton iteration %d" % (fl, it))
            self.per.expand_device(self.dw.ptr, -1.0, self.dU.ptr, s)
        if not converged:
            raise gpu.CeosError(gpu.ERR_LOCAL_NR, "RVE Newton did not converge: norms %s" % norms)
        self.F = F.copy()
        self.per.homogenize_device(self.dFint.ptr, self.dP.ptr, s)
        self.stream.synchronize()
        P = self.dP.numpy()
        A = self._tangent() if tangent else None
        return P, A, {"iterations": len(norms) - 1, "pcg_iterations": pcg_its, "norms": norms}

    def _tangent(self):
        """dP/dF by static condensation at the converged state (K and the BC-applied K_red of the
        last evaluation)."""
        s = self.stream.ptr
        prt, pmax = self.pcg
        A = np.zeros(81)
        for kl in range(9):
            k, L = kl % 3, kl // 3
            E = np.zeros((3, 3))
            E[k, L] = 1.0
            self.stream.synchronize()
            self.dtmp.upload(self._affine(E))
            self.full.spmv_device(self.dK.ptr, self.dtmp.ptr, self.dy.ptr, s)
            self.per.reduce_device(None, self.dy.ptr, None, self.dR.ptr, True, s)
            self.red.linear_solve_device(self.dKred.ptr, self.dR.ptr, self.dw.ptr, prt, pmax, s)
            self.per.expand_device(self.dw.ptr, -1.0, self.dtmp.ptr, s)
            self.full.spmv_device(self.dK.ptr, self.dtmp.ptr, self.dy.ptr, s)
            self.per.homogenize_device(self.dy.ptr, self.dP.ptr, s)
            self.stream.synchronize()
            dP = self.dP.numpy()
            for iJ in range(9):
                A[9 * iJ + kl] = dP[iJ]
        return A

    def accept(self):
        """The solved state becomes the start of the next solve (a converged substep: the
        Gauss-point states stay committed at the start of the step)."""
        self.stream.synchronize()
        hiprt.check(hiprt.hip().hipMemcpy(self.dUc.ptr, self.dU.ptr, self.dU.nbytes, 3), "hipMemcpy")
        self.Fconv = self.F.copy()

    def commit(self):
        """SwitchConvergedState of the RVE's Gauss points; the solved state becomes the start of
        the next solve."""
        self.accept()
        self.full.switch_converged_state()

    def set_start(self, U, F):
        """Start the next solve from the displacement U at the macro F (a restart; the Gauss-point
        states are the handle's committed ones)."""
        self.stream.synchronize()
        self.dUc.upload(np.asarray(U, dtype=np.float64).reshape(-1))
        hiprt.check(hiprt.hip().hipMemcpy(self.dU.ptr, self.dUc.ptr, self.dU.nbytes, 3), "hipMemcpy")
        self.Fconv = np.asarray(F, dtype=np.float64).reshape(3, 3).copy()
        self.F = self.Fconv.copy()

    def displacement(self):
        self.stream.synchronize()
        return self.dU.numpy()
