"""FE^2 on the GPU: a hex8 macro mesh whose Gauss points are periodic voxel RVEs.

The reference drives one RVE analysis at a time through a prescribed macro F history
(MultiScale/ModMultiscaleAnalysis.f90:445-470); it has no concurrent macro loop.  Here one macro
evaluation (residual + tangent of the macro mesh) needs the homogenized response of every macro
Gauss point, obtained either

  plain      by the batched periodic RVE solver (ceos_rve_solve_device: every RVE to
             equilibrium at its macro F, then its condensed tangent), or
  surrogate  by the MLP surrogate (ceos_surrogate_eval_device), with the RVE solve as fallback
             for points outside the training hull and as judge on a sampled fraction of points.

Either way the points' S and dS/dE go to the macro element kernel as external material points
(CEOS_MODEL_EXTERNAL) and the macro Fint and CSR tangent are assembled on the GPU.

The surrogate's descriptor of a microstructure is its small-strain homogenized stiffness C_lin
(the condensed tangent at F = I), computed once per RVE at setup (`descriptors`).  The surrogate
predicts S as a base (Neo-Hookean with the Voigt-average moduli of C_lin plus the anisotropic
rest of C_lin, `base_stress`) plus the network's residual.
"""
import math
import os
import time

import numpy as np

from . import gpu, hiprt, meshgen, rve

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SURROGATE_PATH = os.path.join(PKG, "data", "fe2_surrogate.npz")

LAM_LO, LAM_HI = 0.8, 1.25     # principal stretches of the training distribution
U_SCALE = 0.25                 # U features: (U - I) / U_SCALE
E_REF = 0.1                    # output scaling: S = S_base + (|E|^2 / E_REF^2) (t_mean + t_std y)
N_IN, WIDTH, N_OUT = 27, 256, 6
VOIGT = ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2))
PROPS = np.array([rve.MATRIX_PROPS, rve.INCLUSION_PROPS])


# ------------------------------------------------------------------------------------------
# stretch sampling (the surrogate's training distribution)
# ------------------------------------------------------------------------------------------
def haar_rotations(rng, n):
    """n rotation matrices, uniform (Haar) on SO(3)."""
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    neg = np.linalg.det(q) < 0
    q[neg, :, 0] *= -1.0
    return q


def random_stretch(rng, n, lo=LAM_LO, hi=LAM_HI):
    """Symmetric right stretch tensors: principal stretches uniform in [lo, hi], random axes."""
    lam = rng.uniform(lo, hi, size=(n, 3))
    Q = haar_rotations(rng, n)
    return np.einsum("nij,nj,nkj->nik", Q, lam, Q)


def _reflect(x, lo, hi):
    w = hi - lo
    y = np.mod(x - lo, 2.0 * w)
    return lo + np.where(y > w, 2.0 * w - y, y)


def isochoric(rng, lam, sig_logj, lo=LAM_LO, hi=LAM_HI):
    """Principal stretches rescaled to a volume ratio J = exp(N(0, sig_logj)), kept in [lo, hi]."""
    J = np.prod(lam, axis=-1, keepdims=True)
    target = np.exp(rng.normal(scale=sig_logj, size=J.shape))
    return np.clip(lam * (target / J) ** (1.0 / 3.0), lo, hi)


def walk_stretch(rng, U, sig_lam=0.05, sig_rot=0.25, lo=LAM_LO, hi=LAM_HI, sig_logj=None):
    """One step of a reflected random walk of stretch tensors (principal stretches stay in
    [lo, hi]; the principal axes turn by a random small rotation).  With sig_logj the stretches
    are rescaled to a near-isochoric volume ratio (log J ~ N(0, sig_logj))."""
    lam, Q = np.linalg.eigh(U)
    lam = _reflect(lam + rng.normal(scale=sig_lam, size=lam.shape), lo, hi)
    if sig_logj is not None:
        lam = isochoric(rng, lam, sig_logj, lo, hi)
    ax = rng.normal(size=(len(U), 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = rng.normal(scale=sig_rot, size=len(U))
    K = np.zeros((len(U), 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 2] = -ax[:, 2], ax[:, 1], -ax[:, 0]
    K -= K.transpose(0, 2, 1)
    s, c = np.sin(ang)[:, None, None], np.cos(ang)[:, None, None]
    R = np.eye(3)[None] + s * K + (1.0 - c) * (K @ K)  # Rodrigues
    Q = R @ Q
    return np.einsum("nij,nj,nkj->nik", Q, lam, Q)


def green_lagrange(F):
    F = np.asarray(F)
    return 0.5 * (np.einsum("...ki,...kj->...ij", F, F) - np.eye(3))


def to_voigt(T):
    return np.stack([T[..., i, j] for i, j in VOIGT], axis=-1)


def from_up21(c):
    """(..., 21) upper Voigt -> (..., 6, 6) symmetric."""
    c = np.asarray(c)
    out = np.zeros(c.shape[:-1] + (6, 6))
    k = 0
    for a in range(6):
        for b in range(a, 6):
            out[..., a, b] = out[..., b, a] = c[..., k]
            k += 1
    return out


def linear_part(Clin, F):
    """C_lin : E (Voigt, engineering shear in E) for F (..., 3, 3)."""
    E = green_lagrange(F)
    Ee = to_voigt(E) * np.array([1, 1, 1, 2, 2, 2.0])
    return np.einsum("...ab,...b->...a", from_up21(Clin), Ee)


def base_stress(Clin, F):
    """The surrogate's base S (Voigt): the Neo-Hookean model 3 with the Voigt-average bulk and
    shear moduli of C_lin, S = mu J^-2/3 (I - tr C / 3 C^-1) + J p C^-1, p = 3K J^-2/3 (J^1/3 - 1),
    plus (C_lin - C_iso) : E; the stress ceos_surrogate_eval_device adds the network's residual to."""
    Cl = from_up21(Clin)
    d = np.diagonal(Cl, axis1=-2, axis2=-1)
    K = (d[..., 0] + d[..., 1] + d[..., 2] + 2 * (Cl[..., 0, 1] + Cl[..., 0, 2] + Cl[..., 1, 2])) / 9.0
    mu = ((d[..., 0] + d[..., 1] + d[..., 2]) - (Cl[..., 0, 1] + Cl[..., 0, 2] + Cl[..., 1, 2]) +
          3 * (d[..., 3] + d[..., 4] + d[..., 5])) / 15.0
    lam = K - 2 * mu / 3
    Ciso = np.zeros(Cl.shape)
    Ciso[..., :3, :3] = lam[..., None, None]
    for a in range(3):
        Ciso[..., a, a] += 2 * mu
        Ciso[..., a + 3, a + 3] = mu
    Cg = np.einsum("...ki,...kj->...ij", F, F)
    J = np.linalg.det(F)
    Ci = np.linalg.inv(Cg)
    trC = np.trace(Cg, axis1=-2, axis2=-1)
    p = 3.0 * K * J ** (-2.0 / 3.0) * (np.cbrt(J) - 1.0)
    S = (mu * J ** (-2.0 / 3.0))[..., None, None] * (np.eye(3) - (trC / 3.0)[..., None, None] * Ci) + \
        (J * p)[..., None, None] * Ci
    return to_voigt(S) + linear_part(up21(Cl - Ciso), F)


def up21(M):
    """(..., 6, 6) -> (..., 21) upper Voigt."""
    return np.stack([M[..., a, b] for a in range(6) for b in range(a, 6)], axis=-1)


def stress_rel_error(S, Sref, floor=1e-300):
    """Tensor-Frobenius relative error of Voigt stresses (shear counted twice); reference stresses
    below `floor` count as `floor` (a ratio of round-off to round-off means nothing)."""
    w = np.array([1, 1, 1, 2, 2, 2.0])
    d = np.sqrt(np.sum(w * (S - Sref) ** 2, axis=-1))
    return d / np.maximum(np.sqrt(np.sum(w * Sref ** 2, axis=-1)), floor)


# ------------------------------------------------------------------------------------------
# descriptors and surrogate loading
# ------------------------------------------------------------------------------------------
def descriptors(batch, pcg_rtol=1e-8):
    """C_lin (n_rve, 21): the condensed tangent of every RVE at F = I, as upper Voigt dS/dE.
    Leaves the batch reset (undeformed) and with the default solver settings."""
    n = batch.n
    batch.reset()
    batch.set_solver(pcg_rtol=pcg_rtol)
    try:
        P, A, _ = batch.solve(np.repeat(np.eye(3)[None], n, 0), tangent=True)
    finally:
        batch.set_solver()
    dF = hiprt.DeviceArray.from_numpy(np.tile(np.eye(3).reshape(9), n))
    dP = hiprt.DeviceArray.from_numpy(np.ascontiguousarray(P.transpose(0, 2, 1).reshape(-1)))
    dA = hiprt.DeviceArray.from_numpy(np.ascontiguousarray(A.reshape(-1)))
    dext = hiprt.DeviceArray(n * gpu.EXT_PER_GP)
    gpu.rve_to_external_device(n, dF.ptr, dP.ptr, dA.ptr, dext.ptr)
    return dext.numpy().reshape(n, gpu.EXT_PER_GP)[:, 6:].copy()


def surrogate_cfg(meta):
    return {k: meta[k] for k in ("u_scale", "lam_lo", "lam_hi", "e_ref", "d_mean", "d_std", "d_lo", "d_hi",
                                 "t_mean", "t_std")}


def load_surrogate(path=SURROGATE_PATH, max_points=4096):
    """(Mlp, Surrogate, meta) from a trained surrogate file (numpy .npz, no pickles)."""
    with np.load(path, allow_pickle=False) as z:
        meta = {k: z[k] for k in z.files}
    width = int(meta["width"])
    mlp = gpu.Mlp(N_IN, width, N_OUT, max_batch=max_points, seed=1)
    mlp.set_params(meta["params"].astype(np.float32))
    sur = gpu.Surrogate(mlp, surrogate_cfg(meta), max_points)
    return mlp, sur, meta


# ------------------------------------------------------------------------------------------
# the macro problem
# ------------------------------------------------------------------------------------------
class MacroFE2:
    """n^3 hex8 macro cube; Gauss point g of element e is RVE 8 e + g of a seeded batch
    (sides drawn from `side_choices`, 20 % spherical inclusions, Neo-Hookean phases)."""

    def __init__(self, n_macro=8, seed=rve.SEED, side_choices=rve.SIDES, surrogate=None, X=None, conn=None):
        if X is None:
            X, conn = meshgen.cube(n_macro)
        self.X, self.conn = X, conn
        self.mesh = gpu.CeosMesh(X, conn)
        self.mesh.set_material(1, gpu.MODEL_EXT, [])
        self.nn, self.ne, self.ndof = self.mesh.nn, self.mesh.ne, self.mesh.ndof
        self.ngp = 8 * self.ne
        self.sides, self.phases = rve.batch(self.ngp, seed, side_choices)
        self.rves = gpu.RveBatch(self.sides, self.phases, gpu.MODEL_NH, PROPS)
        rm, col = self.mesh.csr()
        self.nnz = len(col)
        D = hiprt.DeviceArray
        self.dU = D(self.ndof)
        self.dF = D(9 * self.ngp)
        self.dP = D(9 * self.ngp)
        self.dA = D(81 * self.ngp)
        self.dinfo = D(4 * self.ngp, np.int32)
        self.dext = D(gpu.EXT_PER_GP * self.ngp)
        self.dFint = D(self.ndof)
        self.dK = D(self.nnz)
        self.dClin = D(21 * self.ngp)
        self.dflags = D(self.ngp, np.int32)
        self.stream = hiprt.Stream()
        self.mesh.set_external(self.dext.ptr)
        self.Clin = None
        self.sur = surrogate  # (Mlp, Surrogate, meta) or None
        self.judge_rng = np.random.default_rng(seed)
        self.judge_ids = None
        self.graphs = {}
        # set up the mesh's device tables now: graph capture must find nothing left to allocate
        self.mesh.prepare()
        self.dU.upload(np.zeros(self.ndof))

    def setup_descriptors(self, pcg_rtol=1e-8):
        t = time.perf_counter()
        self.Clin = descriptors(self.rves, pcg_rtol)
        self.dClin.upload(self.Clin.reshape(-1))
        self.stress_floor = 1e-9 * float(np.abs(self.Clin).max())  # the stress of a 1e-9 strain
        return time.perf_counter() - t

    def _kinematics(self, U):
        self.dU.upload(np.ascontiguousarray(U, dtype=np.float64))

    # device chains of one evaluation, all on self.stream: enqueue only (graph capturable)
    def _chain_kinematics(self):
        s = self.stream.ptr
        self.mesh.update_mesh_device(self.dU.ptr, s)
        self.mesh.gp_deformation_gradient_device(self.dU.ptr, self.dF.ptr, s)

    def _chain_rve(self):
        s = self.stream.ptr
        self._chain_kinematics()
        self.rves.solve_device(self.dF.ptr, self.dP.ptr, self.dA.ptr, self.dinfo.ptr, s)
        gpu.rve_to_external_device(self.ngp, self.dF.ptr, self.dP.ptr, self.dA.ptr, self.dext.ptr, stream=s)
        self._chain_assemble()

    def _chain_surrogate(self):
        self._chain_kinematics()
        self.sur[1].eval_device(self.ngp, self.dF.ptr, self.dClin.ptr, self.dext.ptr, self.dflags.ptr, None,
                                self.stream.ptr)

    def _chain_assemble(self):
        self.mesh.evaluate_device(0.0, self.dU.ptr, self.dFint.ptr, self.dK.ptr, self.stream.ptr)

    def _run(self, key, use_graph):
        """Enqueue a device chain, replaying its captured HIP graph when use_graph."""
        fn = getattr(self, "_chain_" + key)
        if not use_graph:
            fn()
            return
        g = self.graphs.get(key)
        if g is None:
            g = self.graphs[key] = hiprt.Graph(self.stream, fn)
        g.launch(self.stream)

    def _assemble(self, use_graph=False):
        self._run("assemble", use_graph)

    def _finish(self, stats):
        self.stream.synchronize()
        code, st = self.mesh.device_status()
        if code != 0:
            raise gpu.CeosError(code, "macro element evaluation")
        return stats

    def gp_F(self):
        """Gauss-point F of the last evaluation, (ngp, 3, 3)."""
        self.stream.synchronize()
        return self.dF.numpy().reshape(self.ngp, 3, 3).transpose(0, 2, 1)

    def ext(self):
        """(ngp, 27): S (Voigt) and dS/dE (upper 21) of the last evaluation."""
        self.stream.synchronize()
        return self.dext.numpy().reshape(self.ngp, gpu.EXT_PER_GP)

    def fint(self):
        self.stream.synchronize()
        return self.dFint.numpy()

    def tangent(self):
        self.stream.synchronize()
        return self.dK.numpy()

    def evaluate_rve(self, U, use_graph=False):
        """Plain macro evaluation: every RVE solved and condensed, then Fint and K (the device chain
        replayed from a HIP graph when use_graph)."""
        t0 = time.perf_counter()
        self._kinematics(U)
        self._run("rve", use_graph)
        self.stream.synchronize()
        info = self.dinfo.numpy().reshape(-1, 4)
        if np.any(info[:, 0] != 0):
            raise gpu.CeosError(gpu.ERR_LOCAL_NR, "RVE solve (%d points)" % int(np.sum(info[:, 0] != 0)))
        return self._finish({"mode": "rve", "seconds": time.perf_counter() - t0, "rve_solves": self.ngp,
                             "newton_mean": float(info[:, 1].mean()), "cg_total": int(info[:, 2].sum())})

    # ---- the macro Newton step on the device ------------------------------------------------
    def _chain_residual0(self):
        self.mesh.residual_bc_device(0, self.dU.ptr, self.dFint.ptr, self.dK.ptr, None, self.dUbar.ptr, self.dRes.ptr,
                                     self.dnorm.ptr, 2, self.stream.ptr)

    def _chain_residual(self):
        self.mesh.residual_bc_device(1, self.dU.ptr, self.dFint.ptr, self.dK.ptr, None, self.dUbar.ptr, self.dRes.ptr,
                                     self.dnorm.ptr, 2, self.stream.ptr)

    def _chain_solve(self):
        s = self.stream.ptr
        self.mesh.linear_solve_device_async(self.dK.ptr, self.dRes.ptr, self.dDX.ptr, self.dpcg.ptr, self.pcg_rtol,
                                            self.pcg_maxit, s)
        self.mesh.newton_update_device(self.dU.ptr, self.dDX.ptr, -1.0, s)

    def set_load_case(self, presc_dof, fixed_dof, pcg_rtol=1e-10, pcg_maxit=600):
        """Dirichlet DOFs (0-based) of the macro load case (ceos_bc_set) and the settings of the
        Newton step's PCG (ceos_linear_solve_device_async: pcg_maxit iterations enqueued, early
        exit on the device).  Drops graphs captured for an earlier load case."""
        self.presc = np.asarray(presc_dof, dtype=np.int64)
        self.mesh.set_bc(self.presc, fixed_dof)
        D = hiprt.DeviceArray
        if getattr(self, "dRes", None) is None:
            self.dRes, self.dDX, self.dUbar = D(self.ndof), D(self.ndof), D(self.ndof)
            self.dnorm, self.dpcg = D(3), D(3)
        self.pcg_rtol, self.pcg_maxit = float(pcg_rtol), int(pcg_maxit)
        for k in ("residual0", "residual", "solve"):
            self.graphs.pop(k, None)

    def newton_step(self, ubar, U0, mode="rve", tol=1e-6, maxit=12, use_graph=True, judge_frac=0.01,
                    judge_tol=0.01, judge_gate="aggregate"):
        """One macro load step by full Newton on the device, the loop of
        Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:44-168 with the reference's
        max-abs residual test (norm < tol norm_0 after the first iteration): per iteration the
        evaluation chain (`mode` "rve": every RVE solved and condensed; "surrogate": the surrogate
        and the RVE fallback outside the hull), the Dirichlet residual and
        its norm (ceos_residual_bc_device), then PCG and the update U -= DX
        (ceos_linear_solve_device_async, ceos_newton_update_device).  With use_graph every piece is
        a replayed HIP graph and the host synchronises once per iteration, to read the norm and
        the PCG outcome.  ubar: the prescribed DOFs' values at the end of the step (set_load_case
        order); U0: the displacement the step starts from.  In surrogate mode ceil(judge_frac
        ngp) points are re-solved at the converged state and gate the step as in
        evaluate_surrogate (judge_gate "aggregate" or "point"); a failed gate sends the step
        through the RVE path again from U0.  Returns (U, stats)."""
        if mode not in ("rve", "surrogate"):
            raise ValueError("mode must be 'rve' or 'surrogate'")
        if mode == "surrogate" and (self.sur is None or self.Clin is None):
            raise RuntimeError("surrogate mode needs a surrogate and setup_descriptors()")
        t0 = time.perf_counter()
        Ubar = np.zeros(self.ndof)
        Ubar[self.presc] = ubar
        self.dUbar.upload(Ubar)
        self._kinematics(U0)
        recs = []
        norm0 = None
        converged = False
        for it in range(maxit + 1):
            if it:
                self._run("solve", use_graph)
            if mode == "rve":
                self._run("rve", use_graph)
            else:
                self._run("surrogate", use_graph)
                self.stream.synchronize()
                flags = self.dflags.numpy()
                out = np.nonzero(flags)[0].astype(np.int32)
                if len(out):
                    self._fallback_points(out, "surrogate")
                self._assemble(use_graph)
            self._run("residual0" if it == 0 else "residual", use_graph)
            self.stream.synchronize()
            code, _ = self.mesh.device_status()
            if code != 0:
                raise gpu.CeosError(code, "macro element evaluation")
            nrm = float(self.dnorm.numpy()[2])
            rec = {"it": it, "norm": nrm}
            if it:
                pit, prr, pfl = self.dpcg.numpy()
                rec.update({"pcg_iterations": int(pit), "pcg_relres": float(prr), "pcg_flag": int(pfl)})
                if int(pfl) != 1:
                    raise RuntimeError("macro PCG flag %d at Newton iteration %d" % (int(pfl), it))
            if mode == "rve":
                info = self.dinfo.numpy().reshape(-1, 4)
                if np.any(info[:, 0] != 0):
                    raise gpu.CeosError(gpu.ERR_LOCAL_NR, "RVE solve (%d points)" % int(np.sum(info[:, 0] != 0)))
            recs.append(rec)
            norm0 = nrm if it == 0 else norm0
            if it > 0 and nrm < tol * norm0:
                converged = True
                break
        stats = {"mode": mode, "converged": converged, "iterations": len(recs) - 1, "history": recs}
        if converged and mode == "surrogate" and judge_frac > 0:
            ev, agg = self._judge(judge_frac)
            pmax = float(ev.max()) if len(ev) else 0.0
            stats.update({"judged": int(len(ev)), "judge_err": agg, "judge_point_max": pmax,
                          "judge_point_pass": pmax <= judge_tol})
            if len(ev) and (pmax if judge_gate == "point" else agg) > judge_tol:
                U, st2 = self.newton_step(ubar, U0, "rve", tol, maxit, use_graph)
                stats.update({"fallback": True, "rve_step": st2, "seconds": time.perf_counter() - t0})
                return U, stats
            stats["fallback"] = False
        stats["seconds"] = time.perf_counter() - t0
        self.stream.synchronize()
        return self.dU.numpy(), stats

    def _fallback_points(self, out, fallback_tangent):
        """The RVE solve's S (and with fallback_tangent "rve" its tangent) at the points `out`."""
        Fcm = self.dF.numpy().reshape(self.ngp, 9)
        Fm = Fcm.reshape(self.ngp, 3, 3).transpose(0, 2, 1)
        with_t = fallback_tangent == "rve"
        P, A, _ = self.rves.solve_subset(out, Fm[out], tangent=with_t)
        D = hiprt.DeviceArray.from_numpy
        dsel, dFs = D(out), D(np.ascontiguousarray(Fcm[out].reshape(-1)))
        dPs = D(np.ascontiguousarray(P.transpose(0, 2, 1).reshape(-1)))
        dAs = D(np.ascontiguousarray(A.reshape(-1))) if with_t else None
        gpu.rve_to_external_device(len(out), dFs.ptr, dPs.ptr, dAs.ptr if with_t else None, self.dext.ptr,
                                   dsel=dsel.ptr, stream=self.stream.ptr)
        self.stream.synchronize()

    def _judge(self, judge_frac, judge_newton_rtol=1e-3):
        """(relative stress errors, norm-wise estimate) of ceil(judge_frac ngp) sampled inside
        points, re-solved by the RVE path at the current macro state, against the records of the
        last evaluation."""
        self.stream.synchronize()
        flags = self.dflags.numpy()
        inside = np.nonzero(flags == 0)[0]
        nj = min(len(inside), int(math.ceil(judge_frac * self.ngp)))
        if nj == 0:
            return np.zeros(0), 0.0
        if self.judge_ids is None or len(self.judge_ids) != nj:
            self.judge_ids = np.sort(self.judge_rng.choice(inside, size=nj, replace=False)).astype(np.int32)
        jid = self.judge_ids[flags[self.judge_ids] == 0]
        Fm = self.dF.numpy().reshape(self.ngp, 3, 3).transpose(0, 2, 1)
        self.rves.set_solver(newton_rtol=judge_newton_rtol)
        try:
            Pj, _, _ = self.rves.solve_subset(jid, Fm[jid])
        finally:
            self.rves.set_solver()
        Sj = np.linalg.solve(Fm[jid], Pj)
        S_rve = to_voigt(0.5 * (Sj + Sj.transpose(0, 2, 1)))
        S_all = self.ext()[:, :6]
        w = np.array([1, 1, 1, 2, 2, 2.0])
        num = np.mean(np.sum(w * (S_all[jid] - S_rve) ** 2, axis=1))
        den = np.mean(np.sum(w * S_all ** 2, axis=1))
        agg = float(np.sqrt(num / max(den, self.stress_floor ** 2)))
        return stress_rel_error(S_all[jid], S_rve, self.stress_floor), agg

    def new_judge_sample(self):
        """Draw the judged points again at the next surrogate evaluation (e.g. per load step);
        until then the same points are re-solved, each from its own last converged state."""
        self.judge_ids = None

    def evaluate_surrogate(self, U, judge_frac=0.01, judge_tol=0.01, fallback_tangent="surrogate",
                           judge_newton_rtol=1e-3, use_graph=False, judge_gate="aggregate"):
        """Surrogate macro evaluation: the MLP at points inside the training hull; outside it the
        RVE solve gives the stress (with fallback_tangent="surrogate" the point keeps the
        surrogate's tangent, an inexact macro Newton tangent; "rve" condenses the RVE tangent too);
        ceil(judge_frac ngp) sampled inside points are re-solved as judges.  The judges gate the
        evaluation: with judge_gate="aggregate" the norm-wise estimate above judge_tol (the RMS
        error of the judged points over the RMS stress of all points, tensor norms), with "point"
        (the per-point bar, max over points of |S_nn - S_rve| / |S_rve| <= judge_tol) any judged
        point above judge_tol, makes the whole evaluation fall back to the plain path.  judge_err
        reports the norm-wise estimate, judge_point_max the largest judged error, judge_pass
        whether the chosen gate held and judge_point_pass whether the per-point bar held.
        Judged RVEs are converged to judge_newton_rtol (relative Newton test): 1e-3 moves
        the estimate by about 1e-5 against fully converged judges (the Newton iterates converge
        quadratically, so the last one is far more accurate than the test), three orders below the
        1 % judge tolerance."""
        if self.sur is None or self.Clin is None:
            raise RuntimeError("surrogate evaluation needs a surrogate and setup_descriptors()")
        t0 = time.perf_counter()
        s = self.stream.ptr
        self._kinematics(U)
        self._run("surrogate", use_graph)
        self.stream.synchronize()
        flags = self.dflags.numpy()
        Fcm = self.dF.numpy().reshape(self.ngp, 9)
        Fm = Fcm.reshape(self.ngp, 3, 3).transpose(0, 2, 1)
        out = np.nonzero(flags)[0].astype(np.int32)
        if len(out):
            with_t = fallback_tangent == "rve"
            P, A, _ = self.rves.solve_subset(out, Fm[out], tangent=with_t)
            m = len(out)
            D = hiprt.DeviceArray.from_numpy
            dsel, dFs = D(out), D(np.ascontiguousarray(Fcm[out].reshape(-1)))
            dPs = D(np.ascontiguousarray(P.transpose(0, 2, 1).reshape(-1)))
            dAs = D(np.ascontiguousarray(A.reshape(-1))) if with_t else None
            gpu.rve_to_external_device(m, dFs.ptr, dPs.ptr, dAs.ptr if with_t else None, self.dext.ptr,
                                       dsel=dsel.ptr, stream=s)
        # the macro assembly runs while the judged points are re-solved
        self._assemble(use_graph)
        inside = np.nonzero(flags == 0)[0]
        nj = min(len(inside), int(math.ceil(judge_frac * self.ngp))) if judge_frac > 0 else 0
        if nj and (self.judge_ids is None or len(self.judge_ids) != nj):
            self.judge_ids = np.sort(self.judge_rng.choice(inside, size=nj, replace=False)).astype(np.int32)
        err, agg = np.zeros(0), 0.0
        if nj:
            jid = self.judge_ids[flags[self.judge_ids] == 0]
            self.rves.set_solver(newton_rtol=judge_newton_rtol)
            try:
                Pj, _, _ = self.rves.solve_subset(jid, Fm[jid])
            finally:
                self.rves.set_solver()
            Sj = np.linalg.solve(Fm[jid], Pj)
            S_rve = to_voigt(0.5 * (Sj + Sj.transpose(0, 2, 1)))
            S_all = self.ext()[:, :6]
            S_nn = S_all[jid]
            err = stress_rel_error(S_nn, S_rve, self.stress_floor)
            # sample estimate of the norm-wise error over all points: RMS error of the judged points
            # over the RMS stress of all points (a judged point with a tiny stress cannot blow it up)
            w = np.array([1, 1, 1, 2, 2, 2.0])
            num = np.mean(np.sum(w * (S_nn - S_rve) ** 2, axis=1))
            den = np.mean(np.sum(w * S_all ** 2, axis=1))
            agg = float(np.sqrt(num / max(den, self.stress_floor ** 2)))
        stats = {"mode": "surrogate", "outside_hull": int(len(out)), "judged": int(len(err)),
                 "judge_err": agg, "judge_point_max": float(err.max()) if len(err) else 0.0,
                 "judge_point_mean": float(err.mean()) if len(err) else 0.0, "fallback": False}
        if judge_gate not in ("point", "aggregate"):
            raise ValueError("judge_gate must be 'point' or 'aggregate'")
        gate = (err.max() if len(err) else 0.0) if judge_gate == "point" else agg
        stats["judge_pass"] = bool(gate <= judge_tol)
        stats["judge_point_pass"] = bool((err.max() if len(err) else 0.0) <= judge_tol)
        if len(err) and gate > judge_tol:
            r = self.evaluate_rve(U, use_graph)
            stats["fallback"] = True
            stats["rve_solves"] = r["rve_solves"] + int(len(out)) + len(err)
            stats["seconds"] = time.perf_counter() - t0
            return stats
        self._finish(stats)
        stats["rve_solves"] = int(len(out)) + len(err)
        stats["seconds"] = time.perf_counter() - t0
        return stats
