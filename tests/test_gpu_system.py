"""The macro Newton step on the device (ceos_bc_set, ceos_residual_bc(_device),
ceos_linear_solve(_device), ceos_newton_update_device) against the test driver's restatement of
the reference's boundary-condition and Newton code (tests/fe_driver.py: apply_dirichlet and the
max-abs norm, pinned by the reference's own residual histories in test_oracle_golden.py) and
against sparse direct solves.

Bar: K after the boundary conditions bit-exact (cleared rows and columns, unit diagonal, the
rest untouched); R bit-exact where no product enters (it > 0) and within 1e-12 norm-wise at
it = 0 (the product K (Ubar - U) sums in another order); PCG solutions with ||K x - b|| <= rtol
||b|| measured on the host in FP64; whole Newton histories through the device system with the
iteration counts of the oracle path and converged displacements within 1e-9 relative.
"""
import numpy as np
import pytest

import fe_driver as fd
import oracle_lib as ol
from ceos_amd import gpu, hiprt, meshgen

pytestmark = pytest.mark.gpu

LAM = 100.0 - 2.0 / 3.0


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


class DeviceSystem:
    """A CeosMesh whose Newton system (BC, norm, linear solve) runs in the library.

    rtol_minres: None = the library's default solver (PCG, then MINRES at the same rtol when
    PCG meets an indefinite tangent).  A number: PCG at rtol, and after a PCG breakdown MINRES
    at rtol_minres; MINRES measures the residual in the M^-1 norm, which levels off near 1e-13
    on the stiffest systems here, while PCG's 2-norm goes further."""

    def __init__(self, mesh, rtol=1e-12, norm_type=2, rtol_minres=None):
        self.m = mesh
        self.ndof = mesh.ndof
        self.rtol = rtol
        self.rtol_minres = rtol_minres
        self.norm_type = norm_type
        self._bc = None
        self.pcg = []
        self.minres = 0
        if rtol_minres is not None:
            mesh.set_linear_solver(gpu.SOLVER_PCG)

    def __getattr__(self, k):
        return getattr(self.m, k)

    def newton_system(self, it, fext, bc, Ubar, X):
        if self._bc is not bc:
            self.m.set_bc(bc.presc_dof, bc.fixed)
            self._bc = bc
        return self.m.residual_bc(it, fext, Ubar, self.norm_type)

    def newton_solve(self, b):
        x, its, rr, flag = self.m.linear_solve(b, self.rtol, 50000)
        if flag == 2 and self.rtol_minres is not None:
            self.m.set_linear_solver(gpu.SOLVER_MINRES)
            x, its2, rr, flag = self.m.linear_solve(b, self.rtol_minres, 50000)
            self.m.set_linear_solver(gpu.SOLVER_PCG)
            its += its2
        assert flag in (1, 4), (its, rr, flag)
        self.pcg.append(its)
        self.minres += flag == 4
        return x


def _scrambled(n, seed):
    X, conn = meshgen.cube(n, perturb=0.12, seed=seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(X))
    return X[perm], (inv[conn - 1] + 1)[rng.permutation(len(conn))].astype(np.int32)


def _bc_case(X, rng):
    """x=0 face fixed, x=1 face ux prescribed, plus a few DOFs that are both, and loads."""
    fixed, presc = meshgen.stretch_bc(X)
    both = rng.choice(presc, size=2, replace=False)
    fixed = np.concatenate([fixed, both])
    ndof = 3 * len(X)
    loads = [(int(d), (lambda t, v=v: v * t)) for d, v in zip(rng.choice(ndof, 5, replace=False), rng.normal(size=5))]
    return fd.BoundaryConditions(ndof, fixed, [(int(d), (lambda t: 0.1 * t)) for d in presc], loads)


@pytest.mark.parametrize("precision", [gpu.K_FP64, gpu.K_FP32_STORED])
@pytest.mark.parametrize("it", [0, 1])
def test_residual_and_matrix_bc_match_driver(precision, it):
    rng = np.random.default_rng(3 + it)
    X, conn = _scrambled(5, 7)
    g = gpu.CeosMesh(X, conn, precision=precision)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    bc = _bc_case(X, rng)
    U = rng.normal(scale=0.01, size=g.ndof)
    Ubar = U.copy()
    Ubar[bc.presc_dof] += 0.05
    fext = bc.fext(0.7)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    Fint, _ = g.internal_force()
    K = g.tangent()
    rm, col = g.csr()
    Kb, Rb = fd.apply_dirichlet(rm, col, K, Fint - fext, bc, Ubar, U, it)
    g.set_bc(bc.presc_dof, bc.fixed)
    for norm_type in (2, 1):
        R, nrm = g.residual_bc(it, fext, Ubar, norm_type)
        if it == 0:
            assert rel(R, Rb) < 1e-12
        else:
            assert np.array_equal(R, Rb)
        ref = np.max(np.abs(Rb)) if norm_type == 2 else np.linalg.norm(Rb)
        assert abs(nrm - ref) <= 1e-12 * ref
    # the matrix carries the BC once (a second application changes nothing)
    assert np.array_equal(g.tangent(), Kb)
    assert np.any(Kb != K)


def test_node_mask_skips_nothing_on_a_fully_constrained_node():
    """Every DOF of a node constrained, one DOF of another: rows and columns of exactly those."""
    X, conn = meshgen.cube(3, perturb=0.1, seed=2)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_NH, [0.5, 100.0])
    U = np.random.default_rng(2).normal(scale=0.01, size=g.ndof)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    K = g.tangent()
    rm, col = g.csr()
    bc = fd.BoundaryConditions(g.ndof, [30, 31, 32, 7], [])
    Kb, _ = fd.apply_dirichlet(rm, col, K, np.zeros(g.ndof), bc, U, U, 1)
    g.set_bc([], [30, 31, 32, 7])
    g.residual_bc(1)
    assert np.array_equal(g.tangent(), Kb)


@pytest.mark.parametrize("solver_precision", [gpu.K_FP64, gpu.K_FP32_STORED])
def test_pcg_solves_the_bc_system(solver_precision):
    X, conn = _scrambled(6, 11)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    rng = np.random.default_rng(5)
    # smooth tension plus small noise: K stays positive definite (random strains of a few per
    # cent would give pressures far above mu and an indefinite geometric stiffness)
    U = np.stack([0.04 * X[:, 0], -0.01 * X[:, 1], -0.01 * X[:, 2]], axis=1).reshape(-1)
    U = U + rng.normal(scale=1e-4, size=g.ndof)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    fixed, presc = meshgen.stretch_bc(X)
    g.set_bc(presc, fixed)
    g.residual_bc(1)
    Kb = g.tangent()
    rm, col = g.csr()
    b = rng.normal(size=g.ndof)
    g.set_solver_precision(solver_precision)
    rtol = 1e-10 if solver_precision == gpu.K_FP64 else 1e-5
    x, its, rr, flag = g.linear_solve(b, rtol, 20000)
    res = fd.sym_upper_matvec(rm, col, Kb, x) - b
    print("PCG %d iterations, recursive %.2e, true %.2e" % (its, rr, np.linalg.norm(res) / np.linalg.norm(b)))
    assert flag == 1 and rr <= rtol
    assert np.linalg.norm(res) <= 2 * rtol * np.linalg.norm(b)
    xd = fd.spla.spsolve(fd.sym_from_upper(rm, col, Kb, g.ndof), b)
    assert rel(x, xd) < (1e-7 if solver_precision == gpu.K_FP64 else 1e-2)
    # iteration limit and zero right-hand side
    _, its, _, flag = g.linear_solve(b, 1e-14, 3)
    assert flag == 3 and its == 3
    x0, its, _, flag = g.linear_solve(np.zeros(g.ndof), 1e-10, 100)
    assert flag == 1 and its == 0 and not x0.any()


def _indefinite_system():
    """Random nodal strains of ~6 % under lambda/mu = 99: compressive pressures far above mu
    make the geometric stiffness indefinite."""
    X, conn = _scrambled(6, 11)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    rng = np.random.default_rng(5)
    U = rng.normal(scale=0.01, size=g.ndof)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    fixed, presc = meshgen.stretch_bc(X)
    g.set_bc(presc, fixed)
    g.residual_bc(1)
    rm, col = g.csr()
    Kb = g.tangent()
    A = fd.sym_from_upper(rm, col, Kb, g.ndof)
    ev = np.linalg.eigvalsh(A.toarray())
    assert ev[0] < 0.0 < ev[-1]
    return g, A, rng.normal(size=g.ndof)


def test_pcg_reports_an_indefinite_tangent():
    """PCG alone stops with the breakdown flag on an indefinite K, no NaN."""
    g, A, b = _indefinite_system()
    g.set_linear_solver(gpu.SOLVER_PCG)
    x, its, rr, flag = g.linear_solve(b, 1e-10, 20000)
    assert flag == 2 and np.all(np.isfinite(x))


@pytest.mark.parametrize("method", [gpu.SOLVER_AUTO, gpu.SOLVER_MINRES])
def test_minres_solves_an_indefinite_tangent(method):
    """The default solver switches to MINRES after the PCG breakdown; MINRES alone solves it too.
    The true residual and the direct solution are checked on the host."""
    g, A, b = _indefinite_system()
    g.set_linear_solver(method)
    x, its, rr, flag = g.linear_solve(b, 1e-12, 50000)
    res = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    xd = fd.spla.spsolve(A, b)
    print("MINRES (method %d): %d iterations, M^-1-norm %.2e, true %.2e" % (method, its, rr, res))
    assert flag == 4 and rr <= 1e-12
    assert res < 1e-9 and rel(x, xd) < 1e-8


def test_minres_on_a_positive_definite_tangent_matches_pcg():
    X, conn = _scrambled(5, 3)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    U = np.stack([0.03 * X[:, 0], -0.01 * X[:, 1], -0.01 * X[:, 2]], axis=1).reshape(-1)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    fixed, presc = meshgen.stretch_bc(X)
    g.set_bc(presc, fixed)
    g.residual_bc(1)
    b = np.random.default_rng(2).normal(size=g.ndof)
    x1, _, _, f1 = g.linear_solve(b, 1e-12, 20000)
    g.set_linear_solver(gpu.SOLVER_MINRES)
    x2, _, _, f2 = g.linear_solve(b, 1e-12, 20000)
    assert f1 == 1 and f2 == 4 and rel(x2, x1) < 1e-8


def test_golden_svk_case_through_device_system():
    """The reference's StVenant Checker case (residual history of its log, Marc displacements)
    with the boundary conditions, norms and linear solves of the library."""
    from test_oracle_golden import _case, check_svk, run_svk_case
    case = _case("svk_hex8_case.json")
    g = gpu.CeosMesh(np.array(case["coords"]), np.array(case["conn"], dtype=np.int32))
    g.set_material(1, ol.MODEL_SVK, [case["material"]["E"], case["material"]["nu"]])
    ds = DeviceSystem(g, rtol=1e-13)
    case, norms, probe = run_svk_case(ds)
    check_svk(case, norms, probe)
    print("PCG iterations per solve: max %d mean %.0f" % (max(ds.pcg), np.mean(ds.pcg)))


def test_golden_q1p0_case_through_device_system():
    """The reference's mean-dilatation Checker case (K / mu = 1000) with the library's BC,
    norms and linear solves: the logged residual history and the Marc displacements.  The first
    Newton iterate of most steps of this case has an indefinite tangent (step 2, iteration 1: one
    negative eigenvalue); there PCG breaks down and the library's default continues with MINRES."""
    from test_oracle_golden import _case, check_nhq1p0, run_nhq1p0_case
    case = _case("nhq1p0_hex8_case.json")
    g = gpu.CeosMesh(np.array(case["coords"]), np.array(case["conn"], dtype=np.int32),
                     elem_tech=gpu.ELEM_MEAN_DILATATION)
    g.set_material(1, ol.MODEL_NH_Q1P0, [case["material"]["C10"], case["material"]["K"]])
    ds = DeviceSystem(g, rtol=1e-14, rtol_minres=1e-12)
    case, norms, probe = run_nhq1p0_case(ds)
    check_nhq1p0(case, norms, probe)
    print("Krylov iterations per solve: max %d mean %.0f; %d of %d solves finished by MINRES"
          % (max(ds.pcg), np.mean(ds.pcg), ds.minres, len(ds.pcg)))
    assert ds.minres > 0


def test_c1_newton_history_through_device_system_matches_oracle():
    """10^3 cube, compressible Neo-Hookean, stretch 1.2 in 10 steps: the oracle with the driver's
    BC and direct solves against the library with its own BC, norm and PCG."""
    X, conn = meshgen.cube(10)
    fixed, presc = meshgen.stretch_bc(X)
    bc = fd.BoundaryConditions(3 * len(X), fixed, [(int(d), (lambda t: 0.2 * t)) for d in presc])
    o = ol.OracleMesh(X, conn)
    o.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    ds = DeviceSystem(g, rtol=1e-12)
    ho = fd.quasi_static(o, bc, [(0.0, 1.0, 10)], nl=True, tol=1e-8, maxiter=15)
    hg = fd.quasi_static(ds, bc, [(0.0, 1.0, 10)], nl=True, tol=1e-8, maxiter=15)
    assert [h[4] for h in hg] == [h[4] for h in ho]
    for a, b in zip(hg, ho):
        assert rel(a[3], b[3]) < 1e-9
    print("PCG iterations per solve: max %d mean %.0f" % (max(ds.pcg), np.mean(ds.pcg)))


def test_c1_newton_history_with_two_level_pcg_and_reused_coarse_inverse():
    """The same C1 history with the two-level preconditioner whose coarse inverse serves 2 more
    solves after the one that built it: the oracle's Newton iteration counts and displacements."""
    X, conn = meshgen.cube(10)
    fixed, presc = meshgen.stretch_bc(X)
    bc = fd.BoundaryConditions(3 * len(X), fixed, [(int(d), (lambda t: 0.2 * t)) for d in presc])
    o = ol.OracleMesh(X, conn)
    o.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    g.set_preconditioner(gpu.PREC_TWO_LEVEL, 3.0, 2)
    ds = DeviceSystem(g, rtol=1e-12)
    ho = fd.quasi_static(o, bc, [(0.0, 1.0, 10)], nl=True, tol=1e-8, maxiter=15)
    hg = fd.quasi_static(ds, bc, [(0.0, 1.0, 10)], nl=True, tol=1e-8, maxiter=15)
    assert [h[4] for h in hg] == [h[4] for h in ho]
    for a, b in zip(hg, ho):
        assert rel(a[3], b[3]) < 1e-9
    print("two-level PCG iterations per solve: max %d mean %.0f; aggregates %s" % (
        max(ds.pcg), np.mean(ds.pcg), g.preconditioner_info()), flush=True)


def test_j2_cyclic_history_through_device_system_matches_oracle():
    """J2 with linear hardening on a 6^3 cube, linear analysis, u_x on x=1 cycling
    0 -> +2 % -> -2 % -> 0 over 16 steps: the oracle with the driver's BC and direct solves
    against the library's BC, norm and PCG (the path of the J2 workload).  The Gauss points'
    plastic state enters every residual, so identical iteration counts and displacements pin
    the device system on a history-dependent linear analysis."""
    X, conn = meshgen.cube(6)
    props = [210000.0, 0.3, 250.0, 1000.0]
    fixed, presc = meshgen.stretch_bc(X)
    table = [(0.0, 0.0), (1.0, 0.02), (3.0, -0.02), (4.0, 0.0)]
    bc = fd.BoundaryConditions(3 * len(X), fixed, [(int(d), (lambda t: fd.interp_table(table, t))) for d in presc])
    o = ol.OracleMesh(X, conn, nl_analysis=False)
    g = gpu.CeosMesh(X, conn, nl_analysis=False)
    for m in (o, g):
        m.set_material(1, ol.MODEL_J2, props)
    ds = DeviceSystem(g, rtol=1e-12)
    ho = fd.quasi_static(o, bc, [(0.0, 4.0, 16)], nl=False, tol=1e-10, maxiter=15)
    hg = fd.quasi_static(ds, bc, [(0.0, 4.0, 16)], nl=False, tol=1e-10, maxiter=15)
    assert [h[4] for h in hg] == [h[4] for h in ho]
    for a, b in zip(hg, ho):
        assert rel(a[3], b[3]) < 1e-9
    sg, so = g.state(0), o.state(0)
    assert rel(sg[..., :7], so[..., :7]) < 1e-9
    assert sg[..., 6].max() > 0.0


def test_device_resident_newton_matches_host_calls():
    """One load step by device arrays only (evaluate, BC + norm, PCG, update) against the same
    step by the host calls: identical iterates."""
    X, conn = meshgen.cube(6, perturb=0.1, seed=4)
    fixed, presc = meshgen.stretch_bc(X)
    ndof = 3 * len(X)
    Ubar = np.zeros(ndof)
    Ubar[presc] = 0.05

    def mk():
        m = gpu.CeosMesh(X, conn)
        m.set_material(1, ol.MODEL_CNH, [1.0, LAM])
        m.set_bc(presc, fixed)
        return m

    # host calls
    h = mk()
    U = np.zeros(ndof)
    host = []
    for it in range(4):
        assert h.solve_constitutive(1.0, U) == 0
        R, nrm = h.residual_bc(it, None, Ubar, 2)
        x, its, _, flag = h.linear_solve(-R, 1e-12, 20000)
        U = U + x
        h.update_mesh(U)
        host.append((nrm, U.copy()))
    # device arrays
    d = mk()
    _, nnz = d.info()
    D = hiprt.DeviceArray
    dU, dF, dR, dx = D(ndof), D(ndof), D(ndof), D(ndof)
    dK = D(nnz)
    dUbar = D.from_numpy(Ubar)
    dn = D(3)
    dU.upload(np.zeros(ndof))
    for it in range(4):
        d.evaluate_device(1.0, dU.ptr, dF.ptr, dK.ptr)
        d.residual_bc_device(it, dU.ptr, dF.ptr, dK.ptr, None, dUbar.ptr, dR.ptr, dn.ptr)
        code, _ = d.device_status()
        assert code == 0
        d.linear_solve_device(dK.ptr, dR.ptr, dx.ptr, 1e-12, 20000)
        d.newton_update_device(dU.ptr, dx.ptr, -1.0)
        hiprt.synchronize()
        nrm = dn.numpy()[2]
        assert nrm == host[it][0], (it, nrm, host[it][0])
        assert np.array_equal(dU.numpy(), host[it][1]), it
    norms = [n for n, _ in host]
    assert all(b < a for a, b in zip(norms, norms[1:])) and norms[-1] < 1e-5 * norms[0], norms


def test_entry_points_refuse_bad_arguments():
    """Out-of-range DOFs, a residual before ceos_bc_set, a non-positive tolerance and an unknown
    solver or precision are refused with CEOS_ERR_ARG, and the handle stays usable."""
    X, conn = meshgen.cube(2)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    U = np.zeros(g.ndof)
    assert g.solve_constitutive(0.0, U) == 0
    with pytest.raises(gpu.CeosError) as e:
        g.residual_bc(1)  # no ceos_bc_set yet
    assert e.value.code == gpu.ERR_ARG
    for bad in ([g.ndof], [-1]):
        with pytest.raises(gpu.CeosError) as e:
            g.set_bc(bad, [])
        assert e.value.code == gpu.ERR_ARG
    fixed, presc = meshgen.stretch_bc(X)
    g.set_bc(presc, fixed)
    with pytest.raises(gpu.CeosError) as e:
        g.residual_bc(0)  # prescribed DOFs at it = 0 need Ubar
    assert e.value.code == gpu.ERR_ARG
    R, nrm = g.residual_bc(0, None, np.zeros(g.ndof))
    assert nrm == 0.0 and not R.any()
    for call in (lambda: g.linear_solve(np.ones(g.ndof), 0.0), lambda: g.set_linear_solver(7),
                 lambda: g.set_solver_precision(5)):
        with pytest.raises(gpu.CeosError) as e:
            call()
        assert e.value.code == gpu.ERR_ARG
    x, its, _, flag = g.linear_solve(np.ones(g.ndof), 1e-10)
    assert flag == 1 and np.all(np.isfinite(x))


def _tension_system(n, seed, precision=gpu.K_FP64):
    """A scrambled-numbering cube in smooth tension with the stretch-test BC applied to K."""
    X, conn = _scrambled(n, seed)
    g = gpu.CeosMesh(X, conn, precision=precision)
    g.set_material(1, ol.MODEL_CNH, [1.0, LAM])
    rng = np.random.default_rng(seed)
    U = np.stack([0.04 * X[:, 0], -0.01 * X[:, 1], -0.01 * X[:, 2]], axis=1).reshape(-1)
    U = U + rng.normal(scale=1e-4, size=g.ndof)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    fixed, presc = meshgen.stretch_bc(X)
    g.set_bc(presc, fixed)
    g.residual_bc(1)
    return g, X, np.concatenate([fixed, presc]), rng


def _prolongator(agg, B):
    import scipy.sparse as sp
    nn = len(agg)
    rows = np.repeat(np.arange(3 * nn), 6)
    cols = (6 * np.repeat(agg, 18) + np.tile(np.arange(6), 3 * nn))
    return sp.csr_matrix((B.reshape(-1), (rows, cols)), shape=(3 * nn, 6 * (agg.max() + 1)))


def test_two_level_coarse_space_is_the_galerkin_product():
    """The coarse matrix of the two-level preconditioner equals P^T K P of the BC-applied tangent
    (sparse products on the host, 1e-12); P is orthonormal per aggregate (to its FP32 storage),
    zero on the constrained DOFs, and spans each aggregate's rigid-body modes restricted to its
    free DOFs."""
    g, X, cons, rng = _tension_system(8, 13)
    g.set_preconditioner(gpu.PREC_TWO_LEVEL, 3.0)
    _, its, _, flag = g.linear_solve(rng.normal(size=g.ndof), 1e-10, 20000)
    assert flag == 1
    agg, B, Ac = g.coarse_space()
    na, nc = g.preconditioner_info()
    assert nc == 6 * na == len(Ac) and agg.max() + 1 == na
    rm, col = g.csr()
    A = fd.sym_from_upper(rm, col, g.tangent(), g.ndof)
    P = _prolongator(agg, B)
    Ag = (P.T @ A @ P).toarray()
    dead = np.asarray(np.abs(P).sum(axis=0)).ravel() == 0
    Ag[dead, dead] = 1.0
    print("aggregates %d, coarse DOFs %d (%d empty modes), |Ac - P^T K P| / |P^T K P| %.2e" % (
        na, nc, dead.sum(), rel(Ac, Ag)), flush=True)
    assert np.abs(Ac - Ag).max() <= 1e-12 * np.abs(Ag).max()
    PtP = (P.T @ P).toarray()
    assert np.abs(PtP - np.diag((~dead).astype(float))).max() < 1e-6  # P is stored in FP32
    assert np.abs(B.reshape(-1, 6)[cons]).max() == 0.0
    free = np.ones(g.ndof)
    free[cons] = 0.0
    for I in range(na):
        idx = np.nonzero(agg == I)[0]
        d = X[idx] - X[idx].mean(axis=0)
        M = np.zeros((len(idx), 3, 6))
        for k in range(3):
            M[:, k, k] = 1.0
        M[:, 1, 3], M[:, 2, 3] = -d[:, 2], d[:, 1]
        M[:, 0, 4], M[:, 2, 4] = d[:, 2], -d[:, 0]
        M[:, 0, 5], M[:, 1, 5] = -d[:, 1], d[:, 0]
        M = (M * free.reshape(-1, 3)[idx][:, :, None]).reshape(-1, 6)
        Q = B[idx].reshape(-1, 6)
        assert np.abs(M - Q @ (Q.T @ M)).max() < 1e-6 * max(np.abs(M).max(), 1.0)


@pytest.mark.parametrize("precision", [gpu.K_FP64, gpu.K_FP32_STORED])
def test_two_level_pcg_matches_direct_and_restatement(precision):
    """Two-level PCG: the direct solution to the solve tolerance, bitwise the same on a repeat, fewer
    iterations than block Jacobi (this small, nearly incompressible case: about half; the count
    stays flat as the mesh grows at a fixed aggregate size), and the iteration count of a numpy
    restatement of the same preconditioned CG (block Jacobi + P A_c^-1 P^T with the library's A_c
    and P)."""
    g, X, cons, rng = _tension_system(14, 17, precision)
    rm, col = g.csr()
    Kb = g.tangent().astype(np.float64)
    A = fd.sym_from_upper(rm, col, Kb, g.ndof).tocsr()
    b = rng.normal(size=g.ndof)
    rtol = 1e-10 if precision == gpu.K_FP64 else 1e-5
    _, its_bj, _, fl = g.linear_solve(b, rtol, 50000)
    assert fl == 1
    g.set_preconditioner(gpu.PREC_TWO_LEVEL, 4.0)
    x, its, rr, fl = g.linear_solve(b, rtol, 50000)
    x2, its2, _, _ = g.linear_solve(b, rtol, 50000)
    xd = fd.spla.spsolve(A.tocsc(), b)
    agg, B, Ac = g.coarse_space()
    P = _prolongator(agg, B)
    Aci = np.linalg.inv(Ac)
    Dn = np.zeros((g.nn, 3, 3))
    for k in range(3):
        for l in range(3):
            Dn[:, k, l] = A[3 * np.arange(g.nn) + k, 3 * np.arange(g.nn) + l].A1
    Dinv = np.linalg.inv(Dn)

    def M(r):
        return np.einsum("nij,nj->ni", Dinv, r.reshape(-1, 3)).ravel() + P @ (Aci @ (P.T @ r))

    xr, r = np.zeros(g.ndof), b.copy()
    z = M(r)
    p, rz, its_np = z.copy(), r @ z, 0
    while r @ r > rtol ** 2 * (b @ b) and its_np < 50000:
        q = A @ p
        al = rz / (p @ q)
        xr += al * p
        r -= al * q
        z = M(r)
        rz, rz_old = r @ z, rz
        p = z + (rz / rz_old) * p
        its_np += 1
    print("iterations: block Jacobi %d, two-level %d (restatement %d); |x - x_direct| / |x| %.2e" % (
        its_bj, its, its_np, rel(x, xd)), flush=True)
    assert fl == 1 and rr <= rtol
    assert rel(x, xd) < (1e-7 if precision == gpu.K_FP64 else 1e-2)
    assert np.array_equal(x, x2) and its == its2
    assert its < 0.75 * its_bj
    assert abs(its - its_np) <= max(2, its // 20)
