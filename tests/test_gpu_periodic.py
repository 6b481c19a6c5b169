"""Periodic RVEs on general hex8 meshes on the GPU (ceos_amd.periodic: the library's T products,
boundary conditions, PCG and element evaluation) against: the closed form on a homogeneous RVE;
the voxel RVE batch (an independent implementation on the torus grid) on the same voxel
microstructure, stress and condensed tangent; the CPU restatement with the oracle's element loops
on a perturbed two-phase mesh whose face nodes are off the grid; finite differences of the
homogenized stress; and the periodicity of the converged displacement."""
import numpy as np
import pytest

import oracle_lib as ol
import periodic_oracle as po
from ceos_amd import gpu, meshgen, periodic, rve

pytestmark = pytest.mark.gpu

MAT = {1: (gpu.MODEL_NH, rve.MATRIX_PROPS), 2: (gpu.MODEL_NH, rve.INCLUSION_PROPS)}
F1 = np.array([[1.06, 0.04, -0.02], [0.01, 0.97, 0.03], [0.02, -0.01, 1.01]])
F2 = np.array([[1.12, 0.07, -0.03], [0.02, 0.93, 0.05], [0.03, -0.02, 1.04]])


def _pmat(P9):
    return np.asarray(P9).reshape(3, 3).T


def _two_phase(n, rve_id, perturb):
    X, conn = meshgen.cube(n, perturb=perturb, seed=11)
    if perturb:
        X = po.periodic_perturbation(X, 1.0 / n)
    ph = rve.phases(n, rve_id)
    return X, conn, ph.astype(np.int32) + 1, ph


def test_homogeneous_rve_gives_the_material_response():
    X, conn = meshgen.cube(4, perturb=0.2, seed=3)
    X = po.periodic_perturbation(X, 0.25)
    r = periodic.PeriodicRve(X, conn, None, {1: (gpu.MODEL_NH, rve.MATRIX_PROPS)})
    P, A, info = r.solve(F1)
    s, _, _ = ol.material_point(ol.MODEL_NH, rve.MATRIX_PROPS, F1)
    sig = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    Pref = np.linalg.det(F1) * sig @ np.linalg.inv(F1).T
    print("homogeneous: iterations", info["iterations"], "norms", info["norms"])
    assert np.allclose(_pmat(P), Pref, rtol=0, atol=1e-11 * np.abs(Pref).max())
    w = r.displacement() - (X @ (F1 - np.eye(3)).T).reshape(-1)
    assert np.max(np.abs(w)) < 1e-11 and info["iterations"] <= 1
    r.close()


def test_voxel_mesh_matches_the_voxel_batch():
    n, rid = 6, 3
    X, conn, mid, ph = _two_phase(n, rid, 0.0)
    assert 0 < ph.sum() < n ** 3
    r = periodic.PeriodicRve(X, conn, mid, MAT)
    P, A, info = r.solve(F1, tangent=True)
    b = gpu.RveBatch(np.array([n]), ph, gpu.MODEL_NH, np.array([rve.MATRIX_PROPS, rve.INCLUSION_PROPS]))
    b.set_solver(newton_rtol=1e-12, newton_atol=1e-14, pcg_rtol=1e-12)
    Pb, Ab, ib = b.solve(F1[None], tangent=True)
    print("general mesh: iterations", info["iterations"], "PCG", info["pcg_iterations"], "batch info", ib[0])
    Pg = _pmat(P)
    err = np.max(np.abs(Pg - Pb[0])) / np.max(np.abs(Pb[0]))
    errA = np.max(np.abs(A.reshape(9, 9) - Ab[0])) / np.max(np.abs(Ab[0]))
    print("P rel diff %.3e, A rel diff %.3e" % (err, errA))
    assert err < 1e-9 and errA < 1e-6
    r.close()
    b.close()


def test_perturbed_two_phase_mesh_matches_the_oracle_over_two_steps():
    n, rid = 4, 5
    X, conn, mid, ph = _two_phase(n, rid, 0.15)
    r = periodic.PeriodicRve(X, conn, mid, MAT)
    o = po.OraclePeriodicRve(X, conn, mid, MAT)
    for F in (F1, F2):
        P, _, info = r.solve(F)
        Po, Uo, _ = o.solve(F)
        U = r.displacement()
        print("step: iterations", info["iterations"], "PCG", info["pcg_iterations"])
        assert np.max(np.abs(P - Po)) <= 1e-9 * np.max(np.abs(Po))
        assert np.max(np.abs(U - Uo)) <= 1e-9 * np.max(np.abs(Uo))
        # periodicity: U(slave) - U(master) = (F - I)(X(slave) - X(master))
        pm = r.map
        s = np.nonzero(pm.slave_of >= 0)[0]
        Um = U.reshape(-1, 3)
        jump = Um[s] - Um[pm.slave_of[s]] - (X[s] - X[pm.slave_of[s]]) @ (F - np.eye(3)).T
        assert np.max(np.abs(jump)) < 1e-13
        r.commit()
        o.commit()
    r.close()


def test_two_level_preconditioner_matches_block_jacobi():
    """The reduced PCG with the two-level preconditioner (rigid-body modes of aggregates of the
    reduced mesh) reaches the same equilibrium and condensed tangent as block Jacobi on a 10^3
    perturbed two-phase mesh, in fewer PCG iterations."""
    n, rid = 10, 7
    X, conn, mid, ph = _two_phase(n, rid, 0.15)
    out = {}
    for prec in (gpu.PREC_BLOCK_JACOBI, gpu.PREC_TWO_LEVEL):
        r = periodic.PeriodicRve(X, conn, mid, MAT, preconditioner=prec)
        P, A, info = r.solve(F1, tangent=True)
        print("preconditioner", prec, "iterations", info["iterations"], "PCG", info["pcg_iterations"],
              r.red.preconditioner_info(), flush=True)
        out[prec] = (P, A, sum(info["pcg_iterations"]))
        r.close()
    (Pb, Ab, itb), (Pt, At, itt) = out[gpu.PREC_BLOCK_JACOBI], out[gpu.PREC_TWO_LEVEL]
    assert np.max(np.abs(Pt - Pb)) <= 1e-9 * np.max(np.abs(Pb))
    assert np.max(np.abs(At - Ab)) <= 1e-7 * np.max(np.abs(Ab))
    assert itt < itb


def test_condensed_tangent_matches_finite_differences():
    n, rid = 4, 2
    X, conn, mid, ph = _two_phase(n, rid, 0.15)
    r = periodic.PeriodicRve(X, conn, mid, MAT, newton_rtol=1e-13, newton_atol=1e-15, pcg_rtol=1e-13)
    P, A, _ = r.solve(F1, tangent=True)
    A = A.reshape(9, 9)
    h = 1e-6
    for kl in range(9):
        k, L = kl % 3, kl // 3
        E = np.zeros((3, 3))
        E[k, L] = h
        Pp, _, _ = r.solve(F1 + E)
        Pm, _, _ = r.solve(F1 - E)
        fd = (Pp - Pm) / (2 * h)
        assert np.max(np.abs(fd - A[:, kl])) <= 1e-6 * np.max(np.abs(A)), (kl, fd, A[:, kl])
    r.close()


def test_multiscale_case_runs_from_its_files_and_restarts(tmp_path):
    """A periodic multiscale case written in the reference's formats runs end to end
    (analysis.run_multiscale): its Micro Structure probes (homogenized P, F, Cauchy stress from the
    replay) and a Gauss-point probe match the oracle's periodic solve of the same macro F history;
    an interrupted run restarted from its FEMAnalysis.result writes the same records bitwise."""
    import case_writer as cw
    from ceos_amd import analysis, caseio, probes
    X, conn, mid, ph = _two_phase(4, 5, 0.15)
    full, part = tmp_path / "full", tmp_path / "part"
    for d in (full, part):
        d.mkdir()
        cw.write_multiscale_case(str(d), X, conn, mid)
    case = caseio.Case(str(full))
    hist, tries = analysis.run_multiscale(case, str(full))
    assert len(hist) == 3 and all(ok for *_, ok in tries)
    Xc = case.mesh["coords"]  # the deck's E16.9 coordinates
    o = po.OraclePeriodicRve(Xc, conn, mid, MAT)
    rows = {n: probes.probe_file_rows(full / n)[1] for n in ("P.probe", "F.probe", "S.probe", "gp.probe")}
    assert [t for t, _ in rows["P.probe"]] == pytest.approx([0.0, 1 / 3, 2 / 3, 1.0], rel=1e-14)
    assert np.all(rows["P.probe"][0][1] == 0.0)
    for k, t in enumerate([1 / 3, 2 / 3, 1.0], start=1):
        F = case.macro_F(t)
        Po, Uo, _ = o.solve(F)
        sig_o = o.m.gp_stress()[4, 2].copy()
        o.commit()
        P = rows["P.probe"][k][1]
        assert np.max(np.abs(P - Po)) <= 1e-8 * np.max(np.abs(Po)), (t, P, Po)
        assert np.allclose(rows["F.probe"][k][1], [F[0, 0], F[0, 1], F[2, 2]], rtol=0, atol=1e-13)
        Pm = P.reshape(3, 3).T
        S = Pm @ F.T / np.linalg.det(F)
        assert np.allclose(rows["S.probe"][k][1], [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[1, 2], S[0, 2]],
                           rtol=0, atol=1e-8 * np.abs(S).max())
        assert np.max(np.abs(rows["gp.probe"][k][1] - sig_o)) <= 1e-8 * np.abs(sig_o).max()
    # interrupted after step 2, restarted: the same records
    case2 = caseio.Case(str(part))
    analysis.run_multiscale(case2, str(part), stop_after=(1, 2))
    hist_b, _ = analysis.run_multiscale(caseio.Case(str(part)), str(part), restart=True)
    assert [h[1] for h in hist_b] == [3]
    _, rf = caseio.read_results(full / "FEMAnalysis.result")
    _, rp = caseio.read_results(part / "FEMAnalysis.result")
    assert len(rf) == len(rp) == 4
    for a, b in zip(rf, rp):
        assert a["time"] == b["time"] and np.array_equal(a["U"], b["U"])
    assert (full / "P.probe").read_text() == (part / "P.probe").read_text()
