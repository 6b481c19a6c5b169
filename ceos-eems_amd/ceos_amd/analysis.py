"""Quasi-static analysis of a case on the GPU: the reference's driver loop around the library.

  load cases, steps, golden-ratio cut-back    FEM/ModFEMAnalysis.f90:1136-1374
  external forces and prescribed displacement FEM/BoundaryConditions/ModBoundaryConditions.f90:540-640
  Newton, relative test on the max-abs norm   Tools/MathTools/NonlinearSolver/ModNewtonRaphsonFull.f90:44-168
  FEMAnalysis.result records                  FEM/ModFEMAnalysis.f90:1094-1129, 1211-1212, 1314-1336
  restart from FEMAnalysis.result             FEM/ModFEMAnalysis.f90:2827-2956 (replay), 1476-1516 (resume)

Every evaluation, the boundary conditions, the residual norm and the linear solves run in
libceos_gpu.so (ceos_solve_constitutive, ceos_residual_bc, ceos_linear_solve: PCG, and MINRES
for indefinite tangents); the host keeps only the loop, U and the tables.

    python -m ceos_amd.analysis <case directory> [output directory]
"""
import math
import os
import sys
import time as _time

import numpy as np

from . import caseio, gpu

GOLDEN_RATIO = (1.0 + math.sqrt(5.0)) / 2.0


class NewtonFailed(Exception):
    pass


def make_mesh(case):
    s = case.settings
    m = case.mesh
    mesh = gpu.CeosMesh(m["coords"], m["conn"], mat_id=m["mat"], nl_analysis=s["nl_analysis"],
                        elem_tech=s["element_technology"])
    for mid, (model, props) in s["materials"].items():
        mesh.set_material(mid, model, props)
    return mesh


def linear_solve(mesh, b, rtol, minres_rtol):
    """PCG at rtol; after a breakdown (indefinite tangent) MINRES at minres_rtol, whose residual
    is measured in the M^-1 norm and levels off earlier than PCG's 2-norm on stiff systems."""
    x, its, rr, flag = mesh.linear_solve(b, rtol, 20000)
    if flag == 2:
        mesh.set_linear_solver(gpu.SOLVER_MINRES)
        try:
            x, its2, rr, flag = mesh.linear_solve(b, minres_rtol, 20000)
        finally:
            mesh.set_linear_solver(gpu.SOLVER_PCG)
    if flag not in (1, 4):
        raise NewtonFailed("linear solve flag %d" % flag)
    return x


def newton(mesh, time, fext, Ubar, Xguess, nl, tol, maxiter, norms, absolute, linear_rtol, minres_rtol):
    """NewtonRaphsonFull_Solve on the device system.  Returns (X, iterations)."""
    X = Xguess.copy()
    it = 0
    norm0 = None
    while True:
        st = mesh.solve_constitutive(time, X)
        if st:
            raise NewtonFailed("evaluation status %d" % st)
        R, normR = mesh.residual_bc(it, fext, Ubar, 2)
        norms.append(normR)
        if it == 0:
            norm0 = normR
        if it > 0 and normR < (tol if absolute else norm0 * tol):
            return X, it
        if it >= maxiter:
            raise NewtonFailed("maximum number of iterations")
        it += 1
        X = X + linear_solve(mesh, -R, linear_rtol, minres_rtol)
        if nl:
            mesh.update_mesh(X)


def replay_results(mesh, path, nl):
    """RestartAnalysis (FEM/ModFEMAnalysis.f90:2827-2956): every end-of-step record of
    FEMAnalysis.result, in file order, sets U, moves the mesh (finite strain), updates the
    constitutive model at the record's time and commits the state.  Returns (U, last load case,
    last step).  The time-zero record is written with load case 1, step 1 and the end-of-step
    flag, so a file holding no finished step gives (1, 1), as in the reference."""
    ndof, recs = caseio.read_results(path)
    if ndof != 3 * mesh.nn:
        raise caseio.CaseError("FEMAnalysis.result holds %d DOFs, the mesh %d" % (ndof, 3 * mesh.nn))
    U, last = np.zeros(ndof), None
    for r in recs:
        if r["flag_end_step"] != 1:
            continue
        U = r["U"].copy()
        last = (r["load_case"], r["step"])
        if nl:
            mesh.update_mesh(U)
        st = mesh.solve_constitutive(r["time"], U)
        if st:
            raise NewtonFailed("restart: evaluation status %d at load case %d step %d" % (st, *last))
        mesh.switch_converged_state()
    if last is None:
        raise caseio.CaseError("FEMAnalysis.result holds no end-of-step record")
    return U, last[0], last[1]


def restart_start(steps, last_lc, last_st):
    """First load case of a restarted analysis and the step to resume within it
    (FEM/ModFEMAnalysis.f90:1476-1514); steps[lc - 1] = number of steps of load case lc.
    Returns (init_lc, init_st), or None when the last step of the last load case is done."""
    nlc = len(steps)
    if nlc > last_lc:
        if steps[last_lc - 1] > last_st:
            init_lc = last_lc
        elif steps[last_lc - 1] == last_st:
            init_lc, last_st = last_lc + 1, 0
        else:
            raise caseio.CaseError("Error restarting load case.")
    elif nlc == last_lc:
        init_lc = last_lc
    else:
        raise caseio.CaseError("Error restarting load case.")
    if steps[init_lc - 1] > last_st:
        return init_lc, last_st + 1
    if steps[init_lc - 1] == last_st:
        return None  # "Last load step already calculated!" (the only load case left)
    raise caseio.CaseError("Error restarting load step.")


def run(case, out_dir=None, linear_rtol=1e-12, minres_rtol=1e-12, absolute=False, verbose=False,
        restart=None, stop_after=None, attempts=None):
    """The quasi-static analysis of `case`.  Writes FEMAnalysis.result, the probe files and
    NumberOfIterationsToConverge.dat into out_dir when given.  Returns (history, norms): history
    rows (load case, step, time, U, iterations) per converged step; norms, the residual norms of
    every Newton solve (one list per substep attempt).  absolute=True stops Newton at
    norm < tol instead of norm < tol * first norm: the test of the binary that wrote the
    reference's Checker logs.

    restart (default: the case's Restart option) resumes from out_dir/FEMAnalysis.result: the
    end-of-step records are replayed through the constitutive update and the state commit, and
    the loop continues after the last finished step, appending to the file.  The reference
    accepts the option for its periodic multiscale analysis only (ModReadInputFile.f90:287-301);
    the replay and the resume rules are the same for this quasi-static loop.  stop_after =
    (load case, step) ends the run after that step, as an interrupted analysis would.
    attempts, a list, receives (load case, step, alpha, converged) for every Newton solve."""
    s = case.settings
    mesh = make_mesh(case)
    mesh.set_bc(case.presc_dof, case.fixed_dof)
    mesh.set_linear_solver(gpu.SOLVER_PCG)
    nl = s["nl_analysis"]
    n = case.ndof
    U = np.zeros(n)
    history, norms = [], []
    res = None
    restart = s.get("restart", False) if restart is None else restart
    steps = [c["steps"] for c in case.load_cases]
    start = (1, 1)
    if restart:
        path = os.path.join(out_dir or "", "FEMAnalysis.result")
        if not out_dir or not os.path.exists(path):
            raise caseio.CaseError("no FEMAnalysis.result found")
        U, last_lc, last_st = replay_results(mesh, path, nl)
        start = restart_start(steps, last_lc, last_st)
        res = open(path, "a")
    elif out_dir:
        os.makedirs(out_dir, exist_ok=True)
        res = open(os.path.join(out_dir, "FEMAnalysis.result"), "w")
        caseio.write_results_header(res, n)
        caseio.write_results_record(res, U, 0.0, 1, 1, 0, 0, 1, 0)
    try:
        for lc, c in enumerate(case.load_cases, start=1):
            if start is None or lc < start[0]:
                continue
            dt = (c["t1"] - c["t0"]) / c["steps"]
            for stp in range(start[1] if lc == start[0] else 1, c["steps"] + 1):
                ta, tb = c["t0"] + (stp - 1) * dt, c["t0"] + stp * dt
                fext0, dfext = case.fext(ta), case.fext(tb) - case.fext(ta)
                dpresc = np.zeros(n)
                if len(case.presc_dof):
                    v0, v1 = case.presc_values(ta), case.presc_values(tb)
                    U[case.presc_dof] = v0
                    dpresc[case.presc_dof] = v1 - v0
                Ubar0, Uconv = U.copy(), U.copy()
                amax, amin = 1.0, 0.0
                alpha = amax
                cutback = substep = 0
                while True:
                    t = ta + alpha * dt
                    nrm = []
                    try:
                        U, its = newton(mesh, t, fext0 + alpha * dfext, Ubar0 + alpha * dpresc, Uconv, nl,
                                        s["newton_tol"], s["newton_maxiter"], nrm, absolute, linear_rtol,
                                        minres_rtol)
                        ok = True
                    except NewtonFailed:
                        ok = False
                    norms.append(nrm)
                    if attempts is not None:
                        attempts.append((lc, stp, alpha, ok))
                    if verbose:
                        print("LC %d step %d alpha %.4f: %s" % (lc, stp, alpha, " ".join("%.3e" % v for v in nrm)),
                              flush=True)
                    if not ok:
                        alpha = amin + (1.0 - 1.0 / GOLDEN_RATIO) * (alpha - amin)
                        U = Uconv.copy()
                        if nl:
                            mesh.update_mesh(U)
                        cutback += 1
                        substep = 1
                        if cutback > s["max_cut_backs"]:
                            raise NewtonFailed("load case %d step %d did not converge with %d cut backs"
                                               % (lc, stp, s["max_cut_backs"]))
                    elif alpha == 1.0:
                        substep += 1
                        history.append((lc, stp, t, U.copy(), its))
                        if res:
                            caseio.write_results_record(res, U, t, lc, stp, cutback, substep, 1, its)
                        break
                    else:
                        substep += 1
                        if res:
                            caseio.write_results_record(res, U, t, lc, stp, cutback, substep, 0, its)
                        aux, amin = amin, alpha
                        alpha = min(alpha + GOLDEN_RATIO * (alpha - aux), 1.0)
                        Uconv = U.copy()
                mesh.switch_converged_state()
                if stop_after is not None and (lc, stp) == tuple(stop_after):
                    start = None
                    break
            if start is None:
                break
    finally:
        if res:
            res.close()
    if out_dir:
        caseio.post_process(case, os.path.join(out_dir, "FEMAnalysis.result"), out_dir)
    mesh.close()
    return history, norms


def run_multiscale(case, out_dir=None, verbose=False, restart=None, stop_after=None, attempts=None,
                   pcg_rtol=1e-12):
    """The periodic multiscale analysis of one RVE (FEM/ModFEMAnalysis.f90:1440-1640): for every
    load case and step the macroscopic deformation gradient of the MACROSCOPIC DEFORMATION
    GRADIENT tables at the step's end, the RVE (the case's mesh, periodic on its box, the first
    boundary node set as BuildT's candidates) solved by ceos_amd.periodic.PeriodicRve with the
    case's Newton tolerance, the golden-ratio cut-back on failure (F = F_start + alpha (F_end -
    F_start), as GetBoundaryConditionsMultiscalePeriodic forms it), one FEMAnalysis.result record
    per converged substep and the commit of the Gauss-point states at the end of the step; then
    the post-processing, whose Micro Structure probes read the homogenized P, F and Cauchy stress.
    restart (default: the case's Restart option, which the reference accepts for this model)
    replays the end-of-step records and resumes after the last finished step.  Returns
    (history, attempts): history rows (load case, step, time, U, iterations, P (9, column-major))."""
    from .periodic import PeriodicRve
    s = case.settings
    if not s.get("multiscale"):
        raise caseio.CaseError("not a multiscale analysis")
    m = case.mesh
    nl = s["nl_analysis"]
    rv = PeriodicRve(m["coords"], m["conn"], m["mat"], s["materials"], nl=nl, boundary=case.boundary_nodes(),
                     newton_rtol=s["newton_tol"], newton_atol=1e-300, newton_maxit=s["newton_maxiter"],
                     pcg_rtol=pcg_rtol)
    n = case.ndof
    history, tries = [], []
    restart = s.get("restart", False) if restart is None else restart
    steps = [c["steps"] for c in case.load_cases]
    start = (1, 1)
    res = None
    if restart:
        path = os.path.join(out_dir or "", "FEMAnalysis.result")
        if not out_dir or not os.path.exists(path):
            raise caseio.CaseError("no FEMAnalysis.result found")
        U, last_lc, last_st = replay_results(rv.full, path, nl)
        c = case.load_cases[last_lc - 1]
        rv.set_start(U, case.macro_F(c["t0"] + last_st * (c["t1"] - c["t0"]) / c["steps"]))
        start = restart_start(steps, last_lc, last_st)
        res = open(path, "a")
    elif out_dir:
        os.makedirs(out_dir, exist_ok=True)
        res = open(os.path.join(out_dir, "FEMAnalysis.result"), "w")
        caseio.write_results_header(res, n)
        caseio.write_results_record(res, np.zeros(n), 0.0, 1, 1, 0, 0, 1, 0)
    try:
        for lc, c in enumerate(case.load_cases, start=1):
            if start is None or lc < start[0]:
                continue
            dt = (c["t1"] - c["t0"]) / c["steps"]
            for stp in range(start[1] if lc == start[0] else 1, c["steps"] + 1):
                ta, tb = c["t0"] + (stp - 1) * dt, c["t0"] + stp * dt
                F0, F1 = case.macro_F(ta), case.macro_F(tb)
                amin, alpha = 0.0, 1.0
                cutback = substep = 0
                while True:
                    t = ta + alpha * dt
                    try:
                        P, _, info = rv.solve(F0 + alpha * (F1 - F0), time=t)
                        ok = True
                    except (gpu.CeosError, RuntimeError):
                        ok = False
                    tries.append((lc, stp, alpha, ok))
                    if verbose:
                        print("LC %d step %d alpha %.4f: %s" % (lc, stp, alpha, "converged" if ok else "cut back"),
                              flush=True)
                    if not ok:
                        alpha = amin + (1.0 - 1.0 / GOLDEN_RATIO) * (alpha - amin)
                        cutback += 1
                        substep = 1
                        if cutback > s["max_cut_backs"]:
                            raise NewtonFailed("load case %d step %d did not converge with %d cut backs"
                                               % (lc, stp, s["max_cut_backs"]))
                        continue
                    substep += 1
                    U = rv.displacement()
                    end = abs(alpha - 1.0) < 1e-6
                    if res:
                        caseio.write_results_record(res, U, t, lc, stp, cutback, substep, 1 if end else 0,
                                                    info["iterations"])
                    if end:
                        history.append((lc, stp, t, U, info["iterations"], P))
                        break
                    rv.accept()
                    aux, amin = amin, alpha
                    alpha = min(alpha + GOLDEN_RATIO * (alpha - aux), 1.0)
                rv.commit()
                if stop_after is not None and (lc, stp) == tuple(stop_after):
                    start = None
                    break
            if start is None:
                break
    finally:
        if res:
            res.close()
        rv.close()
    if attempts is not None:
        attempts.extend(tries)
    if out_dir:
        caseio.post_process(case, os.path.join(out_dir, "FEMAnalysis.result"), out_dir)
    return history, tries


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    case = caseio.Case(argv[1])
    out = argv[2] if len(argv) > 2 else argv[1]
    t0 = _time.perf_counter()
    if case.settings.get("multiscale"):
        hist, _ = run_multiscale(case, out, verbose=True)
    else:
        hist, _ = run(case, out, verbose=True)
    print("%d steps in %.2f s; results in %s" % (len(hist), _time.perf_counter() - t0, out))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
