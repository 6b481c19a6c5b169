"""The reference's hex8 Checker cases run from their own input files (tests/golden/checker/) by
the package's reader and quasi-static runner (ceos_amd.caseio, ceos_amd.analysis): every
evaluation, boundary condition, residual norm and linear solve in the library.  The probe files
written by the runner's post-processing are compared with the Marc results at the tolerances of
the reference's Run_Checker.bat, and the residual norms with the reference's Log_CEOS.txt."""
import json
import os

import numpy as np
import pytest

from ceos_amd import analysis, caseio

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CHECKER = os.path.join(HERE, "golden", "checker")
# case, JSON fixture with the log, Checker tolerance of the probes (Run_Checker.bat /Tol)
CASES = [
    ("NonLinear_FEA_3D_Hexa8_StVenant", "svk_hex8_case.json", 1e-5),
    ("NonLinear_FEA_3D_Hexa8_NeoHookean_Incompressible", "nhq1p0_hex8_case.json", 1e-3),
]


@pytest.mark.parametrize("name, fixture, tol", CASES)
def test_checker_case_from_its_own_files(name, fixture, tol, tmp_path):
    d = os.path.join(CHECKER, name)
    case = caseio.Case(d)
    hist, norms = analysis.run(case, out_dir=str(tmp_path), linear_rtol=1e-14, minres_rtol=1e-12, absolute=True)
    assert len(hist) == 50
    for c in "XYZ":
        mine = caseio.read_table(tmp_path / ("CEOS_Disp_%s.probe" % c))
        marc = caseio.read_table(os.path.join(d, "marc_disp_%s.dat" % c.lower()))
        assert len(mine) == len(marc) == 51
        assert np.allclose([t for t, _ in mine], [t for t, _ in marc], rtol=0, atol=1e-12)
        err = max(abs(a - b) for (_, a), (_, b) in zip(mine, marc))
        print("%s %s: max |probe - Marc| %.3e (Checker tolerance %.0e)" % (name, c, err, tol))
        assert err < tol
    log = json.load(open(os.path.join(HERE, "golden", fixture)))["log_norms"]
    for step in range(3):
        assert len(norms[step]) == len(log[step])
        for k, ref in enumerate(log[step]):
            assert abs(norms[step][k] - ref) <= 2e-8 * ref + 1e-12 * log[step][0], (step, k, norms[step][k], ref)
    # the result file holds the initial record and one per converged step
    ndof, recs = caseio.read_results(tmp_path / "FEMAnalysis.result")
    assert ndof == case.ndof and len(recs) == 51
    assert np.array_equal(recs[-1]["U"], hist[-1][3])
