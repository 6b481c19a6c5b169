"""The package's reader of the reference's case files (ceos_amd.caseio) on the reference's own
Checker files (tests/golden/checker/, copied by tests/golden/make_checker_fixtures.py), against
the JSON fixtures the same script extracted independently; the FEMAnalysis.result writer and
reader; the probe post-processing."""
import json
import os

import numpy as np
import pytest

from ceos_amd import caseio

HERE = os.path.dirname(os.path.abspath(__file__))
CHECKER = os.path.join(HERE, "golden", "checker")
CASES = {
    "NonLinear_FEA_3D_Hexa8_StVenant": "svk_hex8_case.json",
    "NonLinear_FEA_3D_Hexa8_NeoHookean_Incompressible": "nhq1p0_hex8_case.json",
}


@pytest.mark.parametrize("name", list(CASES))
def test_case_reader_matches_fixture(name):
    case = caseio.Case(os.path.join(CHECKER, name))
    fx = json.load(open(os.path.join(HERE, "golden", CASES[name])))
    assert np.array_equal(case.mesh["coords"], np.array(fx["coords"]))
    assert np.array_equal(case.mesh["conn"], np.array(fx["conn"]))
    assert list(case.mesh["mat"]) == fx["mat"]
    assert case.mesh["fixed"] == fx["fixed"]
    assert case.mesh["nodal_disp"] == fx["nodal_disp"]
    assert case.mesh["nodal_load"] == fx["nodal_load"]
    assert {k: [list(r) for r in v] for k, v in case.tables.items()} == fx["tables"]
    assert case.load_cases == fx["time"]
    s = case.settings
    assert s["nl_analysis"] and s["newton_tol"] == fx["newton"]["tol"] and s["newton_maxiter"] == fx["newton"]["maxiter"]
    model, props = s["materials"][1]
    assert model == fx["material"]["model"]
    if model == 5:
        assert props == [fx["material"]["E"], fx["material"]["nu"]] and s["element_technology"] == 1
    else:
        assert props == [fx["material"]["C10"], fx["material"]["K"]] and s["element_technology"] == 2
    assert [(p.location, p.node, p.components) for p in case.probes] == [("node", 11, [1]), ("node", 11, [2]),
                                                                        ("node", 11, [3])]
    # boundary conditions in DOF form
    assert len(case.fixed_dof) == 45 and len(case.presc_dof) == 15
    assert np.all(case.presc_dof % 3 == 2)
    f = case.fext(0.5)
    assert f[3 * 58 + 1] == pytest.approx(-0.5e7) and np.count_nonzero(f) == 1


def test_results_file_round_trip(tmp_path):
    case = caseio.Case(os.path.join(CHECKER, "NonLinear_FEA_3D_Hexa8_StVenant"))
    rng = np.random.default_rng(1)
    path = tmp_path / "FEMAnalysis.result"
    Us = [rng.normal(size=case.ndof) for _ in range(3)]
    with open(path, "w") as f:
        caseio.write_results_header(f, case.ndof)
        caseio.write_results_record(f, Us[0], 0.0, 1, 1, 0, 0, 1, 0)
        caseio.write_results_record(f, Us[1], 0.01, 1, 1, 1, 0, 0, 4)
        caseio.write_results_record(f, Us[2], 0.02, 1, 1, 1, 1, 1, 5)
    ndof, recs = caseio.read_results(path)
    assert ndof == case.ndof and len(recs) == 3
    for r, U in zip(recs, Us):
        assert np.array_equal(r["U"], U)  # 17 significant digits: exact
    assert [r["flag_end_step"] for r in recs] == [1, 0, 1] and recs[2]["iterations"] == 5
    rows = caseio.post_process(case, path, tmp_path)  # node probes only: no replay mesh
    # only the end-of-step records reach the probes
    assert [t for t, _ in rows["CEOS_Disp_Z.probe"]] == [0.0, 0.02]
    assert rows["CEOS_Disp_Z.probe"][1][1][0] == Us[2][3 * 10 + 2]
    back = caseio.read_table(tmp_path / "CEOS_Disp_Z.probe")  # E23.15: 15 significant digits
    assert [t for t, _ in back] == [0.0, 0.02]
    assert np.allclose([v for _, v in back], [Us[0][32], Us[2][32]], rtol=5e-15, atol=0)
    its = open(tmp_path / "NumberOfIterationsToConverge.dat").read().split()
    assert its[-1] == "5"


def test_unsupported_inputs_are_refused(tmp_path):
    src = os.path.join(CHECKER, "NonLinear_FEA_3D_Hexa8_StVenant", "Settings.dat")
    text = open(src, encoding="latin-1").read()
    (tmp_path / "a.dat").write_text(text.replace("Hypothesis of Analysis = 3D", "Hypothesis of Analysis = Axisymmetric"),
                                    encoding="latin-1")
    with pytest.raises(caseio.CaseError):
        caseio.read_settings(tmp_path / "a.dat")
    (tmp_path / "b.dat").write_text(text.replace("St_Venant_Kirchhoff", "Hyperelastic_Tranverse_Isotropic"),
                                    encoding="latin-1")
    with pytest.raises(caseio.CaseError):
        caseio.read_settings(tmp_path / "b.dat")


def test_restart_start_follows_the_reference_rules():
    """Resume point of a restarted analysis (FEM/ModFEMAnalysis.f90:1476-1514)."""
    from ceos_amd import analysis
    rs = analysis.restart_start
    assert rs([50], 1, 25) == (1, 26)
    assert rs([50], 1, 50) is None  # last step of the last load case done
    assert rs([10, 20], 1, 4) == (1, 5)
    assert rs([10, 20], 1, 10) == (2, 1)  # load case finished: the next one from step 1
    assert rs([10, 20], 2, 7) == (2, 8)
    assert rs([50], 1, 1) == (1, 2)  # the time-zero record reads as load case 1, step 1
    for bad in (([10], 2, 1), ([10, 20], 1, 11), ([10], 1, 11)):
        with pytest.raises(caseio.CaseError):
            rs(*bad)


def test_restart_option_is_read(tmp_path):
    src = os.path.join(CHECKER, "NonLinear_FEA_3D_Hexa8_StVenant", "Settings.dat")
    text = open(src, encoding="latin-1").read()
    for value, expect in (("True", True), ("False", False)):
        p = tmp_path / ("Settings_%s.dat" % value)
        p.write_text(text.replace("Maximum Cut Backs = 10", "Maximum Cut Backs = 10\n\tRestart = %s" % value),
                     encoding="latin-1")
        assert caseio.read_settings(str(p))["restart"] is expect
    assert caseio.read_settings(src)["restart"] is False
    p = tmp_path / "Settings_bad.dat"
    p.write_text(text.replace("Maximum Cut Backs = 10", "Maximum Cut Backs = 10\n\tRestart = Maybe"), encoding="latin-1")
    with pytest.raises(caseio.CaseError):
        caseio.read_settings(str(p))
