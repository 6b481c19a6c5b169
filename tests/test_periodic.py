"""The periodic pairing of a general RVE mesh (ceos_amd.periodic.periodic_map, the reference's
BuildT) against a brute-force search of periodic images, its errors, and the CPU restatement of
the periodic RVE analysis (tests/periodic_oracle.py) on a homogeneous RVE, whose answer the
mathematics fixes: zero fluctuation and P = P(F) of the material."""
import numpy as np
import pytest

import oracle_lib as ol
import periodic_oracle as po
from ceos_amd import meshgen, periodic

NH = (0.5, 50.0)


def test_pairing_matches_brute_force_images():
    for n, seed in [(3, 1), (4, 2), (5, 3)]:
        X, conn = meshgen.cube(n, perturb=0.15, seed=seed)
        X = po.periodic_perturbation(X, 1.0 / n)
        pm = periodic.periodic_map(X)
        rep, lo, hi = po.image_classes(X)
        # same classes: two nodes share a reduced node exactly when they are periodic images
        nm = pm.node_map
        for A in range(len(X)):
            same = np.nonzero(nm == nm[A])[0]
            assert set(same) == set(np.nonzero(rep == rep[A])[0])
        assert len(pm.masters) == n ** 3 and len(np.unique(nm)) == n ** 3
        # masters lie on no max face, and the reduced nodes keep the order of their masters
        assert np.all(np.diff(pm.masters) > 0)
        assert np.all(np.abs(X[pm.masters] - hi).min(axis=1) > 1e-6)
        # the corners in the reference's order, all mapped to the first
        sig = np.sign(X[pm.corners] - (lo + hi) / 2)
        assert sig.tolist() == [[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1], [1, 1, 1], [-1, 1, 1],
                                [1, -1, 1], [1, 1, -1]]
        assert len(set(nm[pm.corners])) == 1
        # a slave differs from its master by a lattice vector
        s = np.nonzero(pm.slave_of >= 0)[0]
        d = (X[s] - X[pm.slave_of[s]]) / (hi - lo)
        assert np.allclose(d, np.round(d), atol=1e-12) and np.all(np.round(d) >= 0)


def test_non_periodic_meshes_are_refused():
    X, conn = meshgen.cube(3)
    Y = X.copy()
    Y[np.argmin(np.abs(Y[:, 0] - 1.0) + np.abs(Y[:, 1] - 1 / 3) + np.abs(Y[:, 2] - 1 / 3)), 1] += 0.01
    with pytest.raises(periodic.PeriodicMeshError):
        periodic.periodic_map(Y)
    Z = X.copy()
    Z[0] += 0.01  # corner off the box
    with pytest.raises(periodic.PeriodicMeshError):
        periodic.periodic_map(Z)


def test_oracle_restatement_on_a_homogeneous_rve():
    X, conn = meshgen.cube(3, perturb=0.2, seed=7)
    X = po.periodic_perturbation(X, 1.0 / 3)
    o = po.OraclePeriodicRve(X, conn, None, {1: (ol.MODEL_NH, NH)})
    F = np.array([[1.08, 0.05, -0.02], [0.01, 0.95, 0.04], [0.03, -0.02, 1.02]])
    P, U, its = o.solve(F)
    s, _, _ = ol.material_point(ol.MODEL_NH, NH, F)
    sig = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    Pref = np.linalg.det(F) * sig @ np.linalg.inv(F).T
    assert np.allclose(P, Pref.T.reshape(9), rtol=0, atol=1e-12 * np.abs(Pref).max())
    w = U - (X @ (F - np.eye(3)).T).reshape(-1)
    assert np.max(np.abs(w)) < 1e-12 and its <= 1


def test_multiscale_case_files_are_read(tmp_path):
    """A periodic multiscale case in the reference's files: HyperMesh mesh with its boundary node
    set, the MACROSCOPIC DEFORMATION GRADIENT block (tables, 'One', 'Zero'), the Micro Structure
    and Gauss-point probes."""
    import case_writer as cw
    from ceos_amd import caseio, rve
    X, conn = meshgen.cube(4, perturb=0.15, seed=11)
    X = po.periodic_perturbation(X, 0.25)
    mat = rve.phases(4, 5).astype(np.int32) + 1
    cw.write_multiscale_case(str(tmp_path), X, conn, mat)
    case = caseio.Case(str(tmp_path))
    s = case.settings
    assert s["multiscale"] and s["preprocessor"] == "hypermesh" and s["newton_tol"] == 1e-10
    assert s["materials"] == {1: (3, [0.5, 50.0]), 2: (3, [5.0, 500.0])}
    assert np.allclose(case.mesh["coords"], X, rtol=0, atol=5e-9)  # E16.9 in the deck
    assert np.array_equal(case.mesh["conn"], conn) and np.array_equal(case.mesh["mat"], mat)
    b = case.boundary_nodes()
    lo, hi = X.min(axis=0), X.max(axis=0)
    on = np.any((np.abs(X - lo) < 1e-9) | (np.abs(X - hi) < 1e-9), axis=1)
    assert np.array_equal(np.sort(b), np.nonzero(on)[0])
    F = case.macro_F(0.5)
    assert np.allclose(F, [[1.05, 0.025, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.975]], rtol=0, atol=1e-15)
    assert [p.location for p in case.probes] == ["microstructure"] * 3 + ["gausspoint"]
    assert case.ndof == 3 * len(X) and len(case.fixed_dof) == 0 and len(case.presc_dof) == 0
    # the pairing from the boundary set alone equals the one from every node
    pm_b = periodic.periodic_map(case.mesh["coords"], boundary=b)
    pm_a = periodic.periodic_map(case.mesh["coords"])
    assert np.array_equal(pm_b.node_map, pm_a.node_map)
