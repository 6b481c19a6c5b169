"""FE^2 macro points on the GPU against the oracle.

  * Gauss-point F of the library equals the oracle's (same formula, 1e-14);
  * external material points filled with a Neo-Hookean point's pulled-back (S, dS/dE) reproduce
    the oracle's Neo-Hookean Fint and K (the push forward and the point indexing, 1e-11);
  * the RVE-result conversion kernel equals fe2_oracle.rve_to_ext (1e-12);
  * a subset solve equals the full-batch solve of the same RVEs (bitwise: same arithmetic);
  * the surrogate's S, dS/dE and hull flags equal the FP64 restatement (FP32 inference: 2e-5 on
    S, 2e-4 on dS/dE);
  * a plain macro evaluation (1 element, 8 RVEs) gives the Fint the oracle RVE solves give through
    Fint_a = sum_g detJ_X P(g) grad_X N_a(g) (to the solver tolerances, 1e-8).
"""
import numpy as np
import pytest

import fe2_oracle as fo
import mlp_oracle as mo
import oracle_lib as ol
import rve_oracle as ro
from ceos_amd import fe2, gpu, hiprt, meshgen, rve

pytestmark = pytest.mark.gpu

NH = [0.5, 50.0]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _mesh_state(seed=0):
    X, conn = meshgen.cube(3, perturb=0.1)
    U = np.random.default_rng(seed).normal(scale=0.02, size=3 * len(X))
    o = ol.OracleMesh(X, conn)
    o.set_material(1, ol.MODEL_NH, NH)
    o.update_mesh(U)
    assert o.solve_constitutive(0.0, U) == 0
    return X, conn, U, o


def test_gp_deformation_gradient_matches_oracle():
    X, conn, U, o = _mesh_state(1)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, gpu.MODEL_NH, NH)
    Fg = g.gp_deformation_gradient(U)
    Fo = o.gp_F()
    assert np.max(np.abs(Fg - Fo)) < 1e-14  # same formula; the oracle's C loop rounds its sums in its own order


def test_external_points_reproduce_neo_hookean():
    X, conn, U, o = _mesh_state(2)
    Fo = o.gp_F()
    ext = np.zeros((len(conn), 8, 27))
    for e in range(len(conn)):
        for g in range(8):
            s, D, _ = ol.material_point(ol.MODEL_NH, NH, Fo[e, g])
            S, C = fo.pull_back(s, D, Fo[e, g])
            ext[e, g, :6] = S
            ext[e, g, 6:] = fo.up21(C)
    fo_int, _ = o.internal_force()
    Ko = o.tangent()
    dext = hiprt.DeviceArray.from_numpy(ext.reshape(-1))
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, gpu.MODEL_EXT, [])
    g.set_external(dext.ptr)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    f, st = g.internal_force()
    K = g.tangent()
    print("ext vs NH: Fint %.2e K %.2e" % (rel(f, fo_int), rel(K, Ko)), flush=True)
    assert st == 0 and rel(f, fo_int) < 1e-11 and rel(K, Ko) < 1e-11


def test_external_model_needs_its_points():
    X, conn = meshgen.cube(2)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, gpu.MODEL_EXT, [])
    U = np.zeros(3 * len(X))
    assert g.solve_constitutive(0.0, U) == gpu.ERR_ARG


def test_rve_to_external_matches_oracle():
    rng = np.random.default_rng(3)
    n = 50
    F = np.eye(3)[None] + rng.uniform(-0.2, 0.2, size=(n, 3, 3))
    P = rng.normal(size=(n, 3, 3))
    A = rng.normal(size=(n, 9, 9))
    A = A + A.transpose(0, 2, 1)
    D = hiprt.DeviceArray.from_numpy
    dF = D(np.ascontiguousarray(F.transpose(0, 2, 1).reshape(-1)))
    dP = D(np.ascontiguousarray(P.transpose(0, 2, 1).reshape(-1)))
    dA = D(np.ascontiguousarray(A.reshape(-1)))
    sel = np.arange(n, dtype=np.int32)[::-1].copy()
    dsel = D(sel)
    dext = hiprt.DeviceArray(27 * n)
    gpu.rve_to_external_device(n, dF.ptr, dP.ptr, dA.ptr, dext.ptr, dsel=dsel.ptr)
    ext = dext.numpy().reshape(n, 27)
    for k in range(n):
        S, C = fo.rve_to_ext(F[k], P[k], A[k])
        r = sel[k]
        assert np.allclose(ext[r, :6], S, rtol=1e-12, atol=1e-12 * np.abs(S).max())
        assert np.allclose(fo.from_up21(ext[r, 6:]), C, rtol=1e-12, atol=1e-12 * np.abs(C).max())


def test_subset_solve_equals_full_solve():
    sides = np.array([4, 6, 4, 6, 4], dtype=np.int32)
    ph = np.concatenate([rve.phases(int(n), i) for i, n in enumerate(sides)])
    props = np.array([rve.MATRIX_PROPS, rve.INCLUSION_PROPS])
    rng = np.random.default_rng(4)
    Fs = np.eye(3)[None] + rng.uniform(-0.05, 0.05, size=(5, 3, 3))
    a = gpu.RveBatch(sides, ph, gpu.MODEL_NH, props)
    Pa, Aa, _ = a.solve(Fs, tangent=True)
    b = gpu.RveBatch(sides, ph, gpu.MODEL_NH, props)
    sel = np.array([3, 0], dtype=np.int32)
    Pb, Ab, info = b.solve_subset(sel, Fs[sel], tangent=True)
    assert np.all(info[:, 0] == 0)
    assert np.array_equal(Pb, Pa[sel]) and np.array_equal(Ab, Aa[sel])
    # the others kept their undeformed state
    assert np.max(np.abs(b.fluctuation(1))) == 0.0
    with pytest.raises(gpu.CeosError):
        b.solve_subset(np.array([1, 1], dtype=np.int32), Fs[[1, 1]])


def _random_surrogate(seed, width=64, n=300):
    rng = np.random.default_rng(seed)
    m = gpu.Mlp(27, width, 6, max_batch=n, seed=seed)
    p = m.params().astype(np.float64)
    cfg = {"u_scale": 0.25, "lam_lo": 0.8, "lam_hi": 1.25, "e_ref": 0.1,
           "d_mean": rng.normal(size=21), "d_std": rng.uniform(0.5, 2, 21),
           "d_lo": -np.full(21, 2.0), "d_hi": np.full(21, 2.0),
           "t_mean": rng.normal(size=6), "t_std": rng.uniform(0.5, 2, 6)}
    return m, p, cfg


def test_surrogate_matches_fp64_restatement():
    n, width = 300, 64
    m, p, cfg = _random_surrogate(5, width, n)
    sur = gpu.Surrogate(m, cfg, n)
    rng = np.random.default_rng(6)
    U = fe2.random_stretch(rng, n, 0.75, 1.3)  # some points outside [0.8, 1.25]
    R = fe2.haar_rotations(rng, n)
    F = R @ U
    Clin = cfg["d_mean"] + cfg["d_std"] * rng.uniform(-2.2, 2.2, size=(n, 21)) * (rng.uniform(size=(n, 1)) < 0.1) \
        + cfg["d_std"] * rng.uniform(-1.5, 1.5, size=(n, 21)) * (rng.uniform(size=(n, 1)) >= 0.1)
    D = hiprt.DeviceArray.from_numpy
    dF = D(np.ascontiguousarray(F.transpose(0, 2, 1).reshape(-1)))
    dC = D(np.ascontiguousarray(Clin.reshape(-1)))
    dext = hiprt.DeviceArray(27 * n)
    dflags = hiprt.DeviceArray(n, np.int32)
    dP = hiprt.DeviceArray(9 * n)
    sur.eval_device(n, dF.ptr, dC.ptr, dext.ptr, dflags.ptr, dP.ptr)
    ext = dext.numpy().reshape(n, 27)
    flags = dflags.numpy()
    P = dP.numpy().reshape(n, 3, 3).transpose(0, 2, 1)
    worst_s = worst_c = 0.0
    for k in range(n):
        S, C, inside = fo.surrogate(p, cfg, F[k], Clin[k], width)
        assert flags[k] == (0 if inside else 1), k
        worst_s = max(worst_s, np.abs(ext[k, :6] - S).max() / np.abs(S).max())
        worst_c = max(worst_c, np.abs(fo.from_up21(ext[k, 6:]) - C).max() / np.abs(C).max())
        assert np.allclose(P[k], F[k] @ fo.from_voigt(ext[k, :6]), rtol=1e-12, atol=1e-12 * np.abs(P[k]).max())
    print("surrogate vs FP64: S %.2e  dS/dE %.2e  outside %d" % (worst_s, worst_c, flags.sum()), flush=True)
    assert 0 < flags.sum() < n
    assert worst_s < 2e-5 and worst_c < 2e-4


def _hex8_grad_X(Xe):
    """grad_X N at the 8 Gauss points (reference order) and detJ_X for one element."""
    g = 1.0 / np.sqrt(3.0)
    xi = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
    out = []
    for gp in xi * g:
        dN = 0.125 * np.stack([xi[:, 0] * (1 + xi[:, 1] * gp[1]) * (1 + xi[:, 2] * gp[2]),
                               xi[:, 1] * (1 + xi[:, 0] * gp[0]) * (1 + xi[:, 2] * gp[2]),
                               xi[:, 2] * (1 + xi[:, 0] * gp[0]) * (1 + xi[:, 1] * gp[1])], axis=1)
        J = dN.T @ Xe
        out.append((np.linalg.solve(J, dN.T).T, np.linalg.det(J)))
    return out


def test_plain_macro_evaluation_matches_oracle_rves():
    m = fe2.MacroFE2(n_macro=1, side_choices=(4,))
    rng = np.random.default_rng(7)
    U = rng.uniform(-0.04, 0.04, size=3 * m.nn)
    st = m.evaluate_rve(U)
    print(st, flush=True)
    Fg = m.gp_F()
    Xe = m.X[m.conn[0] - 1]
    fint = np.zeros(3 * m.nn)
    for g, (gX, detX) in enumerate(_hex8_grad_X(Xe)):
        r = g  # element 0, Gauss point g
        o = ro.OracleRVE(int(m.sides[r]), m.phases[64 * r:64 * (r + 1)])
        Po, _, _, _ = o.solve(Fg[r])
        Po = Po.reshape(3, 3).T
        fe = gX @ Po.T * detX  # fe[a, i] = sum_J P_iJ gradX_a,J
        for a in range(8):
            nid = m.conn[0, a] - 1
            fint[3 * nid:3 * nid + 3] += fe[a]
    f = m.fint()
    print("macro Fint vs oracle RVEs %.2e" % rel(f, fint), flush=True)
    assert rel(f, fint) < 1e-8
    K = m.tangent()
    assert np.all(np.isfinite(K)) and K[0] > 0


def test_macro_evaluation_graph_replays_match_eager():
    """Both macro evaluation paths replayed from captured HIP graphs (no memset nodes) give the
    eager results bitwise, replay after replay, with eager evaluations and synchronisations in
    between (every run starts from undeformed RVEs)."""
    mlp = gpu.Mlp(27, 64, 6, max_batch=8, seed=3)
    cfg = {"u_scale": 0.25, "lam_lo": 0.5, "lam_hi": 2.0, "e_ref": 0.1,
           "d_mean": np.full(21, 1.0), "d_std": np.full(21, 10.0),
           "d_lo": np.full(21, -1e30), "d_hi": np.full(21, 1e30),
           "t_mean": np.zeros(6), "t_std": np.full(6, 0.1)}
    sur = gpu.Surrogate(mlp, cfg, 8)
    m = fe2.MacroFE2(n_macro=1, side_choices=(4,), surrogate=(mlp, sur, {}))
    m.setup_descriptors()
    rng = np.random.default_rng(12)
    U1 = rng.uniform(-0.03, 0.03, size=3 * m.nn)
    U2 = rng.uniform(-0.03, 0.03, size=3 * m.nn)

    def run(mode, U, g):
        m.rves.reset()
        if mode == "rve":
            m.evaluate_rve(U, use_graph=g)
        else:
            st = m.evaluate_surrogate(U, judge_frac=0.0, use_graph=g)
            assert st["outside_hull"] == 0
        return m.ext().copy(), m.fint().copy(), m.tangent().copy()

    for mode in ("rve", "surrogate"):
        ref = {0: run(mode, U1, False), 1: run(mode, U2, False)}
        for k in (0, 1, 0, 1):
            r = run(mode, (U1, U2)[k], True)
            assert all(np.array_equal(a, b) for a, b in zip(r, ref[k])), (mode, k)
            run(mode, (U2, U1)[k], False)  # eager work between the replays
    assert set(m.graphs) == {"rve", "surrogate", "assemble"}
