"""Parity of the HIP element path (libceos_gpu.so through its C ABI) with the CPU oracle.

Bar: the CSR pattern bit-exact; Fint, K and Gauss-point stresses in FP64 within 1e-12
relative (norm-wise over the assembled arrays, summation order differs: colour order on the
GPU vs element order in the reference); whole Newton histories with the same iteration counts
and converged displacements within 1e-10 relative.
"""
import json
import os

import numpy as np
import pytest

import fe_driver as fd
import oracle_lib as ol
from ceos_amd import gpu, meshgen

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL = 1e-12

MATS = {
    "cnh": (ol.MODEL_CNH, [1.0, 99.3333333333], True),
    "nh": (ol.MODEL_NH, [0.5, 100.0], True),
    "svk": (ol.MODEL_SVK, [100.0, 0.3], True),
    "ghl": (ol.MODEL_GHL, [210000.0, 0.3], False),
    "j2": (ol.MODEL_J2, [210000.0, 0.3, 250.0, 1000.0], False),
    "nhq1p0": (ol.MODEL_NH_Q1P0, [0.5, 100.0], True),  # mean dilatation
}


def pair(X, conn, model, props, nl, mat_id=None, assembly=gpu.ASSEMBLY_GATHER):
    tech = gpu.ELEM_MEAN_DILATATION if model == ol.MODEL_NH_Q1P0 else gpu.ELEM_FULL_INTEGRATION
    g = gpu.CeosMesh(X, conn, mat_id=mat_id, nl_analysis=nl, assembly=assembly, elem_tech=tech)
    o = ol.OracleMesh(X, conn, mat_id=mat_id, nl_analysis=nl)
    ids = [1] if mat_id is None else sorted(set(int(m) for m in mat_id))
    for i in ids:
        g.set_material(i, model, props)
        o.set_material(i, model, props)
    return g, o


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def scrambled(n=4, seed=5):
    X, conn = meshgen.cube(n, perturb=0.12, seed=seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(X))
    return X[perm], (inv[conn - 1] + 1)[rng.permutation(len(conn))].astype(np.int32)


@pytest.mark.parametrize("mesh", ["cube3", "cube10", "scrambled4"])
def test_csr_pattern_bit_exact(mesh):
    if mesh == "cube3":
        X, conn = meshgen.cube(3)
    elif mesh == "cube10":
        X, conn = meshgen.cube(10)
    else:
        X, conn = scrambled()
    g, o = pair(X, conn, ol.MODEL_CNH, [1.0, 1.0], True)
    rg, cg = g.csr()
    ro, co = o.csr()
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
    if mesh == "cube10":
        assert len(cg) == meshgen.csr_nnz_structured(10)
    # full pattern: the symmetric closure of the upper one
    rf, cf = g.csr(upper=False)
    n = len(rg) - 1
    rows = np.repeat(np.arange(n), np.diff(rg))
    A = np.zeros((n, n), dtype=bool)
    A[rows, cg - 1] = True
    A |= A.T
    assert np.array_equal(np.diff(rf), A.sum(axis=1))
    assert np.array_equal(cf - 1, np.nonzero(A)[1])


@pytest.mark.parametrize("assembly", [gpu.ASSEMBLY_GATHER, gpu.ASSEMBLY_COLORED])
@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("mesh", ["cube4", "scrambled4"])
def test_fint_k_sigma_parity(name, mesh, assembly):
    model, props, nl = MATS[name]
    X, conn = meshgen.cube(4, perturb=0.1, seed=2) if mesh == "cube4" else scrambled()
    g, o = pair(X, conn, model, props, nl, assembly=assembly)
    rng = np.random.default_rng(7)
    scale = 0.03 if nl else (0.0004 if name == "j2" else 0.004)  # J2: mixed elastic/plastic
    U = rng.normal(scale=scale, size=3 * len(X))
    for m in (g, o):
        if nl:
            m.update_mesh(U)
        assert m.solve_constitutive(0.25, U) == 0
    fg, sg = g.internal_force()
    fo, so = o.internal_force()
    assert sg == 0 and so == 0
    assert rel(fg, fo) < RTOL, rel(fg, fo)
    Kg, Ko = g.tangent(), o.tangent()
    assert rel(Kg, Ko) < RTOL, rel(Kg, Ko)
    assert np.max(np.abs(Kg - Ko)) < RTOL * np.max(np.abs(Ko))
    assert rel(g.gp_stress(), o.gp_stress()) < RTOL
    if name == "j2":
        sg_, so_ = g.state(1), o.state(1)
        assert np.array_equal(sg_[..., 7], so_[..., 15])  # plastic flags identical
        assert 0 < sg_[..., 7].sum() < sg_[..., 7].size    # mixed elastic / plastic
        assert rel(sg_[..., :7], so_[..., :7]) < 1e-11


def test_multi_material_mesh():
    X, conn = meshgen.cube(4, perturb=0.1, seed=3)
    cent = X[conn - 1].mean(axis=1)
    mat = np.where(np.linalg.norm(cent - 0.5, axis=1) < 0.3, 2, 1).astype(np.int32)
    g = gpu.CeosMesh(X, conn, mat_id=mat)
    o = ol.OracleMesh(X, conn, mat_id=mat)
    for m in (g, o):
        m.set_material(1, ol.MODEL_NH, [0.5, 50.0])
        m.set_material(2, ol.MODEL_CNH, [10.0, 500.0 - 20.0 / 3.0])
    U = np.random.default_rng(1).normal(scale=0.02, size=3 * len(X))
    for m in (g, o):
        m.update_mesh(U)
        assert m.solve_constitutive(0.0, U) == 0
    assert rel(g.internal_force()[0], o.internal_force()[0]) < RTOL
    assert rel(g.tangent(), o.tangent()) < RTOL


def test_unfused_calls_match_fused():
    X, conn = meshgen.cube(3, perturb=0.1, seed=4)
    U = np.random.default_rng(2).normal(scale=0.02, size=3 * len(X))
    out = []
    for fused in (True, False):
        g = gpu.CeosMesh(X, conn, fused=fused)
        g.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
        g.update_mesh(U)
        assert g.solve_constitutive(0.0, U) == 0
        out.append((g.internal_force()[0], g.tangent()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_deterministic_bitwise():
    X, conn = meshgen.cube(6, perturb=0.1, seed=4)
    U = np.random.default_rng(3).normal(scale=0.02, size=3 * len(X))
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_NH, [0.5, 100.0])
    g.update_mesh(U)
    res = []
    for _ in range(2):
        assert g.solve_constitutive(0.0, U) == 0
        res.append((g.internal_force()[0], g.tangent()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


def test_detj_failure_reports_element():
    X, conn = meshgen.cube(3)
    conn = conn.copy()
    bad = 13
    conn[bad - 1, [0, 1]] = conn[bad - 1, [1, 0]]  # inverted element (reference test idea: swap nodes)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, 1.0])
    st = g.solve_constitutive(0.0, np.zeros(3 * len(X)))
    assert st == gpu.ERR_DETJ
    assert g.last_status.error == 1 and g.last_status.n_failed == 1
    assert list(g.failed_elements()) == [bad]
    o = ol.OracleMesh(X, conn)
    o.set_material(1, ol.MODEL_CNH, [1.0, 1.0])
    assert o.solve_constitutive(0.0, np.zeros(3 * len(X))) == -1
    assert list(o.failed_elements()) == [bad]
    # the next good evaluation clears the flags
    g2 = gpu.CeosMesh(*meshgen.cube(3))
    g2.set_material(1, ol.MODEL_CNH, [1.0, 1.0])
    assert g2.solve_constitutive(0.0, np.zeros(3 * 64)) == 0 and len(g2.failed_elements()) == 0


def _case(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_svk_checker_case_on_gpu():
    from test_oracle_golden import check_svk, run_svk_case
    case = _case("svk_hex8_case.json")
    g = gpu.CeosMesh(np.array(case["coords"]), np.array(case["conn"], dtype=np.int32))
    g.set_material(1, ol.MODEL_SVK, [case["material"]["E"], case["material"]["nu"]])
    case, norms, probe = run_svk_case(g)
    check_svk(case, norms, probe)


def test_nhq1p0_checker_case_on_gpu():
    """NonLinear_FEA_3D_Hexa8_NeoHookean_Incompressible (mean dilatation): the reference's
    residual log and the Marc displacements, through the GPU element path."""
    from test_oracle_golden import _case, check_nhq1p0, run_nhq1p0_case
    case = _case("nhq1p0_hex8_case.json")
    g = gpu.CeosMesh(np.array(case["coords"]), np.array(case["conn"], dtype=np.int32),
                     elem_tech=gpu.ELEM_MEAN_DILATATION)
    g.set_material(1, ol.MODEL_NH_Q1P0, [case["material"]["C10"], case["material"]["K"]])
    case, norms, probe = run_nhq1p0_case(g)
    check_nhq1p0(case, norms, probe)


def test_mean_dilatation_pairs_with_model_4_only():
    """The reference's library gives the mean-dilatation technology its Q1P0 models only."""
    X, conn = meshgen.cube(2)
    U = np.zeros(3 * len(X))
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_NH_Q1P0, [0.5, 100.0])
    assert g.solve_constitutive(0.0, U) == gpu.ERR_ARG
    h = gpu.CeosMesh(X, conn, elem_tech=gpu.ELEM_MEAN_DILATATION)
    h.set_material(1, ol.MODEL_NH, [0.5, 100.0])
    assert h.solve_constitutive(0.0, U) == gpu.ERR_ARG


@pytest.mark.parametrize("model", [ol.MODEL_GHL, ol.MODEL_J2])
def test_ghl_checker_case_on_gpu(model):
    from test_oracle_golden import check_ghl, ghl_backend

    def make(X, conn, model, props):
        m = gpu.CeosMesh(X, conn, nl_analysis=False)
        m.set_material(1, model, props)
        return m

    case, m = ghl_backend(make, model)
    check_ghl(case, m, m.gp_stress)


def _stretch_run(backend, X, nsteps, stretch=1.2):
    fixed, presc = meshgen.stretch_bc(X)
    ndof = 3 * len(X)
    bc = fd.BoundaryConditions(ndof, fixed, [(int(d), (lambda t: (stretch - 1.0) * t)) for d in presc])
    return fd.quasi_static(backend, bc, [(0.0, 1.0, nsteps)], nl=True, tol=1e-8, maxiter=15)


def test_c1_newton_history_matches_oracle():
    """10x10x10 unit cube, compressible Neo-Hookean (mu=1, K=100), x=0 fixed, stretch 1.2 on
    x=1 in 10 steps: same Newton iteration counts, converged U within 1e-10 relative."""
    X, conn = meshgen.cube(10)
    lam = 100.0 - 2.0 / 3.0
    g, o = pair(X, conn, ol.MODEL_CNH, [1.0, lam], True)
    hg = _stretch_run(g, X, 10)
    ho = _stretch_run(o, X, 10)
    assert [h[4] for h in hg] == [h[4] for h in ho]
    for a, b in zip(hg, ho):
        assert rel(a[3], b[3]) < 1e-10


def test_j2_cyclic_history_matches_oracle():
    """J2 with linear hardening on a 6^3 cube, cyclic uniaxial strain +-2% (linear analysis)."""
    X, conn = meshgen.cube(6)
    props = [210000.0, 0.3, 250.0, 1000.0]
    g, o = pair(X, conn, ol.MODEL_J2, props, False)
    fixed, presc = meshgen.stretch_bc(X)
    table = [(0.0, 0.0), (1.0, 0.02), (3.0, -0.02), (4.0, 0.0)]
    bc = fd.BoundaryConditions(3 * len(X), fixed,
                               [(int(d), (lambda t: fd.interp_table(table, t))) for d in presc])
    hs = []
    for m in (g, o):
        hs.append(fd.quasi_static(m, bc, [(0.0, 4.0, 16)], nl=False, tol=1e-10, maxiter=15))
    assert [h[4] for h in hs[0]] == [h[4] for h in hs[1]]
    for a, b in zip(*hs):
        assert rel(a[3], b[3]) < 1e-10
    sg, so = g.state(0), o.state(0)
    assert rel(sg[..., :7], so[..., :7]) < 1e-10
    assert sg[..., 6].max() > 0.0


@pytest.mark.parametrize("assembly", [gpu.ASSEMBLY_GATHER, gpu.ASSEMBLY_COLORED])
def test_assemblies_agree_and_residual_only_path(assembly):
    """Both assemblies give the oracle's K and Fint; the residual-only pass (tangent not
    requested) gives the same Fint."""
    X, conn = meshgen.cube(5, perturb=0.1, seed=21)
    U = np.random.default_rng(21).normal(scale=0.02, size=3 * len(X))
    g, o = pair(X, conn, ol.MODEL_NH, [0.5, 100.0], True, assembly=assembly)
    for m in (g, o):
        m.update_mesh(U)
        assert m.solve_constitutive(0.0, U) == 0
    fo = o.internal_force()[0]
    assert rel(g.internal_force()[0], fo) < RTOL and rel(g.tangent(), o.tangent()) < RTOL
    r = gpu.CeosMesh(X, conn, fused=False, assembly=assembly)
    r.set_material(1, ol.MODEL_NH, [0.5, 100.0])
    r.update_mesh(U)
    assert r.solve_constitutive(0.0, U) == 0
    assert rel(r.internal_force()[0], fo) < RTOL


def test_fp32_stored_tangent_inexact_newton():
    """FP32-stored tangent: K within FP32 rounding of the FP64 one; the Newton iteration driven
    by the FP64 residual converges to the FP64 solution."""
    X, conn = meshgen.cube(6, perturb=0.1, seed=8)
    lam = 100.0 - 2.0 / 3.0
    g64 = gpu.CeosMesh(X, conn)
    g32 = gpu.CeosMesh(X, conn, precision=gpu.K_FP32_STORED)
    for m in (g64, g32):
        m.set_material(1, ol.MODEL_CNH, [1.0, lam])
    U = np.random.default_rng(4).normal(scale=0.01, size=3 * len(X))
    for m in (g64, g32):
        m.update_mesh(U)
        assert m.solve_constitutive(0.0, U) == 0
    K64, K32 = g64.tangent(), g32.tangent()
    assert rel(K32, K64) < 2e-7
    assert np.array_equal(g64.internal_force()[0], g32.internal_force()[0])
    m32 = gpu.CeosMesh(X, conn, precision=gpu.K_FP32_STORED)
    m32.set_material(1, ol.MODEL_CNH, [1.0, lam])
    m64 = gpu.CeosMesh(X, conn)
    m64.set_material(1, ol.MODEL_CNH, [1.0, lam])
    h64 = _stretch_run(m64, X, 2)
    h32 = _stretch_run(m32, X, 2)
    assert rel(h32[-1][3], h64[-1][3]) < 1e-8


def test_homogenized_stress_of_affine_deformation():
    """Affine displacement on a distorted mesh: every GP has F = F_bar, so P_bar = P(F_bar) and
    F_bar is recovered; GPU and oracle agree."""
    X, conn = meshgen.cube(4, perturb=0.15, seed=6)
    Fb = np.array([[1.1, 0.05, 0.0], [0.02, 0.95, 0.03], [0.0, 0.01, 1.02]])
    U = (X @ (Fb - np.eye(3)).T).reshape(-1)
    g, o = pair(X, conn, ol.MODEL_NH, [0.5, 50.0], True)
    for m in (g, o):
        m.update_mesh(U)
        assert m.solve_constitutive(0.0, U) == 0
    Pg, Fg = g.homogenize()
    Po, Fo = o.homogenize()
    assert rel(Pg, Po) < RTOL and rel(Fg, Fo) < RTOL
    assert np.allclose(Fg.reshape(3, 3).T, Fb, atol=1e-13)
    s, _, _ = ol.material_point(ol.MODEL_NH, [0.5, 50.0], Fb)
    sig = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    P = np.linalg.det(Fb) * sig @ np.linalg.inv(Fb).T
    assert np.allclose(Pg.reshape(3, 3).T, P, rtol=1e-12, atol=1e-12)


def test_state_save_load_roundtrip(tmp_path):
    X, conn = meshgen.cube(3)
    g = gpu.CeosMesh(X, conn, nl_analysis=False)
    g.set_material(1, ol.MODEL_J2, [210000.0, 0.3, 250.0, 1000.0])
    U = np.zeros(3 * len(X))
    U[0::3] = 0.03 * X[:, 0]
    assert g.solve_constitutive(0.0, U) == 0
    g.switch_converged_state()
    s0 = g.state(0)
    path = str(tmp_path / "state.bin")
    g.save_state(path)
    g2 = gpu.CeosMesh(X, conn, nl_analysis=False)
    g2.set_material(1, ol.MODEL_J2, [210000.0, 0.3, 250.0, 1000.0])
    g2.load_state(path)
    assert np.array_equal(g2.state(0), s0)
    for m in (g, g2):
        assert m.solve_constitutive(0.0, 1.01 * U) == 0
    assert np.array_equal(g.internal_force()[0], g2.internal_force()[0])


def test_device_resident_path_matches_host_path():
    from ceos_amd import hiprt
    X, conn = meshgen.cube(5, perturb=0.1, seed=1)
    U = np.random.default_rng(5).normal(scale=0.02, size=3 * len(X))
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    f_host, k_host = g.internal_force()[0], g.tangent()
    dU = hiprt.DeviceArray.from_numpy(U)
    dF = hiprt.DeviceArray(len(U))
    dK = hiprt.DeviceArray(len(k_host))
    s = hiprt.Stream()
    g.update_mesh_device(dU.ptr, s.ptr)
    g.evaluate_device(0.0, dU.ptr, dF.ptr, dK.ptr, s.ptr)
    s.synchronize()
    assert np.array_equal(dF.numpy(), f_host)
    assert np.array_equal(dK.numpy(), k_host)
    code, st = g.device_status()
    assert code == 0 and st.error == 0


def test_full_size_c2_properties():
    """80^3 perturbed cube (512,000 hex8): closed-form nnz; for an affine displacement every
    Gauss point carries the same stress, interior residuals vanish and rigid translations are
    in the null space of the assembled tangent."""
    X, conn = meshgen.cube(80, perturb=0.1)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, [1.0, 100.0 - 2.0 / 3.0])
    nc, nnz = g.info()
    assert nnz == meshgen.csr_nnz_structured(80) == 63786006
    assert nc == 8
    A = np.array([[0.05, 0.01, 0.0], [0.0, -0.02, 0.01], [0.01, 0.0, 0.03]])
    U = (X @ A.T).reshape(-1)
    g.update_mesh(U)
    assert g.solve_constitutive(0.0, U) == 0
    f, st = g.internal_force()
    assert st == 0
    onb = np.zeros(len(X), dtype=bool)
    tol = 1e-12
    for k in range(3):
        onb |= (np.abs(X[:, k]) < tol) | (np.abs(X[:, k] - 1.0) < tol)
    fr = f.reshape(-1, 3)
    assert np.max(np.abs(fr[~onb])) < 1e-12 * np.max(np.abs(fr))
    K = g.tangent()
    rm, col = g.csr()
    for k in range(3):
        t = np.zeros(3 * len(X))
        t[k::3] = 1.0
        y = fd.sym_upper_matvec(rm, col, K, t)
        assert np.max(np.abs(y)) < 1e-10 * np.max(np.abs(K))


def _c2_state(X):
    """The step-1 state of the C2 stretch test that the benchmark evaluates (u_x = 0.04 X plus a
    smooth transverse field, F non-trivial at every Gauss point)."""
    U = np.zeros_like(X)
    U[:, 0] = 0.04 * X[:, 0]
    U[:, 1] = -0.006 * X[:, 1] + 0.002 * np.sin(3.1 * X[:, 0]) * X[:, 2]
    U[:, 2] = -0.006 * X[:, 2] + 0.002 * np.cos(2.3 * X[:, 1]) * X[:, 0]
    return U.reshape(-1)


def test_full_size_c2_matches_oracle():
    """The whole 80^3 perturbed cube (512,000 hex8, 63,786,006 tangent entries): the device's Fint,
    K and Gauss-point stresses at the benchmark's state against the oracle's on every entry
    (norm-wise 1e-12, as on the small meshes); the FP32-stored tangent of the same state against the
    FP64 one (2e-7: FP32 rounding of each stored entry) with Fint unchanged."""
    X, conn = meshgen.cube(80, perturb=0.1)
    U = _c2_state(X)
    props = [1.0, 100.0 - 2.0 / 3.0]
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, ol.MODEL_CNH, props)
    g.update_mesh(U)
    assert g.solve_constitutive(0.2, U) == 0
    fg, kg, sg = g.internal_force()[0], g.tangent(), g.gp_stress()
    g.close()
    o = ol.OracleMesh(X, conn, threads=min(len(os.sched_getaffinity(0)), 16))
    o.set_material(1, ol.MODEL_CNH, props)
    o.update_mesh(U)
    assert o.solve_constitutive(0.2, U) == 0
    fo, ko, so = o.internal_force()[0], o.tangent(), o.gp_stress()
    o.close()
    assert rel(fg, fo) < RTOL and rel(kg, ko) < RTOL and rel(sg, so) < RTOL, (rel(fg, fo), rel(kg, ko))
    g32 = gpu.CeosMesh(X, conn, precision=gpu.K_FP32_STORED)
    g32.set_material(1, ol.MODEL_CNH, props)
    g32.update_mesh(U)
    assert g32.solve_constitutive(0.2, U) == 0
    f32, k32 = g32.internal_force()[0], g32.tangent()
    g32.close()
    assert np.array_equal(f32, fg)
    assert rel(k32.astype(np.float64), kg) < 2e-7


def test_schedule_does_not_change_results():
    """Any launch schedule (slab count) gives the oracle's Fint and K to rounding, and a J2
    state committed under one schedule survives a switch to another."""
    X, conn = meshgen.cube(8, perturb=0.1, seed=9)
    U = np.random.default_rng(9).normal(scale=0.02, size=3 * len(X))
    o = ol.OracleMesh(X, conn)
    o.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
    o.update_mesh(U)
    assert o.solve_constitutive(0.0, U) == 0
    fo, ko = o.internal_force()[0], o.tangent()
    for slabs in (1, 3, 7):
        g = gpu.CeosMesh(X, conn)
        g.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
        g.set_schedule(slabs)
        g.update_mesh(U)
        assert g.solve_constitutive(0.0, U) == 0
        assert rel(g.internal_force()[0], fo) < RTOL and rel(g.tangent(), ko) < RTOL
    g = gpu.CeosMesh(X, conn, nl_analysis=False)
    g.set_material(1, ol.MODEL_J2, [210000.0, 0.3, 250.0, 1000.0])
    Ue = 0.05 * U
    assert g.solve_constitutive(0.0, Ue) == 0
    g.switch_converged_state()
    s0 = g.state(0)
    assert s0[..., 6].max() > 0
    g.set_schedule(5)
    assert np.array_equal(g.state(0), s0)


VISCO = [50.0, 1.0, 10.0, 20.0, 2.0, 5.0]  # K_inf, Mu_inf, Lambda_inf, K_e, Mu_e, Ni_v


def test_visco_parity_after_committed_step():
    X, conn = meshgen.cube(3, perturb=0.1, seed=12)
    g, o = pair(X, conn, ol.MODEL_VISCO, VISCO, True)
    rng = np.random.default_rng(12)
    U1 = rng.normal(scale=0.02, size=3 * len(X))
    U2 = U1 + rng.normal(scale=0.02, size=3 * len(X))
    for m in (g, o):
        m.update_mesh(U1)
        assert m.solve_constitutive(0.1, U1) == 0
        m.switch_converged_state()
        m.update_mesh(U2)
        assert m.solve_constitutive(0.35, U2) == 0
    assert rel(g.internal_force()[0], o.internal_force()[0]) < RTOL
    assert rel(g.tangent(), o.tangent()) < 1e-11
    assert rel(g.gp_stress(), o.gp_stress()) < RTOL
    sg, so = g.state(1), o.state(1)
    assert rel(sg, so) < 1e-11
    assert np.all(sg[..., 19] == 0.35)


def test_visco_relaxation_history_matches_oracle():
    """Stretch to 1.1 in 2 steps, then hold for 4 steps (relaxation): same Newton histories."""
    X, conn = meshgen.cube(3)
    g, o = pair(X, conn, ol.MODEL_VISCO, VISCO, True)
    fixed, presc = meshgen.stretch_bc(X)
    table = [(0.0, 0.0), (1.0, 0.1), (3.0, 0.1)]
    bc = fd.BoundaryConditions(3 * len(X), fixed, [(int(d), (lambda t: fd.interp_table(table, t))) for d in presc])
    hs = [fd.quasi_static(m, bc, [(0.0, 1.0, 2), (1.0, 3.0, 4)], nl=True, tol=1e-9, maxiter=15) for m in (g, o)]
    assert [h[4] for h in hs[0]] == [h[4] for h in hs[1]]
    for a, b in zip(*hs):
        assert rel(a[3], b[3]) < 1e-10
    assert rel(g.gp_stress(), o.gp_stress()) < 1e-10


def test_visco_local_newton_failure_status():
    """A NaN viscous modulus (set by the host) keeps the local residual from converging: status
    -2, as the reference's 'Max Iteration in Local Newton-Raphson' error."""
    X, conn = meshgen.cube(2)
    p = list(VISCO)
    p[4] = float("nan")
    g, o = pair(X, conn, ol.MODEL_VISCO, p, True)
    U = np.random.default_rng(1).normal(scale=0.01, size=3 * len(X))
    assert o.solve_constitutive(0.1, U) == -2
    assert g.solve_constitutive(0.1, U) == gpu.ERR_LOCAL_NR
    assert g.last_status.error == 1 and g.last_status.n_failed == len(conn)


@pytest.mark.parametrize("name", ["cnh", "nh", "svk", "visco"])
def test_numerical_tangent_matches_oracle(name):
    """Central-difference tangent (GetTangentModulusBase, h = 1e-8) on both sides: the two
    differ by the rounding of the perturbed stresses, amplified by 1/h, hence the 1e-6 bar;
    both stay within 1e-5 of the analytic tangent."""
    if name == "visco":
        model, props = ol.MODEL_VISCO, VISCO
    else:
        model, props, _ = MATS[name]
    X, conn = meshgen.cube(3, perturb=0.1, seed=4)
    g, o = pair(X, conn, model, props, True)
    rng = np.random.default_rng(4)
    U = rng.normal(scale=0.02, size=3 * len(X))
    for m in (g, o):
        m.update_mesh(U)
        assert m.solve_constitutive(0.2, U) == 0
    Ka = g.tangent()
    g.set_tangent_mode(True)
    o.set_tangent_mode(True)
    Kg, Ko = g.tangent(), o.tangent()
    assert rel(Kg, Ko) < 1e-6, rel(Kg, Ko)
    assert rel(Kg, Ka) < 1e-5, rel(Kg, Ka)
    assert rel(Ko, Ka) < 1e-5
    g.set_tangent_mode(False)
    assert np.array_equal(g.tangent(), Ka)


def test_pipeline_gives_bitwise_identical_results():
    """The pipelined schedule (element chunks on the caller's stream, gather groups on a second
    stream) changes only the order of independent launches: Fint and K are bitwise the same for
    every chunk count, with two materials so that launch ranges cross chunk boundaries, on the
    host-pointer and on the device-resident entries."""
    from ceos_amd import hiprt
    X, conn = meshgen.cube(12, perturb=0.1, seed=5)
    ne = len(conn)
    mat = np.where(np.arange(ne) % 3 == 0, 2, 1).astype(np.int32)
    U = np.random.default_rng(5).normal(scale=0.008, size=3 * len(X))
    out = []
    for chunks in (0, 4, 7, 13):
        g = gpu.CeosMesh(X, conn, mat_id=mat)
        g.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
        g.set_material(2, ol.MODEL_NH, [0.5, 50.0])
        g.set_pipeline(chunks)
        g.update_mesh(U)
        assert g.solve_constitutive(0.0, U) == 0
        f, st = g.internal_force()
        K = g.tangent()
        dU = hiprt.DeviceArray.from_numpy(U)
        dF = hiprt.DeviceArray(len(U))
        dK = hiprt.DeviceArray(len(K))
        g.evaluate_device(0.0, dU.ptr, dF.ptr, dK.ptr)
        hiprt.synchronize()
        out.append((f, K, dF.numpy(), dK.numpy()))
    for f, K, fd, kd in out[1:]:
        assert np.array_equal(f, out[0][0]) and np.array_equal(K, out[0][1])
        assert np.array_equal(fd, out[0][0]) and np.array_equal(kd, out[0][1])


@pytest.mark.parametrize("mesh", ["lexicographic", "scrambled"])
def test_residual_only_pass_matches_fint_of_full_evaluation(mesh):
    """The residual-only evaluation (no K: the element kernel writes only the 24 Fe doubles of
    each element, the Fint gather finds them from the diagonal contribution words) gives the
    Fint of the F+K evaluation bitwise (same Fe, same contribution order), with two materials,
    on a lexicographically numbered and on a scrambled mesh; and the oracle's Fint."""
    from ceos_amd import hiprt
    if mesh == "lexicographic":
        X, conn = meshgen.cube(6, perturb=0.1, seed=9)
    else:
        X, conn = scrambled(5, seed=9)
    ne = len(conn)
    mat = np.where(np.arange(ne) % 4 == 1, 2, 1).astype(np.int32)
    U = np.random.default_rng(9).normal(scale=0.01, size=3 * len(X))
    g = gpu.CeosMesh(X, conn, mat_id=mat)
    o = ol.OracleMesh(X, conn, mat_id=mat)
    for m in (g, o):
        m.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
        m.set_material(2, ol.MODEL_NH, [0.5, 50.0])
    _, nnz = g.info()
    dU = hiprt.DeviceArray.from_numpy(U)
    dF, dF2 = hiprt.DeviceArray(len(U)), hiprt.DeviceArray(len(U))
    dK = hiprt.DeviceArray(nnz)
    g.update_mesh_device(dU.ptr)
    g.evaluate_device(0.0, dU.ptr, dF.ptr, dK.ptr)
    g.evaluate_device(0.0, dU.ptr, dF2.ptr, None)
    hiprt.synchronize()
    assert g.device_status()[0] == 0
    assert np.array_equal(dF2.numpy(), dF.numpy())
    o.update_mesh(U)
    assert o.solve_constitutive(0.0, U) == 0
    assert rel(dF2.numpy(), o.internal_force()[0]) < RTOL
