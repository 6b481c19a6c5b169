"""Parity of the batched periodic RVE solver (ceos_rve_*, one persistent launch) with the CPU
restatement of the reference's periodic analysis (tests/rve_oracle.py: full mesh, T^T K T,
direct solves).

Bar: homogenized P within 1e-9 relative and fluctuations within 1e-8 of the oracle's (both
Newton loops stop at max|R| <= 1e-10 max|R_0|, the GPU's CG at 1e-10, so the two converged
states differ by the solver tolerances, not by rounding); the condensed tangent within 1e-7.
"""
import numpy as np
import pytest

import oracle_lib as ol
import rve_oracle as ro
from ceos_amd import gpu, rve

pytestmark = pytest.mark.gpu

PROPS = np.array([rve.MATRIX_PROPS, rve.INCLUSION_PROPS])


def _F(seed, scale=0.06):
    rng = np.random.default_rng(seed)
    return np.eye(3) + rng.uniform(-scale, scale, size=(3, 3))


def _batch(sides, ids):
    ph = np.concatenate([rve.phases(n, i) for n, i in zip(sides, ids)])
    return gpu.RveBatch(sides, ph, gpu.MODEL_NH, PROPS), ph


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_homogeneous_rve_exact():
    b = gpu.RveBatch([4], np.zeros(64, dtype=np.uint8), gpu.MODEL_NH, PROPS)
    F = _F(1)
    P, _, info = b.solve(F[None])
    print("info", info, flush=True)
    s, _, _ = ol.material_point(ol.MODEL_NH, list(rve.MATRIX_PROPS), F)
    sig = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    Pm = np.linalg.det(F) * sig @ np.linalg.inv(F).T
    assert info[0, 0] == 0 and info[0, 1] == 0
    assert np.allclose(P[0], Pm, rtol=0, atol=1e-12 * np.abs(Pm).max())
    assert np.max(np.abs(b.fluctuation(0))) == 0.0


def test_mixed_batch_matches_oracle():
    sides = [6, 4, 8, 4, 6]
    ids = [3, 4, 5, 6, 7]
    b, ph = _batch(sides, ids)
    Fs = np.stack([_F(10 + k) for k in range(len(sides))])
    P, _, info = b.solve(Fs)
    print("info", info.tolist(), flush=True)
    assert np.all(info[:, 0] == 0)
    off = 0
    for k, (n, i) in enumerate(zip(sides, ids)):
        o = ro.OracleRVE(n, ph[off:off + n ** 3])
        off += n ** 3
        Po, _, _, _ = o.solve(Fs[k])
        Po = Po.reshape(3, 3).T
        print(k, n, rel(P[k], Po), flush=True)
        assert rel(P[k], Po) < 1e-9
        assert np.max(np.abs(b.fluctuation(k) - o.w)) < 1e-8 * max(np.max(np.abs(o.w)), 1e-3)


def test_condensed_tangent_matches_oracle():
    sides = [4, 6]
    b, ph = _batch(sides, [11, 12])
    b.set_solver(newton_rtol=1e-12, pcg_rtol=1e-12)
    Fs = np.stack([_F(20), _F(21)])
    P, A, info = b.solve(Fs, tangent=True)
    print("info", info.tolist(), flush=True)
    off = 0
    for k, n in enumerate(sides):
        o = ro.OracleRVE(n, ph[off:off + n ** 3])
        off += n ** 3
        o.solve(Fs[k], rtol=1e-12)
        Ao = o.tangent()
        print(k, rel(A[k], Ao), flush=True)
        assert rel(A[k], Ao) < 1e-7
        assert np.allclose(A[k], A[k].T, atol=1e-8 * np.abs(A[k]).max())  # hyperelastic: major symmetry


def test_condensed_tangent_matches_oracle_at_side_12():
    """The condensation at a workload side: three aggregates per axis in the coarse level, nine
    solves in three groups of three with FP32 search directions; FP64 tangent storage."""
    b, ph = _batch([12], [13])
    b.set_solver(newton_rtol=1e-12, pcg_rtol=1e-12)
    F = _F(22, 0.04)
    P, A, info = b.solve(F[None], tangent=True)
    print("info", info.tolist(), flush=True)
    o = ro.OracleRVE(12, ph)
    o.solve(F, rtol=1e-12)
    Ao = o.tangent()
    print("tangent rel", rel(A[0], Ao), flush=True)
    assert rel(A[0], Ao) < 1e-7


def test_incremental_solves_and_cut_back():
    """Two macro increments from the converged state, then a large stretch reached by the
    golden-ratio cut-back; the states match the oracle driven through the same increments."""
    sides = [4, 6]
    b, ph = _batch(sides, [7, 8])
    os_ = [ro.OracleRVE(n, ph[o:o + n ** 3]) for n, o in zip(sides, [0, 64])]
    F1 = np.stack([_F(30, 0.03), _F(31, 0.03)])
    F2 = np.stack([np.diag([1.25, 0.8, 1.1]), np.diag([0.85, 1.2, 1.05])])
    for Fs in (F1, F2):
        P, _, info = b.solve(Fs)
        print("info", info.tolist(), flush=True)
        for k, o in enumerate(os_):
            Po, _, _, _ = o.solve(Fs[k])
            assert rel(P[k], Po.reshape(3, 3).T) < 1e-9
    assert info[0, 3] >= 1  # diag(1.25, 0.8, 1.1) from a mild state needs a cut-back


def test_queue_longer_than_the_machine():
    """More RVEs than workgroup slots: every RVE converges and a sample matches the oracle."""
    n_rve = 600
    sides = np.full(n_rve, 4, dtype=np.int32)
    sides[::7] = 6
    ids = np.arange(n_rve)
    b, ph = _batch(sides, ids)
    rng = np.random.default_rng(5)
    Fs = np.eye(3)[None] + rng.uniform(-0.04, 0.04, size=(n_rve, 3, 3))
    P, _, info = b.solve(Fs)
    assert np.all(info[:, 0] == 0)
    offs = np.concatenate([[0], np.cumsum(sides.astype(np.int64) ** 3)])
    for k in (0, 1, 7, 299, 599):
        n = int(sides[k])
        o = ro.OracleRVE(n, ph[offs[k]:offs[k] + n ** 3])
        Po, _, _, _ = o.solve(Fs[k])
        assert rel(P[k], Po.reshape(3, 3).T) < 1e-9


@pytest.mark.parametrize("n", [12, 16])
def test_workload_sizes_match_oracle(n):
    b, ph = _batch([n], [40 + n])
    F = _F(50 + n, 0.04)
    P, _, info = b.solve(F[None])
    print("info", info.tolist(), flush=True)
    o = ro.OracleRVE(n, ph)
    Po, _, _, _ = o.solve(F)
    assert rel(P[0], Po.reshape(3, 3).T) < 1e-9


def test_fp32_stored_tangent_matches_fp64():
    """FP32 storage of the reduced tangent: the Newton residual stays FP64, so P reaches the same
    tolerance; the condensed tangent carries the storage rounding."""
    sides = [6, 8]
    out = []
    for mode in (gpu.K_FP64, gpu.K_FP32_STORED):
        b, _ = _batch(sides, [21, 22])
        b.set_precision(mode)
        Fs = np.stack([_F(60), _F(61)])
        P, A, info = b.solve(Fs, tangent=True)
        print(mode, info.tolist(), flush=True)
        assert np.all(info[:, 0] == 0)
        out.append((P, A))
    (P64, A64), (P32, A32) = out
    for k in range(len(sides)):
        print(k, rel(P32[k], P64[k]), rel(A32[k], A64[k]), flush=True)
        assert rel(P32[k], P64[k]) < 1e-9
        assert rel(A32[k], A64[k]) < 1e-5


def test_two_level_preconditioner_agrees_and_cuts_iterations():
    """The coarse level of the CG preconditioner (translations of 4^3-node aggregates) changes
    only the CG iterates: solves and condensed tangents with and without it agree to the solver
    tolerances on every aggregate layout of sides 4..16 (one aggregate; end aggregates two nodes
    wide; 2, 3 and 4 aggregates per axis), in FP64 and FP32 tangent storage, and from side 8 on
    the CG needs fewer iterations."""
    sides = [4, 6, 8, 10, 12, 14, 16]
    Fs = np.stack([_F(90 + k, 0.05) for k in range(len(sides))])
    for mode in (gpu.K_FP64, gpu.K_FP32_STORED):
        out = {}
        for on in (False, True):
            b, _ = _batch(sides, [80 + k for k in range(len(sides))])
            b.set_precision(mode)
            b.set_coarse(on)
            P, A, info = b.solve(Fs, tangent=True)
            print("precision", mode, "coarse", on, info.tolist(), flush=True)
            assert np.all(info[:, 0] == 0)
            out[on] = (P, A, info)
        for k, n in enumerate(sides):
            print(n, rel(out[True][0][k], out[False][0][k]), rel(out[True][1][k], out[False][1][k]), flush=True)
            assert rel(out[True][0][k], out[False][0][k]) < 1e-9
            assert rel(out[True][1][k], out[False][1][k]) < 1e-7
            if n >= 8:
                assert out[True][2][k, 2] < out[False][2][k, 2]


def test_condensation_warm_start_agrees_and_saves_iterations():
    """The RVE's last condensation solutions (default on) start the next condensation and, as
    dw/dF, predict the next Newton solve's first iterate: P and the tangent agree with the cold
    solves to the solver tolerances, with no more Newton steps and fewer CG iterations; FP32-stored
    tangent as in the FE^2 workload."""
    sides = [8, 12]
    F1 = np.stack([_F(81, 0.04), _F(82, 0.04)])
    F2 = F1 + 2e-3 * np.stack([np.eye(3) + 0.3 * np.ones((3, 3)), np.eye(3) - 0.2 * np.ones((3, 3))])
    out = {}
    for warm in (False, True):
        b, _ = _batch(sides, [91, 92])
        b.set_precision(gpu.K_FP32_STORED)
        b.set_tangent_rtol(1e-8)
        b.set_warm_start(warm)
        b.solve(F1, tangent=True)
        P, A, info = b.solve(F2, tangent=True)
        print("warm", warm, info.tolist(), flush=True)
        assert np.all(info[:, 0] == 0)
        out[warm] = (P, A, info)
    for k in range(len(sides)):
        print(k, rel(out[True][1][k], out[False][1][k]), flush=True)
        assert rel(out[True][0][k], out[False][0][k]) < 1e-9
        assert rel(out[True][1][k], out[False][1][k]) < 1e-6
        assert out[True][2][k, 1] <= out[False][2][k, 1]
        assert out[True][2][k, 2] < out[False][2][k, 2]


def test_autotune_picks_a_tile_and_every_tile_agrees():
    """The CG sweep tile changes only the order of independent loads: solves and condensed
    tangents with every tile agree to the solver tolerances; the autotuner times all three."""
    sides = [4, 6, 8, 6]
    b, _ = _batch(sides, [31, 32, 33, 34])
    g, t = b.autotune(max_rves=4)
    print("autotune", g, t, flush=True)
    assert g in (2, 4, 8) and np.all(t > 0)
    Fs = np.stack([_F(70 + k) for k in range(len(sides))])
    out = {}
    for tile in (2, 4, 8):
        b.reset()
        b.set_cg_group(tile)
        P, A, info = b.solve(Fs, tangent=True)
        assert np.all(info[:, 0] == 0)
        out[tile] = (P, A)
    for tile in (2, 8):
        for k in range(len(sides)):
            assert rel(out[tile][0][k], out[4][0][k]) < 1e-9
            assert rel(out[tile][1][k], out[4][1][k]) < 1e-7


def test_state_save_load_continues_bitwise(tmp_path):
    """A batch restarted from ceos_rve_state_save continues exactly as the uninterrupted batch
    (the restart of a multiscale analysis, FEM/ModFEMAnalysis.f90:2827-2956, rebuilds the same
    converged RVE state); a state file does not load into a batch of other RVEs."""
    sides, ids = [6, 4, 8], [11, 12, 13]
    F1 = np.stack([_F(40 + k, 0.04) for k in range(3)])
    F2 = np.stack([_F(50 + k, 0.05) for k in range(3)])
    a, _ = _batch(sides, ids)
    a.solve(F1, tangent=True)  # the condensation's solutions warm-start the next one, and travel too
    path = str(tmp_path / "rve.state")
    a.state_save(path)
    P_a, A_a, info_a = a.solve(F2, tangent=True)
    b, _ = _batch(sides, ids)
    b.state_load(path)
    P_b, A_b, info_b = b.solve(F2, tangent=True)
    print("info", info_a.tolist(), info_b.tolist(), flush=True)
    assert np.array_equal(P_a, P_b) and np.array_equal(A_a, A_b) and np.array_equal(info_a, info_b)
    for r in range(3):
        assert np.array_equal(a.fluctuation(r), b.fluctuation(r))
    other, _ = _batch([6, 4, 6], ids)
    with pytest.raises(gpu.CeosError):
        other.state_load(path)
