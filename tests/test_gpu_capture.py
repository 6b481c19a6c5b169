"""Graph capture of the device-resident evaluation and the ordering of the handle's own stream
after work queued on a caller's stream (libceos_gpu.so through its C ABI).

* After ceos_mesh_prepare the very first ceos_evaluate_device of a handle can be captured in a
  HIP graph; every replay (with eager work and synchronisations between replays) gives the eager
  result bitwise.
* Capturing the first evaluation of an unprepared handle is refused with CEOS_ERR_ARG (building
  the tables allocates and synchronises) and leaves the handle usable.
* ceos_switch_converged_state / ceos_get_state / ceos_device_status issued right after an
  evaluation queued on a caller's stream, with no synchronisation by the caller, see that
  evaluation's results.
"""
import numpy as np
import pytest

import oracle_lib as ol
from ceos_amd import gpu, hiprt, meshgen

pytestmark = pytest.mark.gpu

J2 = [210000.0, 0.3, 250.0, 1000.0]


def _j2_mesh(n):
    X, conn = meshgen.cube(n, perturb=0.1, seed=3)
    g = gpu.CeosMesh(X, conn, nl_analysis=False)
    g.set_material(1, gpu.MODEL_J2, J2)
    U = np.zeros(3 * len(X))
    U[0::3] = 0.004 * X[:, 0]  # past first yield everywhere
    U[1::3] = -0.001 * X[:, 1] + 0.0005 * np.sin(3.0 * X[:, 2])
    return X, conn, g, U


@pytest.mark.parametrize("model", ["cnh", "j2"])
def test_first_evaluation_after_prepare_is_capturable(model):
    X, conn = meshgen.cube(6, perturb=0.1, seed=4)
    nl = model == "cnh"
    g = gpu.CeosMesh(X, conn, nl_analysis=nl)
    if model == "cnh":
        g.set_material(1, gpu.MODEL_CNH, [1.0, 99.0])
    else:
        g.set_material(1, gpu.MODEL_J2, J2)
    U = np.random.default_rng(4).normal(scale=0.003, size=3 * len(X))
    _, nnz = g.info()
    dU = hiprt.DeviceArray.from_numpy(U)
    dF, dK = hiprt.DeviceArray(len(U)), hiprt.DeviceArray(nnz)
    s = hiprt.Stream()
    g.prepare()
    g.update_mesh_device(dU.ptr, s.ptr)
    s.synchronize()
    graph = hiprt.Graph(s, lambda: g.evaluate_device(0.0, dU.ptr, dF.ptr, dK.ptr, s.ptr))
    # the eager result on a second pair of buffers
    dF2, dK2 = hiprt.DeviceArray(len(U)), hiprt.DeviceArray(nnz)
    g.evaluate_device(0.0, dU.ptr, dF2.ptr, dK2.ptr, s.ptr)
    s.synchronize()
    assert g.device_status()[0] == 0
    f_ref, k_ref = dF2.numpy(), dK2.numpy()
    assert np.linalg.norm(f_ref) > 0 and np.linalg.norm(k_ref) > 0
    o = ol.OracleMesh(X, conn, nl_analysis=nl)
    o.set_material(1, ol.MODEL_CNH if model == "cnh" else ol.MODEL_J2, [1.0, 99.0] if model == "cnh" else J2)
    o.update_mesh(U)
    assert o.solve_constitutive(0.0, U) == 0
    assert np.linalg.norm(f_ref - o.internal_force()[0]) < 1e-12 * np.linalg.norm(f_ref)
    for rep in range(4):
        dF.upload(np.full(len(U), np.nan))
        dK.upload(np.full(nnz, np.nan))
        graph.launch(s)
        s.synchronize()
        assert np.array_equal(dF.numpy(), f_ref), rep
        assert np.array_equal(dK.numpy(), k_ref), rep
        # eager work between replays
        g.evaluate_device(0.0, dU.ptr, dF2.ptr, None, s.ptr)
        s.synchronize()
        assert g.device_status()[0] == 0


def test_capture_of_unprepared_first_evaluation_is_refused():
    X, conn = meshgen.cube(4, perturb=0.1, seed=6)
    g = gpu.CeosMesh(X, conn)
    g.set_material(1, gpu.MODEL_CNH, [1.0, 99.0])
    U = np.random.default_rng(6).normal(scale=0.003, size=3 * len(X))
    _, nnz = g.info()
    dU = hiprt.DeviceArray.from_numpy(U)
    dF, dK = hiprt.DeviceArray(len(U)), hiprt.DeviceArray(nnz)
    s = hiprt.Stream()
    with pytest.raises(gpu.CeosError) as ei:
        hiprt.Graph(s, lambda: g.evaluate_device(0.0, dU.ptr, dF.ptr, dK.ptr, s.ptr))
    assert ei.value.code == gpu.ERR_ARG
    # the handle is still usable eagerly
    g.update_mesh_device(dU.ptr, s.ptr)
    g.evaluate_device(0.0, dU.ptr, dF.ptr, dK.ptr, s.ptr)
    s.synchronize()
    assert g.device_status()[0] == 0
    o = ol.OracleMesh(X, conn)
    o.set_material(1, ol.MODEL_CNH, [1.0, 99.0])
    o.update_mesh(U)
    assert o.solve_constitutive(0.0, U) == 0
    f = o.internal_force()[0]
    assert np.linalg.norm(dF.numpy() - f) < 1e-12 * np.linalg.norm(f)


def test_handle_stream_entries_wait_for_caller_stream():
    """J2 at 24^3: evaluate on a caller's stream, then commit and read the state with no
    synchronisation in between; the committed state is that evaluation's trial state."""
    X, conn, g, U = _j2_mesh(24)
    _, nnz = g.info()
    g.prepare()
    dU = hiprt.DeviceArray.from_numpy(U)
    dF, dK = hiprt.DeviceArray(len(U)), hiprt.DeviceArray(nnz)
    s = hiprt.Stream()
    for _ in range(3):  # queue enough work that an unordered read would overtake it
        g.evaluate_device(0.0, dU.ptr, dF.ptr, dK.ptr, s.ptr)
    g.switch_converged_state()
    committed = g.state(0)
    code, _ = g.device_status()
    assert code == 0
    s.synchronize()
    trial = g.state(1)
    assert np.array_equal(committed, trial)
    assert trial[..., 7].mean() > 0.5  # plastic at most Gauss points
    # a detJ failure queued on the caller's stream is seen by ceos_device_status: a finite-strain
    # handle whose current coordinates fold the elements around one interior node
    Xc, connc = meshgen.cube(24, perturb=0.1, seed=3)
    h = gpu.CeosMesh(Xc, connc)
    h.set_material(1, gpu.MODEL_CNH, [1.0, 99.0])
    h.prepare()
    Ubad = np.zeros(3 * len(Xc))
    mid = int(np.argmin(np.linalg.norm(Xc - 0.5, axis=1)))
    Ubad[3 * mid] = 5.0
    dUb = hiprt.DeviceArray.from_numpy(Ubad)
    dFb, dKb = hiprt.DeviceArray(len(Ubad)), hiprt.DeviceArray(h.info()[1])
    h.update_mesh_device(dUb.ptr, s.ptr)
    h.evaluate_device(0.0, dUb.ptr, dFb.ptr, dKb.ptr, s.ptr)
    code, st = h.device_status()
    assert code == gpu.ERR_DETJ and 1 <= st.n_failed <= 8
