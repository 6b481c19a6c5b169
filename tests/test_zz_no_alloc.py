"""ceos_evaluate_device_static (the evaluation with its element records in the library's static
device buffer, no allocation in the call) against ceos_evaluate_device through the C ABI:
bitwise equal Fint and K, two calls in a row, the second on a smaller mesh; residual only; a mesh
whose F+K records exceed the buffer is refused while its residual-only evaluation runs; an
unprepared handle is refused."""
import numpy as np
import pytest

from ceos_amd import gpu, hiprt, meshgen

pytestmark = pytest.mark.gpu


def _mesh(n, seed):
    X, conn = meshgen.cube(n, perturb=0.1, seed=seed)
    ne = len(conn)
    mat = np.where(np.arange(ne) % 5 == 2, 2, 1).astype(np.int32)
    g = gpu.CeosMesh(X, conn, mat_id=mat)
    g.set_material(1, gpu.MODEL_CNH, [1.0, 99.0])
    g.set_material(2, gpu.MODEL_NH, [0.5, 50.0])
    U = np.random.default_rng(seed).normal(scale=0.02 / n, size=3 * len(X))
    return g, U


def _both(g, U, want_k, s):
    """(Fint, K) of the allocating entry and of the static one on the same inputs."""
    _, nnz = g.info()
    dU = hiprt.DeviceArray.from_numpy(U)
    out = []
    for static in (False, True):
        dF = hiprt.DeviceArray(len(U))
        dF.upload(np.full(len(U), np.nan))
        dK = None
        if want_k:
            dK = hiprt.DeviceArray(nnz)
            dK.upload(np.full(nnz, np.nan))
        g.update_mesh_device(dU.ptr, s.ptr)
        fn = g.evaluate_device_static if static else g.evaluate_device
        fn(0.0, dU.ptr, dF.ptr, dK.ptr if want_k else None, s.ptr)
        s.synchronize()
        assert g.device_status()[0] == 0
        out.append((dF.numpy(), dK.numpy() if want_k else None))
    return out


def test_static_entry_matches_allocating_entry_twice_then_smaller():
    s = hiprt.Stream()
    for n, seed in ((14, 21), (9, 22)):  # the second call on a smaller mesh
        g, U = _mesh(n, seed)
        g.prepare(records=False)
        for want_k in (True, False):
            (f0, k0), (f1, k1) = _both(g, U, want_k, s)
            assert np.linalg.norm(f0) > 0
            assert np.array_equal(f0, f1), (n, want_k)
            if want_k:
                assert np.linalg.norm(k0) > 0 and np.array_equal(k0, k1), n
        g.close()


def test_static_entry_refuses_oversized_and_unprepared():
    s = hiprt.Stream()
    # 48^3 = 110,592 elements: F+K patch records need about 162 MiB > 128 MiB; residual only 20 MiB
    g, U = _mesh(48, 23)
    _, nnz = g.info()
    dU = hiprt.DeviceArray.from_numpy(U)
    dF, dK = hiprt.DeviceArray(len(U)), hiprt.DeviceArray(nnz)
    with pytest.raises(gpu.CeosError) as ei:
        g.evaluate_device_static(0.0, dU.ptr, dF.ptr, None, s.ptr)  # not prepared yet
    assert ei.value.code == gpu.ERR_ARG
    g.prepare(records=False)
    with pytest.raises(gpu.CeosError) as ei:
        g.evaluate_device_static(0.0, dU.ptr, dF.ptr, dK.ptr, s.ptr)
    assert ei.value.code == gpu.ERR_ARG
    (f0, _), (f1, _) = _both(g, U, False, s)
    assert np.array_equal(f0, f1)
