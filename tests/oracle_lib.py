"""ctypes wrapper of the CPU oracle (oracle/libceos_oracle.so) -- test infrastructure only.

The oracle restates the reference element loops in plain C (see oracle/ceos_oracle.c for the
file:line each routine follows).  Tests use it as the checker of the HIP library.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "libceos_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

MODEL_GHL, MODEL_J2, MODEL_NH, MODEL_SVK, MODEL_CNH, MODEL_VISCO = 1, 2, 3, 5, 7, 12
MODEL_NH_Q1P0 = 4  # mean dilatation (Q1P0) Neo-Hookean
STATE_STRIDE = 20  # per Gauss point in or_get_state / or_set_state

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.or_mesh_create.restype = C.c_void_p
        L.or_mesh_create.argtypes = [C.c_int, _dp, C.c_int, _ip, _ip]
        L.or_mesh_destroy.argtypes = [C.c_void_p]
        L.or_material_set.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_int]
        L.or_set_analysis.argtypes = [C.c_void_p, C.c_int]
        L.or_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.or_set_tangent_mode.argtypes = [C.c_void_p, C.c_int]
        L.or_csr_nnz.restype = C.c_long
        L.or_csr_nnz.argtypes = [C.c_void_p]
        L.or_csr_pattern.argtypes = [C.c_void_p, _ip, _ip]
        L.or_solve_constitutive.argtypes = [C.c_void_p, C.c_double, _dp]
        L.or_internal_force.argtypes = [C.c_void_p, _dp]
        L.or_tangent.argtypes = [C.c_void_p, _dp]
        L.or_update_mesh.argtypes = [C.c_void_p, _dp]
        L.or_switch_converged_state.argtypes = [C.c_void_p]
        L.or_failed_elements.argtypes = [C.c_void_p, _ip, C.c_int]
        L.or_get_gp_stress.argtypes = [C.c_void_p, _dp]
        L.or_get_gp_F.argtypes = [C.c_void_p, _dp]
        L.or_get_state.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_set_state.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_element_stiffness.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_element_internal_force.argtypes = [C.c_void_p, C.c_int, _dp]
        L.or_element_volume.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_homogenize.argtypes = [C.c_void_p, _dp, _dp]
        L.or_hex8_gauss.argtypes = [_dp, _dp]
        L.or_hex8_shape.argtypes = [_dp, _dp]
        L.or_hex8_dshape.argtypes = [_dp, _dp]
        L.or_matrix_bg.argtypes = [_dp, _dp, _dp, _dp, C.POINTER(C.c_double)]
        L.or_deformation_gradient.argtypes = [_dp, _dp, _dp, _dp]
        L.or_model_nstate.argtypes = [C.c_int]
        L.or_material_point.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_void_p, _dp, _dp, C.c_void_p]
        _lib = L
    return _lib


def material_point(model, props, F, time=0.0, state_old=None):
    """One update + tangent at a material point. Returns (sigma6, D 6x6, state_new)."""
    L = lib()
    ns = L.or_model_nstate(model)
    props = np.ascontiguousarray(props, dtype=np.float64)
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(9))
    s = np.zeros(6)
    D = np.zeros(36)
    nw = np.zeros(max(ns, 1))
    old = None
    if state_old is not None:
        old = np.ascontiguousarray(state_old, dtype=np.float64)
    st = L.or_material_point(model, props, F, time,
                             old.ctypes.data if old is not None else None, s, D,
                             nw.ctypes.data)
    if st:
        raise RuntimeError("or_material_point status %d" % st)
    return s, D.reshape(6, 6), nw[:ns]


class OracleMesh:
    """Mirror of the reference element loops on one mesh (1-based CSR, DOF = 3(node-1)+k)."""

    def __init__(self, X, conn1, mat_id=None, nl_analysis=True, threads=1):
        L = lib()
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        conn1 = np.ascontiguousarray(conn1, dtype=np.int32).reshape(-1, 8)
        self.nn, self.ne = X.shape[0], conn1.shape[0]
        if mat_id is None:
            mat_id = np.ones(self.ne, dtype=np.int32)
        mat_id = np.ascontiguousarray(mat_id, dtype=np.int32)
        self._h = L.or_mesh_create(self.nn, X.reshape(-1), self.ne, conn1.reshape(-1), mat_id)
        if not self._h:
            raise ValueError("or_mesh_create failed")
        self.ndof = 3 * self.nn
        L.or_set_analysis(self._h, 1 if nl_analysis else 0)
        L.or_set_threads(self._h, threads)
        self._csr = None

    def close(self):
        if self._h:
            lib().or_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_material(self, mat_id, model, props):
        p = np.ascontiguousarray(props, dtype=np.float64)
        st = lib().or_material_set(self._h, mat_id, model, p, len(p))
        if st:
            raise ValueError("or_material_set status %d" % st)

    def set_tangent_mode(self, numerical):
        lib().or_set_tangent_mode(self._h, 1 if numerical else 0)

    def set_threads(self, n):
        lib().or_set_threads(self._h, n)

    def csr(self):
        if self._csr is None:
            L = lib()
            nnz = L.or_csr_nnz(self._h)
            if nnz < 0:
                raise RuntimeError("pattern failed")
            rm = np.zeros(self.ndof + 1, dtype=np.int32)
            col = np.zeros(nnz, dtype=np.int32)
            st = L.or_csr_pattern(self._h, rm, col)
            if st:
                raise RuntimeError("or_csr_pattern status %d" % st)
            self._csr = (rm, col)
        return self._csr

    def solve_constitutive(self, time, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        return lib().or_solve_constitutive(self._h, time, U)

    def internal_force(self):
        F = np.zeros(self.ndof)
        st = lib().or_internal_force(self._h, F)
        return F, st

    def tangent(self):
        rm, col = self.csr()
        K = np.zeros(len(col))
        st = lib().or_tangent(self._h, K)
        if st:
            raise RuntimeError("or_tangent status %d" % st)
        return K

    def update_mesh(self, U):
        lib().or_update_mesh(self._h, np.ascontiguousarray(U, dtype=np.float64))

    def switch_converged_state(self):
        lib().or_switch_converged_state(self._h)

    def failed_elements(self):
        L = lib()
        n = L.or_failed_elements(self._h, np.zeros(1, dtype=np.int32), 0)
        ids = np.zeros(max(n, 1), dtype=np.int32)
        L.or_failed_elements(self._h, ids, n)
        return ids[:n]

    def gp_stress(self):
        out = np.zeros(self.ne * 8 * 6)
        lib().or_get_gp_stress(self._h, out)
        return out.reshape(self.ne, 8, 6)

    def gp_F(self):
        out = np.zeros(self.ne * 8 * 9)
        lib().or_get_gp_F(self._h, out)
        return out.reshape(self.ne, 8, 3, 3)

    def gp_deformation_gradient(self, U):
        """The library's accessor: F at the Gauss points for U; here those of the last
        solve_constitutive, which the callers made at the same U."""
        return self.gp_F()

    def state(self, which=1):
        out = np.zeros(self.ne * 8 * STATE_STRIDE)
        lib().or_get_state(self._h, which, out)
        return out.reshape(self.ne, 8, STATE_STRIDE)

    def set_state(self, which, arr):
        lib().or_set_state(self._h, which, np.ascontiguousarray(arr, dtype=np.float64).reshape(-1))

    def element_stiffness(self, e):
        Ke = np.zeros(576)
        st = lib().or_element_stiffness(self._h, e, Ke)
        if st:
            raise RuntimeError("or_element_stiffness status %d" % st)
        return Ke.reshape(24, 24)

    def element_internal_force(self, e):
        Fe = np.zeros(24)
        st = lib().or_element_internal_force(self._h, e, Fe)
        if st:
            raise RuntimeError("or_element_internal_force status %d" % st)
        return Fe

    def element_volume(self, e):
        V, VX = C.c_double(), C.c_double()
        lib().or_element_volume(self._h, e, C.byref(V), C.byref(VX))
        return V.value, VX.value

    def homogenize(self):
        P, F = np.zeros(9), np.zeros(9)
        lib().or_homogenize(self._h, P, F)
        return P, F
