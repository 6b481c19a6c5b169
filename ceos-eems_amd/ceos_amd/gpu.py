"""Python mirror of the reference element-layer interface, bound to libceos_gpu.so (C ABI).

Method names follow the reference subroutines they replace (paths relative to SourceCode/):
  solve_constitutive      SolveConstitutiveModel   FEM/Subroutines/SubSolveConstitutiveModel.f90:13
  internal_force          InternalForce            FEM/Subroutines/SubInternalForce.f90:13
  tangent                 TangentStiffnessMatrix   FEM/Subroutines/SubTangentStiffnessMatrix.f90:13
  update_mesh             UpdateMeshCoordinates    FEM/Subroutines/SubUpdateMeshCoordinates.f90:1
  switch_converged_state  SwitchConvergedState     FEM/ModFEMAnalysis.f90:1349-1353
  homogenize              HomogenizeStress / HomogenizeDeformationGradient
                                                   MultiScale/ModMultiscaleAnalysis.f90:140-441
  csr                     AllocateKgSparseUpperTriangular  FEM/ModFEMAnalysis.f90:209-273
Status codes mirror ClassStatus (Tools/ModStatus.f90): 0 ok, -1 detJ <= 1e-13, ...

There is no CPU fallback: constructing a mesh without the built library or without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libceos_gpu.so")

OK, ERR_DETJ, ERR_LOCAL_NR, ERR_SLOT, ERR_ARG, ERR_HIP = 0, -1, -2, -3, -4, -5
PREC_BLOCK_JACOBI, PREC_TWO_LEVEL = 0, 1
MODEL_GHL, MODEL_J2, MODEL_NH, MODEL_SVK, MODEL_CNH, MODEL_VISCO = 1, 2, 3, 5, 7, 12
MODEL_EXT = 99  # FE^2 macro point: S and dS/dE per Gauss point from the caller
MODEL_NH_Q1P0 = 4  # mean-dilatation Neo-Hookean (with ELEM_MEAN_DILATATION)
ELEM_FULL_INTEGRATION, ELEM_MEAN_DILATATION = 1, 2
SOLVER_AUTO, SOLVER_PCG, SOLVER_MINRES = 0, 1, 2
EXT_PER_GP = 27
STATE_STRIDE = 20
K_FP64, K_FP32_STORED = 0, 1
ASSEMBLY_GATHER, ASSEMBLY_COLORED = 0, 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class CeosStatus(C.Structure):
    _fields_ = [("error", C.c_int), ("error_id", C.c_int), ("n_failed", C.c_int),
                ("description", C.c_char * 256)]


class CeosError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s failed with status %d" % (what, code))
        self.code = code


EXPORTS = [
    "ceos_mesh_create", "ceos_mesh_destroy", "ceos_material_set", "ceos_set_analysis",
    "ceos_set_precision", "ceos_csr_pattern", "ceos_solve_constitutive", "ceos_internal_force",
    "ceos_tangent", "ceos_tangent_f32", "ceos_update_mesh", "ceos_switch_converged_state",
    "ceos_failed_elements", "ceos_get_gp_stress", "ceos_get_state", "ceos_set_state",
    "ceos_homogenize", "ceos_state_save", "ceos_state_load", "ceos_set_fused",
    "ceos_evaluate_device", "ceos_update_mesh_device", "ceos_device_status", "ceos_mesh_info",
    "ceos_set_schedule", "ceos_set_assembly", "ceos_set_tangent_mode",
    "ceos_timing_start", "ceos_timing_stop",
    "ceos_rve_batch_create", "ceos_rve_batch_destroy", "ceos_rve_set_solver", "ceos_rve_reset",
    "ceos_rve_solve", "ceos_rve_solve_device", "ceos_rve_get_fluctuation", "ceos_rve_set_inexact",
    "ceos_rve_profile",
    "ceos_mlp_create", "ceos_mlp_destroy", "ceos_mlp_num_params", "ceos_mlp_get_params",
    "ceos_mlp_set_params", "ceos_mlp_gradients", "ceos_mlp_train_step", "ceos_mlp_forward",
    "ceos_set_external_gp", "ceos_gp_deformation_gradient", "ceos_gp_deformation_gradient_device",
    "ceos_rve_solve_subset", "ceos_rve_to_external_device", "ceos_mlp_shape",
    "ceos_surrogate_create", "ceos_surrogate_destroy", "ceos_surrogate_features_device",
    "ceos_surrogate_eval_device", "ceos_rve_set_precision", "ceos_rve_set_tangent_rtol", "ceos_rve_set_coarse", "ceos_rve_set_warm_start",
    "ceos_mlp_gradients_weighted", "ceos_mlp_train_step_weighted", "ceos_mlp_set_training_precision",
    "ceos_set_pipeline", "ceos_rve_autotune", "ceos_rve_set_cg_group",
    "ceos_bc_set", "ceos_residual_bc", "ceos_residual_bc_device", "ceos_set_solver_precision",
    "ceos_linear_solve", "ceos_linear_solve_device", "ceos_newton_update_device", "ceos_set_linear_solver",
    "ceos_mesh_prepare", "ceos_evaluate_device_static", "ceos_assembly_info",
    "ceos_rve_state_save", "ceos_rve_state_load", "ceos_linear_solve_device_async",
    "ceos_set_preconditioner", "ceos_preconditioner_info", "ceos_coarse_space",
    "ceos_periodic_create", "ceos_periodic_destroy", "ceos_periodic_info", "ceos_periodic_reduce_device",
    "ceos_periodic_expand_device", "ceos_periodic_homogenize_device", "ceos_spmv_device",
]


class SurrogateCfg(C.Structure):
    _fields_ = [("u_scale", C.c_double), ("lam_lo", C.c_double), ("lam_hi", C.c_double), ("e_ref", C.c_double),
                ("d_mean", C.c_double * 21), ("d_std", C.c_double * 21), ("d_lo", C.c_double * 21),
                ("d_hi", C.c_double * 21), ("t_mean", C.c_double * 6), ("t_std", C.c_double * 6)]

_lib = None


def load_library(path=LIB_PATH):
    """Load libceos_gpu.so and declare the C ABI. Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libceos_gpu.so is not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.ceos_mesh_create.argtypes = [i, _dp, i, i, _ip, _ip, C.POINTER(vp)]
    L.ceos_mesh_destroy.argtypes = [vp]
    L.ceos_mesh_destroy.restype = None
    L.ceos_material_set.argtypes = [vp, i, i, _dp, i]
    L.ceos_set_analysis.argtypes = [vp, i, i]
    L.ceos_set_precision.argtypes = [vp, i]
    L.ceos_set_fused.argtypes = [vp, i]
    L.ceos_csr_pattern.argtypes = [vp, i, C.POINTER(i), C.POINTER(C.c_longlong), vp, vp]
    L.ceos_solve_constitutive.argtypes = [vp, d, _dp, C.POINTER(CeosStatus)]
    L.ceos_internal_force.argtypes = [vp, _dp, C.POINTER(CeosStatus)]
    L.ceos_tangent.argtypes = [vp, _dp]
    L.ceos_tangent_f32.argtypes = [vp, _fp]
    L.ceos_update_mesh.argtypes = [vp, _dp]
    L.ceos_switch_converged_state.argtypes = [vp]
    L.ceos_failed_elements.argtypes = [vp, vp, i]
    L.ceos_get_gp_stress.argtypes = [vp, _dp]
    L.ceos_get_state.argtypes = [vp, i, _dp]
    L.ceos_set_state.argtypes = [vp, i, _dp]
    L.ceos_homogenize.argtypes = [vp, _dp, _dp]
    L.ceos_state_save.argtypes = [vp, C.c_char_p]
    L.ceos_state_load.argtypes = [vp, C.c_char_p]
    L.ceos_evaluate_device.argtypes = [vp, d, vp, vp, vp, vp]
    L.ceos_mesh_prepare.argtypes = [vp, i]
    L.ceos_evaluate_device_static.argtypes = [vp, d, vp, vp, vp, vp]
    L.ceos_update_mesh_device.argtypes = [vp, vp, vp]
    L.ceos_device_status.argtypes = [vp, C.POINTER(CeosStatus)]
    L.ceos_mesh_info.argtypes = [vp, C.POINTER(i), C.POINTER(C.c_longlong)]
    L.ceos_assembly_info.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.ceos_rve_state_save.argtypes = [vp, C.c_char_p]
    L.ceos_rve_state_load.argtypes = [vp, C.c_char_p]
    L.ceos_linear_solve_device_async.argtypes = [vp, vp, vp, vp, d, i, vp, vp]
    L.ceos_set_preconditioner.argtypes = [vp, i, d, i]
    L.ceos_preconditioner_info.argtypes = [vp, C.POINTER(i), C.POINTER(i)]
    L.ceos_coarse_space.argtypes = [vp, vp, vp, vp]
    L.ceos_set_schedule.argtypes = [vp, i]
    L.ceos_set_pipeline.argtypes = [vp, i]
    L.ceos_set_assembly.argtypes = [vp, i]
    L.ceos_set_tangent_mode.argtypes = [vp, i]
    L.ceos_timing_start.argtypes = [vp, i]
    L.ceos_timing_stop.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.ceos_rve_batch_create.argtypes = [i, _ip, vp, i, _dp, i, C.POINTER(vp)]
    L.ceos_rve_batch_destroy.argtypes = [vp]
    L.ceos_rve_batch_destroy.restype = None
    L.ceos_rve_set_solver.argtypes = [vp, d, d, i, d, i, i]
    L.ceos_rve_reset.argtypes = [vp]
    L.ceos_rve_set_inexact.argtypes = [vp, i]
    L.ceos_rve_set_precision.argtypes = [vp, i]
    L.ceos_rve_set_tangent_rtol.argtypes = [vp, d]
    L.ceos_rve_set_coarse.argtypes = [vp, i]
    L.ceos_rve_set_warm_start.argtypes = [vp, i]
    L.ceos_rve_autotune.argtypes = [vp, i, C.POINTER(i), _dp]
    L.ceos_rve_set_cg_group.argtypes = [vp, i]
    L.ceos_rve_profile.argtypes = [vp, vp]
    f = C.c_float
    L.ceos_mlp_create.argtypes = [i, i, i, i, C.c_ulonglong, C.POINTER(vp)]
    L.ceos_mlp_destroy.argtypes = [vp]
    L.ceos_mlp_destroy.restype = None
    L.ceos_mlp_num_params.argtypes = [vp]
    L.ceos_mlp_num_params.restype = C.c_longlong
    L.ceos_mlp_get_params.argtypes = [vp, vp]
    L.ceos_mlp_set_params.argtypes = [vp, vp]
    L.ceos_mlp_gradients.argtypes = [vp, vp, vp, i, vp, C.POINTER(f)]
    L.ceos_mlp_train_step.argtypes = [vp, vp, vp, i, f, C.POINTER(f)]
    L.ceos_mlp_gradients_weighted.argtypes = [vp, vp, vp, vp, i, vp, C.POINTER(f)]
    L.ceos_mlp_train_step_weighted.argtypes = [vp, vp, vp, vp, i, f, C.POINTER(f)]
    L.ceos_mlp_set_training_precision.argtypes = [vp, i]
    L.ceos_mlp_forward.argtypes = [vp, vp, i, vp, vp, vp]
    L.ceos_rve_solve.argtypes = [vp, _dp, _dp, vp, vp]
    L.ceos_rve_solve_device.argtypes = [vp, vp, vp, vp, vp, vp]
    L.ceos_rve_get_fluctuation.argtypes = [vp, i, _dp]
    L.ceos_set_external_gp.argtypes = [vp, vp]
    L.ceos_gp_deformation_gradient.argtypes = [vp, _dp, _dp]
    L.ceos_gp_deformation_gradient_device.argtypes = [vp, vp, vp, vp]
    L.ceos_rve_solve_subset.argtypes = [vp, i, _ip, _dp, _dp, vp, vp]
    L.ceos_rve_to_external_device.argtypes = [i, vp, vp, vp, vp, vp, vp]
    L.ceos_mlp_shape.argtypes = [vp, vp, vp, vp, vp]
    L.ceos_surrogate_create.argtypes = [vp, C.POINTER(SurrogateCfg), i, C.POINTER(vp)]
    L.ceos_surrogate_destroy.argtypes = [vp]
    L.ceos_surrogate_destroy.restype = None
    L.ceos_surrogate_features_device.argtypes = [vp, i, vp, vp, vp, vp, vp]
    L.ceos_surrogate_eval_device.argtypes = [vp, i, vp, vp, vp, vp, vp, vp]
    L.ceos_bc_set.argtypes = [vp, i, vp, i, vp]
    L.ceos_residual_bc.argtypes = [vp, i, vp, vp, vp, i, C.POINTER(d)]
    L.ceos_residual_bc_device.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, i, vp, vp]
    L.ceos_set_solver_precision.argtypes = [vp, i]
    L.ceos_set_linear_solver.argtypes = [vp, i]
    L.ceos_linear_solve.argtypes = [vp, _dp, _dp, d, i, C.POINTER(i), C.POINTER(d), C.POINTER(i)]
    L.ceos_linear_solve_device.argtypes = [vp, vp, vp, vp, d, i, C.POINTER(i), C.POINTER(d), C.POINTER(i), vp]
    L.ceos_newton_update_device.argtypes = [vp, vp, vp, d, vp]
    L.ceos_periodic_create.argtypes = [vp, vp, _ip, C.POINTER(vp)]
    L.ceos_periodic_destroy.argtypes = [vp]
    L.ceos_periodic_destroy.restype = None
    L.ceos_periodic_info.argtypes = [vp, C.POINTER(i), C.POINTER(C.c_longlong), C.POINTER(d)]
    L.ceos_periodic_reduce_device.argtypes = [vp, vp, vp, vp, vp, i, vp]
    L.ceos_periodic_expand_device.argtypes = [vp, vp, d, vp, vp]
    L.ceos_periodic_homogenize_device.argtypes = [vp, vp, vp, vp]
    L.ceos_spmv_device.argtypes = [vp, vp, vp, vp, vp]
    _lib = L
    return L


def _check(code, what):
    if code != OK:
        raise CeosError(code, what)


class CeosMesh:
    """One hex8 mesh on the GPU. X (nn,3), conn (ne,8) 1-based node ids, mat_id (ne,)."""

    def __init__(self, X, conn1, mat_id=None, nl_analysis=True, precision=K_FP64, fused=True,
                 assembly=ASSEMBLY_GATHER, elem_tech=ELEM_FULL_INTEGRATION):
        L = load_library()
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        conn1 = np.ascontiguousarray(conn1, dtype=np.int32).reshape(-1, 8)
        self.nn, self.ne = X.shape[0], conn1.shape[0]
        self.ndof = 3 * self.nn
        if mat_id is None:
            mat_id = np.ones(self.ne, dtype=np.int32)
        mat_id = np.ascontiguousarray(mat_id, dtype=np.int32)
        h = C.c_void_p()
        _check(L.ceos_mesh_create(self.nn, X.reshape(-1), self.ne, 8, conn1.reshape(-1), mat_id,
                                  C.byref(h)), "ceos_mesh_create")
        self._h = h
        _check(L.ceos_set_analysis(self._h, 1 if nl_analysis else 0, elem_tech), "ceos_set_analysis")
        _check(L.ceos_set_precision(self._h, precision), "ceos_set_precision")
        _check(L.ceos_set_fused(self._h, 1 if fused else 0), "ceos_set_fused")
        _check(L.ceos_set_assembly(self._h, assembly), "ceos_set_assembly")
        self.precision = precision
        self._csr = None
        self.last_status = CeosStatus()

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            load_library().ceos_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_material(self, mat_id, model, props):
        p = np.ascontiguousarray(props, dtype=np.float64)
        _check(load_library().ceos_material_set(self._h, mat_id, model, p, len(p)), "ceos_material_set")

    def set_tangent_mode(self, numerical):
        _check(load_library().ceos_set_tangent_mode(self._h, 1 if numerical else 0), "ceos_set_tangent_mode")

    def timing_start(self, max_evaluations):
        _check(load_library().ceos_timing_start(self._h, max_evaluations), "ceos_timing_start")

    def timing_stop(self):
        """(element_ms, assembly_ms, n_evaluations, element_launches) summed over the
        evaluations timed since timing_start."""
        te, ta, n, nl = C.c_double(), C.c_double(), C.c_int(), C.c_longlong()
        _check(load_library().ceos_timing_stop(self._h, C.byref(te), C.byref(ta), C.byref(n), C.byref(nl)),
               "ceos_timing_stop")
        return te.value, ta.value, n.value, nl.value

    def set_assembly(self, mode):
        _check(load_library().ceos_set_assembly(self._h, mode), "ceos_set_assembly")

    def set_pipeline(self, chunks):
        _check(load_library().ceos_set_pipeline(self._h, chunks), "ceos_set_pipeline")

    def set_schedule(self, slabs):
        _check(load_library().ceos_set_schedule(self._h, slabs), "ceos_set_schedule")

    def info(self):
        nc, nnz = C.c_int(), C.c_longlong()
        _check(load_library().ceos_mesh_info(self._h, C.byref(nc), C.byref(nnz)), "ceos_mesh_info")
        return nc.value, nnz.value

    def assembly_info(self):
        """ceos_assembly_info of a prepared handle: patch and record counts of the gather assembly."""
        v = (C.c_longlong * 8)()
        _check(load_library().ceos_assembly_info(self._h, v), "ceos_assembly_info")
        keys = ("patches", "boundary_records", "boundary_diag_records", "complete_records",
                "complete_diag_records", "node_blocks", "gather_blocks", "gather_diag_blocks")
        return dict(zip(keys, (int(x) for x in v)))

    def csr(self, upper=True):
        if upper and self._csr is not None:
            return self._csr
        L = load_library()
        n, nnz = C.c_int(), C.c_longlong()
        _check(L.ceos_csr_pattern(self._h, 1 if upper else 0, C.byref(n), C.byref(nnz), None, None),
               "ceos_csr_pattern")
        rm = np.zeros(n.value + 1, dtype=np.int32)
        col = np.zeros(nnz.value, dtype=np.int32)
        _check(L.ceos_csr_pattern(self._h, 1 if upper else 0, C.byref(n), C.byref(nnz),
                                  rm.ctypes.data, col.ctypes.data), "ceos_csr_pattern")
        if upper:
            self._csr = (rm, col)
        return rm, col

    def solve_constitutive(self, time, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        return load_library().ceos_solve_constitutive(self._h, time, U, C.byref(self.last_status))

    def internal_force(self):
        F = np.zeros(self.ndof)
        st = load_library().ceos_internal_force(self._h, F, C.byref(self.last_status))
        return F, st

    def tangent(self):
        rm, col = self.csr()
        L = load_library()
        if self.precision == K_FP32_STORED:
            K = np.zeros(len(col), dtype=np.float32)
            _check(L.ceos_tangent_f32(self._h, K), "ceos_tangent_f32")
            return K.astype(np.float64)
        K = np.zeros(len(col))
        _check(L.ceos_tangent(self._h, K), "ceos_tangent")
        return K

    def update_mesh(self, U):
        _check(load_library().ceos_update_mesh(self._h, np.ascontiguousarray(U, dtype=np.float64)),
               "ceos_update_mesh")

    def switch_converged_state(self):
        _check(load_library().ceos_switch_converged_state(self._h), "ceos_switch_converged_state")

    def failed_elements(self):
        L = load_library()
        n = L.ceos_failed_elements(self._h, None, 0)
        ids = np.zeros(max(n, 1), dtype=np.int32)
        L.ceos_failed_elements(self._h, ids.ctypes.data, n)
        return ids[:n]

    def gp_stress(self):
        out = np.zeros(self.ne * 48)
        _check(load_library().ceos_get_gp_stress(self._h, out), "ceos_get_gp_stress")
        return out.reshape(self.ne, 8, 6)

    def state(self, which=1):
        out = np.zeros(self.ne * 8 * STATE_STRIDE)
        _check(load_library().ceos_get_state(self._h, which, out), "ceos_get_state")
        return out.reshape(self.ne, 8, STATE_STRIDE)

    def set_state(self, which, arr):
        _check(load_library().ceos_set_state(self._h, which,
                                             np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)),
               "ceos_set_state")

    def homogenize(self):
        P, F = np.zeros(9), np.zeros(9)
        _check(load_library().ceos_homogenize(self._h, P, F), "ceos_homogenize")
        return P, F

    def save_state(self, path):
        _check(load_library().ceos_state_save(self._h, path.encode()), "ceos_state_save")

    def load_state(self, path):
        _check(load_library().ceos_state_load(self._h, path.encode()), "ceos_state_load")

    # device-resident path (pointers are integers, e.g. torch.Tensor.data_ptr())
    def evaluate_device(self, time, dU, dFint, dK, stream=None):
        """ceos_evaluate_device on device pointers (ints); dK None = residual only (Fint without
        forming the tangent).  Asynchronous on `stream` (None = the handle's stream)."""
        _check(load_library().ceos_evaluate_device(self._h, time, dU, dFint, dK, stream),
               "ceos_evaluate_device")

    def prepare(self, records=True):
        """ceos_mesh_prepare: build the evaluation tables (and the element-record buffer), so the
        next evaluate_device allocates nothing and may be captured in a HIP graph."""
        _check(load_library().ceos_mesh_prepare(self._h, 1 if records else 0), "ceos_mesh_prepare")

    def evaluate_device_static(self, time, dU, dFint, dK, stream=None):
        """ceos_evaluate_device_static: evaluate_device with the element records in the library's
        static device buffer (no allocation in the call); the handle must be prepared."""
        _check(load_library().ceos_evaluate_device_static(self._h, time, dU, dFint, dK, stream),
               "ceos_evaluate_device_static")

    def set_external(self, dext):
        """Device pointer (int) of the [ne][8][27] external material points (MODEL_EXT)."""
        _check(load_library().ceos_set_external_gp(self._h, dext), "ceos_set_external_gp")

    def gp_deformation_gradient(self, U):
        """F at every Gauss point, (ne, 8, 3, 3), element order."""
        F = np.zeros(72 * self.ne)
        _check(load_library().ceos_gp_deformation_gradient(self._h, np.ascontiguousarray(U, dtype=np.float64), F),
               "ceos_gp_deformation_gradient")
        return F.reshape(self.ne, 8, 3, 3).transpose(0, 1, 3, 2)

    def gp_deformation_gradient_device(self, dU, dF, stream=None):
        _check(load_library().ceos_gp_deformation_gradient_device(self._h, dU, dF, stream),
               "ceos_gp_deformation_gradient_device")

    def update_mesh_device(self, dU, stream=None):
        _check(load_library().ceos_update_mesh_device(self._h, dU, stream), "ceos_update_mesh_device")

    def device_status(self):
        st = CeosStatus()
        code = load_library().ceos_device_status(self._h, C.byref(st))
        return code, st

    # Newton system: boundary conditions, residual norm, linear solve
    def set_bc(self, presc_dof, fixed_dof):
        """Constrained DOFs, 0-based (converted to the ABI's 1-based ids)."""
        p = np.ascontiguousarray(np.asarray(presc_dof, dtype=np.int64) + 1, dtype=np.int32)
        f = np.ascontiguousarray(np.asarray(fixed_dof, dtype=np.int64) + 1, dtype=np.int32)
        _check(load_library().ceos_bc_set(self._h, len(p), p.ctypes.data if len(p) else None,
                                          len(f), f.ctypes.data if len(f) else None), "ceos_bc_set")

    def residual_bc(self, it, fext=None, ubar=None, norm_type=2):
        """(R, norm) of Newton iteration `it` after solve_constitutive; K in the handle gets the BC."""
        R = np.zeros(self.ndof)
        nrm = C.c_double()
        fe = None if fext is None else np.ascontiguousarray(fext, dtype=np.float64)
        ub = None if ubar is None else np.ascontiguousarray(ubar, dtype=np.float64)
        _check(load_library().ceos_residual_bc(self._h, it, None if fe is None else fe.ctypes.data,
                                               None if ub is None else ub.ctypes.data, R.ctypes.data,
                                               norm_type, C.byref(nrm)), "ceos_residual_bc")
        return R, nrm.value

    def residual_bc_device(self, it, dU, dFint, dK, dFext, dUbar, dR, dnorm, norm_type=2, stream=None):
        _check(load_library().ceos_residual_bc_device(self._h, it, dU, dFint, dK, dFext, dUbar, dR, norm_type,
                                                      dnorm, stream), "ceos_residual_bc_device")

    def set_linear_solver(self, method):
        _check(load_library().ceos_set_linear_solver(self._h, method), "ceos_set_linear_solver")

    def set_solver_precision(self, mode):
        _check(load_library().ceos_set_solver_precision(self._h, mode), "ceos_set_solver_precision")

    def linear_solve(self, b, rtol=1e-10, maxit=20000):
        """x = K^-1 b with the handle's (BC-applied) tangent. Returns (x, iterations, relres, flag)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.ndof)
        it, rr, fl = C.c_int(), C.c_double(), C.c_int()
        _check(load_library().ceos_linear_solve(self._h, b, x, rtol, maxit, C.byref(it), C.byref(rr), C.byref(fl)),
               "ceos_linear_solve")
        return x, it.value, rr.value, fl.value

    def linear_solve_device(self, dK, db, dx, rtol=1e-10, maxit=20000, stream=None):
        it, rr, fl = C.c_int(), C.c_double(), C.c_int()
        _check(load_library().ceos_linear_solve_device(self._h, dK, db, dx, rtol, maxit, C.byref(it), C.byref(rr),
                                                       C.byref(fl), stream), "ceos_linear_solve_device")
        return it.value, rr.value, fl.value

    def set_preconditioner(self, kind, aggregate_size=0.0, coarse_reuse=0):
        """ceos_set_preconditioner: PREC_BLOCK_JACOBI or PREC_TWO_LEVEL (aggregate edge in mean
        element edges, <= 0: 8; a coarse inverse serves coarse_reuse further solves)."""
        _check(load_library().ceos_set_preconditioner(self._h, kind, float(aggregate_size), int(coarse_reuse)),
               "ceos_set_preconditioner")

    def preconditioner_info(self):
        na, nc = C.c_int(), C.c_int()
        _check(load_library().ceos_preconditioner_info(self._h, C.byref(na), C.byref(nc)),
               "ceos_preconditioner_info")
        return na.value, nc.value

    def coarse_space(self):
        """(aggregate of every node, P rows (nn, 3, 6), P^T K P (nc, nc)) of the last two-level solve."""
        na, nc = self.preconditioner_info()
        agg = np.zeros(self.nn, dtype=np.int32)
        B = np.zeros((self.nn, 3, 6))
        Ac = np.zeros((nc, nc))
        _check(load_library().ceos_coarse_space(self._h, agg.ctypes.data, B.ctypes.data, Ac.ctypes.data),
               "ceos_coarse_space")
        return agg, B, Ac

    def linear_solve_device_async(self, dK, db, dx, dresult, rtol=1e-10, maxit=1000, stream=None):
        """ceos_linear_solve_device_async: PCG of exactly maxit enqueued iterations (early exit on
        the device), no host synchronisation; dresult (3 doubles) receives {iterations, relative
        residual, flag}."""
        _check(load_library().ceos_linear_solve_device_async(self._h, dK, db, dx, rtol, maxit, dresult, stream),
               "ceos_linear_solve_device_async")

    def newton_update_device(self, dU, dDX, scale=1.0, stream=None):
        _check(load_library().ceos_newton_update_device(self._h, dU, dDX, scale, stream),
               "ceos_newton_update_device")

    def spmv_device(self, dK, dx, dy, stream=None):
        """ceos_spmv_device: dy = K dx for the symmetric tangent in the handle's upper CSR."""
        _check(load_library().ceos_spmv_device(self._h, dK, dx, dy, stream), "ceos_spmv_device")


class Periodic:
    """ceos_periodic_*: T^T K T, T^T R and T w between the handle of an RVE mesh and the handle
    of its periodic mesh (node_map: reduced node of every node, 0-based)."""

    def __init__(self, full, red, node_map):
        L = load_library()
        nm = np.ascontiguousarray(node_map, dtype=np.int32)
        if len(nm) != full.nn:
            raise ValueError("node_map needs one entry per node")
        h = C.c_void_p()
        _check(L.ceos_periodic_create(full._h, red._h, nm, C.byref(h)), "ceos_periodic_create")
        self._h = h
        self.full, self.red = full, red  # the handles must outlive this object

    def close(self):
        if self._h:
            load_library().ceos_periodic_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n, c, v = C.c_int(), C.c_longlong(), C.c_double()
        _check(load_library().ceos_periodic_info(self._h, C.byref(n), C.byref(c), C.byref(v)), "ceos_periodic_info")
        return n.value, c.value, v.value

    def reduce_device(self, dK, dF, dKred, dRred, mask_constrained=False, stream=None):
        _check(load_library().ceos_periodic_reduce_device(self._h, dK, dF, dKred, dRred, 1 if mask_constrained else 0,
                                                          stream), "ceos_periodic_reduce_device")

    def expand_device(self, dw, scale, dU, stream=None):
        _check(load_library().ceos_periodic_expand_device(self._h, dw, float(scale), dU, stream),
               "ceos_periodic_expand_device")

    def homogenize_device(self, dF, dP, stream=None):
        _check(load_library().ceos_periodic_homogenize_device(self._h, dF, dP, stream),
               "ceos_periodic_homogenize_device")


class RveBatch:
    """Batch of periodic voxel RVEs solved on the GPU (ceos_rve_*).

    sides (n_rve,) even in 4..16; phase: uint8 voxel phases, RVEs concatenated (x fastest);
    props: (2, nprops) constants of the matrix (phase 0) and inclusion (phase 1) for `model`.
    solve(Fbar) takes (n_rve, 3, 3) macro F and returns (P (n_rve, 3, 3), A (n_rve, 9, 9) or
    None, info (n_rve, 4)); A[r, 3J+i, 3L+k] = dP_iJ/dF_kL.  Raises CeosError if an RVE did not
    converge (info[:, 0] = -1 marks it)."""

    def __init__(self, sides, phase, model, props):
        L = load_library()
        self.sides = np.ascontiguousarray(sides, dtype=np.int32)
        self.n = len(self.sides)
        ph = np.ascontiguousarray(phase, dtype=np.uint8)
        if ph.size != int(np.sum(self.sides.astype(np.int64) ** 3)):
            raise ValueError("phase has %d voxels, sides need %d" % (ph.size, np.sum(self.sides.astype(np.int64) ** 3)))
        pr = np.ascontiguousarray(props, dtype=np.float64).reshape(2, -1)
        h = C.c_void_p()
        _check(L.ceos_rve_batch_create(self.n, self.sides, ph.ctypes.data, int(model), pr.reshape(-1),
                                       pr.shape[1], C.byref(h)), "ceos_rve_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().ceos_rve_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solver(self, newton_rtol=1e-10, newton_atol=1e-13, newton_maxit=12, pcg_rtol=1e-10,
                   pcg_maxit=5000, max_cutbacks=20):
        _check(load_library().ceos_rve_set_solver(self._h, newton_rtol, newton_atol, newton_maxit, pcg_rtol,
                                                  pcg_maxit, max_cutbacks), "ceos_rve_set_solver")

    def set_inexact(self, on):
        _check(load_library().ceos_rve_set_inexact(self._h, 1 if on else 0), "ceos_rve_set_inexact")

    def reset(self):
        _check(load_library().ceos_rve_reset(self._h), "ceos_rve_reset")

    def set_precision(self, mode):
        _check(load_library().ceos_rve_set_precision(self._h, mode), "ceos_rve_set_precision")

    def autotune(self, max_rves=256):
        """Choose the CG sweep tile by timing the condensation of a sample; returns (tile, seconds
        per candidate tile 2, 4, 8)."""
        g = C.c_int()
        t = np.zeros(3)
        _check(load_library().ceos_rve_autotune(self._h, max_rves, C.byref(g), t), "ceos_rve_autotune")
        return g.value, t

    def set_cg_group(self, g):
        _check(load_library().ceos_rve_set_cg_group(self._h, g), "ceos_rve_set_cg_group")

    def set_tangent_rtol(self, rtol):
        _check(load_library().ceos_rve_set_tangent_rtol(self._h, rtol), "ceos_rve_set_tangent_rtol")

    def set_warm_start(self, on):
        """Condensation solves from the RVE's last condensation solutions (on) or from zero."""
        _check(load_library().ceos_rve_set_warm_start(self._h, 1 if on else 0), "ceos_rve_set_warm_start")

    def set_coarse(self, on):
        """Two-level CG preconditioner (aggregate translations) on or off (block Jacobi alone)."""
        _check(load_library().ceos_rve_set_coarse(self._h, 1 if on else 0), "ceos_rve_set_coarse")

    def solve(self, Fbar, tangent=False, check=True):
        F = np.asarray(Fbar, dtype=np.float64).reshape(self.n, 3, 3)
        Fc = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(self.n, 9))  # column-major 9-vectors
        P = np.zeros((self.n, 9))
        A = np.zeros((self.n, 81)) if tangent else None
        info = np.zeros((self.n, 4), dtype=np.int32)
        code = load_library().ceos_rve_solve(self._h, Fc.reshape(-1), P.reshape(-1),
                                             A.ctypes.data if tangent else None, info.ctypes.data)
        if check:
            _check(code, "ceos_rve_solve")
        Pm = P.reshape(self.n, 3, 3).transpose(0, 2, 1)
        return Pm, (A.reshape(self.n, 9, 9) if tangent else None), info

    def solve_subset(self, sel, Fbar, tangent=False, check=True):
        """Solve only the RVEs `sel` (distinct ids) at Fbar (len(sel), 3, 3); returns P, A, info for
        those RVEs in the order of sel."""
        sel = np.ascontiguousarray(sel, dtype=np.int32)
        m = len(sel)
        F = np.asarray(Fbar, dtype=np.float64).reshape(m, 3, 3)
        Fc = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(m, 9))
        P = np.zeros((max(m, 1), 9))
        A = np.zeros((max(m, 1), 81)) if tangent else None
        info = np.zeros((max(m, 1), 4), dtype=np.int32)
        code = load_library().ceos_rve_solve_subset(self._h, m, sel, Fc.reshape(-1) if m else np.zeros(9),
                                                    P.reshape(-1), A.ctypes.data if tangent else None,
                                                    info.ctypes.data)
        if check:
            _check(code, "ceos_rve_solve_subset")
        Pm = P[:m].reshape(m, 3, 3).transpose(0, 2, 1)
        return Pm, (A[:m].reshape(m, 9, 9) if tangent else None), info[:m]

    def solve_device(self, dF, dP, dA=None, dinfo=None, stream=None):
        """Device arrays (pointers): dF, dP 9 per RVE (column-major), dA 81 per RVE or None."""
        _check(load_library().ceos_rve_solve_device(self._h, dF, dP, dA, dinfo, stream), "ceos_rve_solve_device")

    def profile(self):
        """(n_rve, 2) seconds in the Newton assemblies and in their linear solves (preconditioner
        set-up included) since the last reset."""
        t = np.zeros((self.n, 2), dtype=np.int64)
        _check(load_library().ceos_rve_profile(self._h, t.ctypes.data), "ceos_rve_profile")
        return t / 1.0e8

    def state_save(self, path):
        """ceos_rve_state_save: F_conv, fluctuations and the last P, dP/dF of every RVE to a file."""
        _check(load_library().ceos_rve_state_save(self._h, os.fsencode(path)), "ceos_rve_state_save")

    def state_load(self, path):
        """ceos_rve_state_load: the state of ceos_rve_state_save, into a batch of the same RVEs."""
        _check(load_library().ceos_rve_state_load(self._h, os.fsencode(path)), "ceos_rve_state_load")

    def fluctuation(self, r):
        n = int(self.sides[r])
        w = np.zeros(3 * n ** 3)
        _check(load_library().ceos_rve_get_fluctuation(self._h, r, w), "ceos_rve_get_fluctuation")
        return w


class Mlp:
    """MLP n_in -> width x3 (SiLU) -> n_out on the GPU (ceos_mlp_*).  Arrays passed to the
    device methods are hiprt.DeviceArray (float32); params() / set_params() use numpy."""

    def __init__(self, n_in, width, n_out, max_batch, seed=1):
        L = load_library()
        h = C.c_void_p()
        _check(L.ceos_mlp_create(n_in, width, n_out, max_batch, seed, C.byref(h)), "ceos_mlp_create")
        self._h = h
        self.n_in, self.width, self.n_out, self.max_batch = n_in, width, n_out, max_batch
        self.nparams = L.ceos_mlp_num_params(h)

    def close(self):
        if getattr(self, "_h", None):
            load_library().ceos_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_training_fp32(self, on):
        _check(load_library().ceos_mlp_set_training_precision(self._h, 1 if on else 0),
               "ceos_mlp_set_training_precision")

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(load_library().ceos_mlp_get_params(self._h, p.ctypes.data), "ceos_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check(load_library().ceos_mlp_set_params(self._h, p.ctypes.data), "ceos_mlp_set_params")

    def gradients(self, dX, dT, batch, dW=None):
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float()
        _check(load_library().ceos_mlp_gradients_weighted(self._h, dX.ptr, dT.ptr, dW.ptr if dW is not None else None,
                                                          batch, g.ctypes.data, C.byref(loss)),
               "ceos_mlp_gradients_weighted")
        return g, loss.value

    def train_step(self, dX, dT, batch, lr, want_loss=False, dW=None):
        loss = C.c_float()
        _check(load_library().ceos_mlp_train_step_weighted(self._h, dX.ptr, dT.ptr,
                                                           dW.ptr if dW is not None else None, batch, lr,
                                                           C.byref(loss) if want_loss else None),
               "ceos_mlp_train_step_weighted")
        return loss.value if want_loss else None

    def forward(self, dX, n, dY, dJ=None, stream=None):
        _check(load_library().ceos_mlp_forward(self._h, dX.ptr, n, dY.ptr, dJ.ptr if dJ is not None else None,
                                               stream), "ceos_mlp_forward")


def rve_to_external_device(n, dF, dP, dA, dext, dsel=None, stream=None):
    """S = F^-1 P and dS/dE from dP/dF for n points into ext rows (dsel[k] or k); dA None writes
    S only."""
    _check(load_library().ceos_rve_to_external_device(n, dsel, dF, dP, dA, dext, stream),
           "ceos_rve_to_external_device")


class Surrogate:
    """ceos_surrogate over an Mlp(27, width, 6): macro F + RVE descriptor -> S, dS/dE, hull flag.
    cfg: dict with u_scale, lam_lo, lam_hi, e_ref, d_mean, d_std, d_lo, d_hi (21 each), t_mean,
    t_std (6 each)."""

    def __init__(self, mlp, cfg, max_points):
        c = SurrogateCfg()
        c.u_scale, c.lam_lo, c.lam_hi, c.e_ref = (float(cfg[k]) for k in ("u_scale", "lam_lo", "lam_hi", "e_ref"))
        for k, n in (("d_mean", 21), ("d_std", 21), ("d_lo", 21), ("d_hi", 21), ("t_mean", 6), ("t_std", 6)):
            v = np.asarray(cfg[k], dtype=np.float64).reshape(-1)
            assert v.size == n, k
            getattr(c, k)[:] = v.tolist()
        self.cfg = c
        self.mlp = mlp  # kept alive: the surrogate uses it
        h = C.c_void_p()
        _check(load_library().ceos_surrogate_create(mlp._h, C.byref(c), max_points, C.byref(h)),
               "ceos_surrogate_create")
        self._h = h
        self.max_points = max_points

    def close(self):
        if getattr(self, "_h", None):
            load_library().ceos_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def features_device(self, n, dF, dClin, dX, dflags=None, stream=None):
        _check(load_library().ceos_surrogate_features_device(self._h, n, dF, dClin, dX, dflags, stream),
               "ceos_surrogate_features_device")

    def eval_device(self, n, dF, dClin, dext, dflags=None, dP=None, stream=None):
        _check(load_library().ceos_surrogate_eval_device(self._h, n, dF, dClin, dext, dflags, dP, stream),
               "ceos_surrogate_eval_device")
