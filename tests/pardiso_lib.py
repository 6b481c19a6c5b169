"""MKL PARDISO through ctypes, set up as the reference's linear solver of the Newton step
(test infrastructure and CPU baseline only; the library never calls it).

PardisoConstructor / Get_iparm / PardisoSolve / CallPardiso
(Tools/MathTools/LinearSolver/ModPardisoSolver.f90:84-262, 343-420): mtype -2 (real symmetric
indefinite, upper CSR, 1-based), maxfct = mnum = nrhs = 1, iparm(1) = 1, iparm(2) = 3 (parallel
nested dissection), iparm(4) = 62, iparm(18) = -1, iparm(24) = 1 (two-level factorization), the
rest 0; every solve is phase 13 (analysis, factorization, solve) on a fresh handle followed by
phase -1 (release), as PardisoSolve does.  Error -4 (zero pivot) is accepted like the reference.

MKL comes from the image's conda tree (libmkl_rt with the Intel OpenMP runtime), LP64.
"""
import ctypes as C
import os

import numpy as np

MKL_DIR = "/opt/conda/lib"
_lib = None


def available():
    return os.path.exists(os.path.join(MKL_DIR, "libmkl_rt.so.1")) or os.path.exists(os.path.join(MKL_DIR, "libmkl_rt.so"))


def load():
    global _lib
    if _lib is not None:
        return _lib
    for dep in ("libiomp5.so",):
        p = os.path.join(MKL_DIR, dep)
        if os.path.exists(p):
            C.CDLL(p, mode=C.RTLD_GLOBAL)
    name = "libmkl_rt.so.1" if os.path.exists(os.path.join(MKL_DIR, "libmkl_rt.so.1")) else "libmkl_rt.so"
    L = C.CDLL(os.path.join(MKL_DIR, name), mode=C.RTLD_GLOBAL)
    ip = C.POINTER(C.c_int)
    L.pardiso.argtypes = [C.c_void_p, ip, ip, ip, ip, ip, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, ip, ip,
                          ip, C.c_void_p, C.c_void_p, ip]
    L.pardiso.restype = None
    L.MKL_Set_Num_Threads.argtypes = [C.c_int]
    L.MKL_Get_Max_Threads.restype = C.c_int
    _lib = L
    return L


def set_threads(n):
    load().MKL_Set_Num_Threads(int(n))


def iparm_reference():
    ip = np.zeros(64, dtype=np.int32)
    ip[0] = 1     # iparm(1): no solver defaults
    ip[1] = 3     # iparm(2): parallel nested dissection
    ip[3] = 62    # iparm(4)
    ip[17] = -1   # iparm(18): report nonzeros of the factors
    ip[23] = 1    # iparm(24): two-level factorization
    return ip


def solve(rowmap, col, val, b):
    """x = K^-1 b, K symmetric given by its upper CSR (1-based rowmap/col, FP64 values).
    Returns (x, error, nnz of the factors)."""
    L = load()
    n = len(rowmap) - 1
    ia = np.ascontiguousarray(rowmap, dtype=np.int32)
    ja = np.ascontiguousarray(col, dtype=np.int32)
    a = np.ascontiguousarray(val, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64).copy()
    x = np.zeros(n)
    pt = np.zeros(64, dtype=np.int64)
    iparm = iparm_reference()
    perm = np.zeros(n, dtype=np.int32)
    one, mtype, msglvl = C.c_int(1), C.c_int(-2), C.c_int(0)
    nn = C.c_int(n)
    err = C.c_int(0)
    ipp = iparm.ctypes.data_as(C.POINTER(C.c_int))

    def call(phase):
        ph = C.c_int(phase)
        L.pardiso(pt.ctypes.data, C.byref(one), C.byref(one), C.byref(mtype), C.byref(ph), C.byref(nn),
                  a.ctypes.data, ia.ctypes.data, ja.ctypes.data, perm.ctypes.data, C.byref(one), ipp,
                  C.byref(msglvl), bb.ctypes.data, x.ctypes.data, C.byref(err))
        return err.value

    e = call(13)
    nnz_factors = int(iparm[17])
    call(-1)
    return x, e, nnz_factors
