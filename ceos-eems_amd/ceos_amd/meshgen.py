"""Synthetic hex8 meshes and boundary conditions for the workloads the library is measured on.

The reference reads meshes from GiD / HyperMesh files; the workloads here are structured cubes
generated in memory with the same conventions (1-based node ids, hex8 node order
(---,+--,++-,-+-,--+,+-+,+++,-++) of FEM/Elements/ElementLibrary/ModElementHexa8.f90:150-200).

Workloads:
  cube(10)                          1,000 hex8, 1,331 nodes (unit cube, unperturbed)
  cube(80, perturb=0.1, seed=...)   512,000 hex8, 531,441 nodes, interior nodes moved
                                    uniformly within +-10% of the spacing (numpy PCG64 seeded)
  cube(64)                          262,144 hex8 (J2 cyclic case)
"""
import numpy as np

PERTURB_SEED = 20261203


def cube(n, length=1.0, perturb=0.0, seed=PERTURB_SEED):
    """Unit cube with n cells per side. Returns (X (nn,3) float64, conn (ne,8) int32 1-based)."""
    m = n + 1
    h = length / n
    g = np.arange(m, dtype=np.float64) * h
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    X = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    if perturb > 0.0:
        rng = np.random.Generator(np.random.PCG64(seed))
        d = rng.uniform(-perturb * h, perturb * h, size=X.shape)
        idx = np.arange(m)
        interior = (idx > 0) & (idx < n)
        mask = (interior[None, None, :] & interior[None, :, None] & interior[:, None, None]).ravel()
        X[mask] += d[mask]
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = k.ravel(), j.ravel(), i.ravel()  # x fastest

    def nid(a, b, c):
        return 1 + a + m * (b + m * c)

    conn = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                     nid(i, j + 1, k + 1)], axis=1).astype(np.int32)
    return X, conn


def stretch_bc(X, length=1.0, tol=1e-12):
    """x=0 face: all DOFs fixed; x=L face: ux prescribed. Returns (fixed_dofs0, presc_dofs0) as
    0-based DOF indices (DOF = 3*node + k)."""
    x0 = np.nonzero(np.abs(X[:, 0]) < tol)[0]
    x1 = np.nonzero(np.abs(X[:, 0] - length) < tol)[0]
    fixed = np.sort(np.concatenate([3 * x0, 3 * x0 + 1, 3 * x0 + 2]))
    presc = np.sort(3 * x1)
    return fixed.astype(np.int64), presc.astype(np.int64)


def csr_nnz_structured(n):
    """Closed-form nnz of the upper CSR of a structured n^3 hex8 cube (node-blocked 27-point)."""
    m = n + 1
    ndof = 3 * m ** 3
    full = 9 * (3 * m - 2) ** 3
    return (full - ndof) // 2 + ndof
