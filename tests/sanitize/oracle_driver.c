/* Drives every entry point of the CPU oracle (oracle/ceos_oracle.c) on small perturbed hex8 cubes,
 * for every material model, full and numerical tangents, state commits and failures, so that a
 * build with -fsanitize=address,undefined checks the oracle's memory accesses and arithmetic.
 * Test infrastructure: built and run by tests/test_sanitized_oracle.py.  Exit status 0 = clean. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/ceos_oracle.h"

static unsigned long long rng_state = 20261204ULL;
static double urand(void) {  /* splitmix64 */
    unsigned long long z = (rng_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;
}

static void cube(int n, double **X, int **conn, int *nn, int *ne) {
    const int m = n + 1;
    *nn = m * m * m;
    *ne = n * n * n;
    *X = malloc(sizeof(double) * 3 * (size_t)(*nn));
    *conn = malloc(sizeof(int) * 8 * (size_t)(*ne));
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                const int id = i + m * (j + m * k);
                const int inner = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                (*X)[3 * id] = (double)i / n + (inner ? 0.1 * (urand() - 0.5) / n : 0.0);
                (*X)[3 * id + 1] = (double)j / n + (inner ? 0.1 * (urand() - 0.5) / n : 0.0);
                (*X)[3 * id + 2] = (double)k / n + (inner ? 0.1 * (urand() - 0.5) / n : 0.0);
            }
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int *c = *conn + 8 * (i + n * (j + n * k));
                const int b = 1 + i + m * (j + m * k);
                c[0] = b; c[1] = b + 1; c[2] = b + 1 + m; c[3] = b + m;
                c[4] = b + m * m; c[5] = b + 1 + m * m; c[6] = b + 1 + m + m * m; c[7] = b + m + m * m;
            }
}

static int run_model(int model, const double *props, int nprops, int nl, int numt) {
    double *X;
    int *conn, nn, ne;
    cube(3, &X, &conn, &nn, &ne);
    int *mat = malloc(sizeof(int) * ne);
    for (int e = 0; e < ne; ++e) mat[e] = 1 + (e % 3 == 1);
    or_mesh *m = or_mesh_create(nn, X, ne, conn, mat);
    int rc = 0;
    if (!m) return 1;
    rc |= or_material_set(m, 1, model, props, nprops);
    rc |= or_material_set(m, 2, model, props, nprops);
    or_set_analysis(m, nl);
    or_set_tangent_mode(m, numt);
    or_set_threads(m, 2);
    const int ndof = or_ndof(m);
    const long nnz = or_csr_nnz(m);
    int *rowmap = malloc(sizeof(int) * (ndof + 1));
    int *col = malloc(sizeof(int) * nnz);
    rc |= or_csr_pattern(m, rowmap, col);
    double *U = calloc(ndof, sizeof(double)), *F = malloc(sizeof(double) * ndof);
    double *K = malloc(sizeof(double) * nnz);
    double *buf = malloc(sizeof(double) * (size_t)ne * 8 * OR_STATE_STRIDE);
    for (int step = 1; step <= 3; ++step) {
        for (int d = 0; d < ndof; ++d) U[d] = 0.002 * step * X[d] + 0.0005 * (urand() - 0.5);
        if (nl) or_update_mesh(m, U);
        rc |= or_solve_constitutive(m, 0.1 * step, U);
        rc |= or_internal_force(m, F);
        rc |= or_tangent(m, K);
        or_switch_converged_state(m);
    }
    or_get_gp_stress(m, buf);
    or_get_gp_F(m, buf);
    or_get_state(m, 0, buf);
    or_set_state(m, 1, buf);
    double Ke[576], Fe[24], V, VX, P[9], Fh[9];
    rc |= or_element_stiffness(m, 5, Ke);
    rc |= or_element_internal_force(m, 5, Fe);
    or_element_volume(m, 5, &V, &VX);
    rc |= or_homogenize(m, P, Fh);
    int ids[8];
    (void)or_failed_elements(m, ids, 8);
    or_mesh_destroy(m);
    free(rowmap); free(col); free(U); free(F); free(K); free(buf); free(mat); free(X); free(conn);
    return rc;
}

int main(void) {
    const double ghl[] = {210000.0, 0.3}, j2[] = {210000.0, 0.3, 250.0, 1000.0}, nh[] = {0.5, 100.0};
    const double svk[] = {100.0, 0.3}, cnh[] = {1.0, 99.0};
    const double visco[] = {100.0, 1.0, 99.0, 50.0, 0.5, 0.2};
    int rc = 0;
    rc |= run_model(OR_MODEL_GHL, ghl, 2, 0, 0);
    rc |= run_model(OR_MODEL_J2, j2, 4, 0, 0);
    rc |= run_model(OR_MODEL_NEOHOOKEAN, nh, 2, 1, 0);
    rc |= run_model(OR_MODEL_NEOHOOKEAN, nh, 2, 1, 1);
    rc |= run_model(OR_MODEL_NH_Q1P0, nh, 2, 1, 0);
    rc |= run_model(OR_MODEL_SVK, svk, 2, 1, 0);
    rc |= run_model(OR_MODEL_COMP_NEOHOOKEAN, cnh, 2, 1, 1);
    rc |= run_model(OR_MODEL_VISCO_MATRIX, visco, 6, 1, 0);
    /* a failing element: two swapped nodes give detJ < 0 */
    {
        double *X;
        int *conn, nn, ne;
        cube(2, &X, &conn, &nn, &ne);
        int t = conn[8 * 3]; conn[8 * 3] = conn[8 * 3 + 1]; conn[8 * 3 + 1] = t;
        or_mesh *m = or_mesh_create(nn, X, ne, conn, NULL);
        or_material_set(m, 1, OR_MODEL_COMP_NEOHOOKEAN, cnh, 2);
        double *U = calloc(3 * nn, sizeof(double));
        const int st = or_solve_constitutive(m, 0.0, U);
        int ids[4];
        const int nf = or_failed_elements(m, ids, 4);
        if (st == 0 || nf != 1 || ids[0] != 4) rc |= 2;
        or_mesh_destroy(m);
        free(U); free(X); free(conn);
    }
    printf("oracle driver: %s\n", rc ? "FAILED" : "ok");
    return rc ? 1 : 0;
}
