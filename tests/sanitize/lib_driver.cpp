// Host-side checker of libceos_gpu built with AddressSanitizer and UndefinedBehaviorSanitizer on its
// host code (tests/sanitize/build_lib_asan.sh): drives the C ABI through the host paths that build
// tables (CSR pattern, patch and gather maps, colours, launch schedules, BC tables, coarse spaces,
// periodic contribution lists, RVE batches) and the host-pointer entry points, on small meshes with
// scrambled node numbering.  The values are checked only for being finite and for the calls'
// status codes; parity is the other tests' business.  Prints "lib driver: ok" at the end.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "ceos_gpu.h"

#define CHECK(call)                                                                  \
    do {                                                                             \
        const int rc_ = (call);                                                      \
        if (rc_ != CEOS_OK) {                                                        \
            std::fprintf(stderr, "%s:%d: %s returned %d\n", __FILE__, __LINE__, #call, rc_); \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

static void require(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        std::exit(1);
    }
}

static bool finite(const std::vector<double>& v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

// structured nx x ny x nz hex8 box with perturbed interior nodes; node ids scrambled by perm
struct Box {
    int nn = 0, ne = 0;
    std::vector<double> X;
    std::vector<int> conn;  // 1-based
    std::vector<int> mat;
};

static Box make_box(int nx, int ny, int nz, double perturb, bool scramble, unsigned seed) {
    Box b;
    const int mx = nx + 1, my = ny + 1, mz = nz + 1;
    b.nn = mx * my * mz;
    b.ne = nx * ny * nz;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(-perturb, perturb);
    std::vector<int> perm(b.nn);
    std::iota(perm.begin(), perm.end(), 0);
    if (scramble) std::shuffle(perm.begin(), perm.end(), rng);
    b.X.assign(3 * (size_t)b.nn, 0.0);
    for (int k = 0; k < mz; ++k)
        for (int j = 0; j < my; ++j)
            for (int i = 0; i < mx; ++i) {
                const int g = i + mx * (j + my * k);
                const bool inner = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
                const double d[3] = {inner ? u(rng) : 0.0, inner ? u(rng) : 0.0, inner ? u(rng) : 0.0};
                b.X[3 * perm[g] + 0] = (i + d[0]) / nx;
                b.X[3 * perm[g] + 1] = (j + d[1]) / ny;
                b.X[3 * perm[g] + 2] = (k + d[2]) / nz;
            }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int c[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
                for (int a = 0; a < 8; ++a) {
                    const int g = (i + c[a][0]) + mx * ((j + c[a][1]) + my * (k + c[a][2]));
                    b.conn.push_back(perm[g] + 1);
                }
                b.mat.push_back(((i + j + k) % 3 == 0) ? 2 : 1);
            }
    return b;
}

static void mesh_paths(const std::string& dir) {
    Box b = make_box(5, 4, 3, 0.15, true, 7u);
    ceos_handle* h = nullptr;
    CHECK(ceos_mesh_create(b.nn, b.X.data(), b.ne, 8, b.conn.data(), b.mat.data(), &h));
    const double cnh1[2] = {1.0, 100.0 - 2.0 / 3.0}, cnh2[2] = {5.0, 500.0 - 10.0 / 3.0};
    CHECK(ceos_material_set(h, 1, CEOS_MODEL_COMPRESSIBLE_NEO_HOOKEAN, cnh1, 2));
    CHECK(ceos_material_set(h, 2, CEOS_MODEL_COMPRESSIBLE_NEO_HOOKEAN, cnh2, 2));
    // CSR patterns, two-call protocol
    for (int upper = 0; upper <= 1; ++upper) {
        int n = 0;
        long long nnz = 0;
        CHECK(ceos_csr_pattern(h, upper, &n, &nnz, nullptr, nullptr));
        std::vector<int> rowmap(n + 1), col(nnz);
        CHECK(ceos_csr_pattern(h, upper, &n, &nnz, rowmap.data(), col.data()));
        require(rowmap[n] - rowmap[0] == nnz, "CSR row map");
    }
    int n = 0;
    long long nnz = 0;
    CHECK(ceos_csr_pattern(h, 1, &n, &nnz, nullptr, nullptr));
    const int ndof = 3 * b.nn;
    std::mt19937 rng(11u);
    std::normal_distribution<double> nd(0.0, 0.001);  // strains of a few 1e-3: K stays positive definite
    std::vector<double> U(ndof), Fint(ndof), K(nnz), sig(6 * 8 * (size_t)b.ne);
    for (double& x : U) x = nd(rng);
    ceos_status st;
    // every assembly mode, launch schedule and pipeline depth rebuilds its tables
    const int modes[][3] = {{CEOS_ASSEMBLY_GATHER, 0, 0}, {CEOS_ASSEMBLY_GATHER, 2, 0}, {CEOS_ASSEMBLY_GATHER, 0, 3},
                            {CEOS_ASSEMBLY_COLORED, 0, 0}, {CEOS_ASSEMBLY_GATHER, 0, 0}};
    for (const auto& m : modes) {
        CHECK(ceos_set_assembly(h, m[0]));
        CHECK(ceos_set_schedule(h, m[1]));
        CHECK(ceos_set_pipeline(h, m[2]));
        CHECK(ceos_update_mesh(h, U.data()));
        CHECK(ceos_solve_constitutive(h, 0.0, U.data(), &st));
        require(st.error == 0, "constitutive status");
        CHECK(ceos_internal_force(h, Fint.data(), &st));
        CHECK(ceos_tangent(h, K.data()));
        require(finite(Fint) && finite(K), "finite Fint and K");
    }
    long long info[8];
    CHECK(ceos_assembly_info(h, info));
    int ncol = 0;
    long long nnz2 = 0;
    CHECK(ceos_mesh_info(h, &ncol, &nnz2));
    require(nnz2 == nnz, "mesh info nnz");
    // numerical tangent, FP32 storage, Gauss-point stress and state
    CHECK(ceos_set_tangent_mode(h, CEOS_TANGENT_NUMERICAL));
    CHECK(ceos_solve_constitutive(h, 0.0, U.data(), &st));
    CHECK(ceos_tangent(h, K.data()));
    CHECK(ceos_set_tangent_mode(h, CEOS_TANGENT_ANALYTIC));
    CHECK(ceos_set_precision(h, CEOS_K_FP32_STORED));
    std::vector<float> Kf(nnz);
    CHECK(ceos_solve_constitutive(h, 0.0, U.data(), &st));
    CHECK(ceos_tangent_f32(h, Kf.data()));
    CHECK(ceos_set_precision(h, CEOS_K_FP64));
    CHECK(ceos_solve_constitutive(h, 0.0, U.data(), &st));
    CHECK(ceos_internal_force(h, Fint.data(), &st));
    CHECK(ceos_tangent(h, K.data()));
    CHECK(ceos_get_gp_stress(h, sig.data()));
    std::vector<double> state(8 * (size_t)b.ne * CEOS_STATE_STRIDE);
    CHECK(ceos_get_state(h, 1, state.data()));
    CHECK(ceos_set_state(h, 0, state.data()));
    CHECK(ceos_switch_converged_state(h));
    double P[9], F[9];
    CHECK(ceos_homogenize(h, P, F));
    const std::string path = dir + "/lib_driver_state.bin";
    CHECK(ceos_state_save(h, path.c_str()));
    CHECK(ceos_state_load(h, path.c_str()));
    std::remove(path.c_str());
    std::vector<double> Fgp(9 * 8 * (size_t)b.ne);
    CHECK(ceos_gp_deformation_gradient(h, U.data(), Fgp.data()));
    // boundary conditions, residual, linear solves with every preconditioner and method, at the
    // reference state (K positive definite: a random state of this nearly incompressible material
    // can make K indefinite through its geometric part, and PCG then hands over to MINRES)
    std::vector<double> U0(ndof, 0.0);
    CHECK(ceos_update_mesh(h, U0.data()));
    CHECK(ceos_solve_constitutive(h, 0.0, U0.data(), &st));
    std::vector<int> presc, fixed;  // 1-based DOF numbers
    for (int A = 0; A < b.nn; ++A) {
        if (b.X[3 * A] < 1e-12) {
            fixed.push_back(3 * A + 1);
            fixed.push_back(3 * A + 2);
            fixed.push_back(3 * A + 3);
        } else if (b.X[3 * A] > 1.0 - 1e-12) {
            presc.push_back(3 * A + 1);
        }
    }
    CHECK(ceos_bc_set(h, (int)presc.size(), presc.data(), (int)fixed.size(), fixed.data()));
    std::vector<double> Ubar(ndof, 0.0), R(ndof), x(ndof);
    for (int d : presc) Ubar[d - 1] = 0.01;
    double normR = 0.0;
    CHECK(ceos_residual_bc(h, 0, nullptr, Ubar.data(), R.data(), 2, &normR));
    require(std::isfinite(normR) && normR > 0.0, "residual norm");
    for (double& r : R) r = -r;
    const int precs[][2] = {{CEOS_PREC_BLOCK_JACOBI, 0}, {CEOS_PREC_TWO_LEVEL, 0}, {CEOS_PREC_TWO_LEVEL, 2}};
    for (const auto& pr : precs) {
        CHECK(ceos_set_preconditioner(h, pr[0], 2.0, pr[1]));
        for (int method : {CEOS_SOLVER_AUTO, CEOS_SOLVER_PCG, CEOS_SOLVER_MINRES}) {
            CHECK(ceos_set_linear_solver(h, method));
            int its = 0, flag = 0;
            double rr = 0.0;
            CHECK(ceos_linear_solve(h, R.data(), x.data(), 1e-8, 5000, &its, &rr, &flag));
            std::printf("linear solve: preconditioner %d reuse %d method %d: %d iterations, relres %.2e, flag %d\n",
                        pr[0], pr[1], method, its, rr, flag);
            require(finite(x), "finite solution");
        }
    }
    int n_agg = 0, n_coarse = 0;
    CHECK(ceos_set_preconditioner(h, CEOS_PREC_TWO_LEVEL, 2.0, 0));
    CHECK(ceos_set_linear_solver(h, CEOS_SOLVER_PCG));
    {
        int its = 0, flag = 0;
        double rr = 0.0;
        CHECK(ceos_linear_solve(h, R.data(), x.data(), 1e-8, 5000, &its, &rr, &flag));
    }
    CHECK(ceos_preconditioner_info(h, &n_agg, &n_coarse));
    require(n_agg > 0 && n_coarse > 0, "coarse space built");
    std::vector<int> agg(b.nn);
    std::vector<double> basis(18 * (size_t)b.nn), Ac((size_t)n_coarse * n_coarse);
    CHECK(ceos_coarse_space(h, agg.data(), basis.data(), Ac.data()));
    // an inverted element: the status reports it and the failed ids come back
    std::vector<double> Ubad(U);
    {
        // node 7 of element 0 mirrored through its node 1: the element turns inside out
        const int A1 = b.conn[0] - 1, A7 = b.conn[6] - 1;
        for (int d = 0; d < 3; ++d) Ubad[3 * A7 + d] = 2.0 * (b.X[3 * A1 + d] - b.X[3 * A7 + d]);
    }
    CHECK(ceos_update_mesh(h, Ubad.data()));
    const int rc_bad = ceos_solve_constitutive(h, 0.0, Ubad.data(), &st);
    std::vector<int> ids(b.ne);
    const int nf = ceos_failed_elements(h, ids.data(), b.ne);
    std::printf("inverted element: status %d, error %d, %d failed elements\n", rc_bad, st.error, nf);
    require(nf >= 1, "the inverted element is reported");
    ceos_mesh_destroy(h);
}

// n^3 voxel cube and its torus: the periodic contribution lists of ceos_periodic_create
static void periodic_paths() {
    const int nc = 3;
    Box full = make_box(nc, nc, nc, 0.0, false, 3u);
    // torus: node (i,j,k) of the cube -> (i mod n, j mod n, k mod n)
    const int m = nc + 1;
    std::vector<int> node_map(full.nn);
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) node_map[i + m * (j + m * k)] = (i % nc) + nc * ((j % nc) + nc * (k % nc));
    const int nr = nc * nc * nc;
    std::vector<double> Xr(3 * (size_t)nr);
    for (int A = 0; A < full.nn; ++A)
        for (int d = 0; d < 3; ++d) Xr[3 * node_map[A] + d] = full.X[3 * A + d];
    std::vector<int> conn_r(full.conn.size());
    for (size_t q = 0; q < full.conn.size(); ++q) conn_r[q] = node_map[full.conn[q] - 1] + 1;
    ceos_handle *hf = nullptr, *hr = nullptr;
    CHECK(ceos_mesh_create(full.nn, full.X.data(), full.ne, 8, full.conn.data(), full.mat.data(), &hf));
    CHECK(ceos_mesh_create(nr, Xr.data(), full.ne, 8, conn_r.data(), full.mat.data(), &hr));
    const double nh[2] = {0.5, 50.0};
    for (ceos_handle* h : {hf, hr}) {
        CHECK(ceos_material_set(h, 1, CEOS_MODEL_NEO_HOOKEAN, nh, 2));
        CHECK(ceos_material_set(h, 2, CEOS_MODEL_NEO_HOOKEAN, nh, 2));
    }
    CHECK(ceos_bc_set(hf, 0, nullptr, 0, nullptr));
    const int fixed[3] = {1, 2, 3};  // node 0, 1-based DOFs
    CHECK(ceos_bc_set(hr, 0, nullptr, 3, fixed));
    ceos_periodic* per = nullptr;
    CHECK(ceos_periodic_create(hf, hr, node_map.data(), &per));
    int nn_red = 0;
    long long contributions = 0;
    double volume = 0.0;
    CHECK(ceos_periodic_info(per, &nn_red, &contributions, &volume));
    require(nn_red == nr && contributions > 0 && std::fabs(volume - 1.0) < 1e-12, "periodic info");
    ceos_periodic_destroy(per);
    ceos_mesh_destroy(hr);
    ceos_mesh_destroy(hf);
}

// a small RVE batch: create, solve with tangents, subset solve, tile probe, state file, destroy
static void rve_paths(const std::string& dir) {
    const int side[3] = {6, 4, 8};
    size_t tot = 0;
    for (int s : side) tot += (size_t)s * s * s;
    std::vector<unsigned char> phase(tot);
    for (size_t q = 0; q < tot; ++q) phase[q] = (q % 7 == 0) ? 1 : 0;
    const double props[4] = {0.5, 50.0, 5.0, 500.0};
    ceos_rve_batch* b = nullptr;
    CHECK(ceos_rve_batch_create(3, side, phase.data(), CEOS_MODEL_NEO_HOOKEAN, props, 2, &b));
    std::vector<double> Fb(27), P(27), A(243);
    std::vector<int> info(12);
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 9; ++q) Fb[9 * r + q] = (q % 4 == 0 ? 1.0 : 0.0) + 0.01 * (r + 1) * ((q * 7) % 5 - 2) / 2.0;
    for (int coarse = 1; coarse >= 0; --coarse) {
        CHECK(ceos_rve_reset(b));
        CHECK(ceos_rve_set_coarse(b, coarse));
        CHECK(ceos_rve_solve(b, Fb.data(), P.data(), A.data(), info.data()));
        for (int r = 0; r < 3; ++r) require(info[4 * r] == 0, "RVE converged");
    }
    const int sel[1] = {1};
    double Ps[9], As[81];
    int infs[4];
    CHECK(ceos_rve_solve_subset(b, 1, sel, Fb.data() + 9, Ps, As, infs));
    int chosen = 0;
    double secs[3];
    CHECK(ceos_rve_autotune(b, 2, &chosen, secs));
    const std::string path = dir + "/lib_driver_rve.bin";
    CHECK(ceos_rve_state_save(b, path.c_str()));
    CHECK(ceos_rve_state_load(b, path.c_str()));
    std::remove(path.c_str());
    std::vector<double> w(3 * 64);
    CHECK(ceos_rve_get_fluctuation(b, 1, w.data()));
    std::vector<long long> ticks(6);
    CHECK(ceos_rve_profile(b, ticks.data()));
    ceos_rve_batch_destroy(b);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    mesh_paths(dir);
    std::printf("mesh paths done\n");
    std::fflush(stdout);
    periodic_paths();
    std::printf("periodic paths done\n");
    std::fflush(stdout);
    rve_paths(dir);
    std::printf("lib driver: ok\n");
    return 0;
}
