// Golden-vector harness over the reference's METIS-free sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// unmodified reference headers and translation units (core/gmres.cpp,
// core/fgmres.cpp, core/bicgstab.h, precond/ilu0.h, precond/jacobi.h,
// precond/coarse_solve.h, mesh/poisson.h, io/metrics_logger.h,
// io/matrix_market.h, analysis/*.h, core/sparse_matrix.h).  The binary goes
// to oracle/_ref/ and is driven by tests/golden/make_golden.py, which writes the
// committed fixtures.  Nothing under the reference's decomp/ (needs METIS) is
// included: RAS/ASM goldens are produced from the reference's own
// extract_submatrix and ILU0Precond on subdomain index sets that the caller
// supplies.
#include "core/sparse_matrix.h"
#include "core/vector.h"
#include "core/gmres.h"
#include "core/fgmres.h"
#include "core/bicgstab.h"
#include "precond/ilu0.h"
#include "precond/jacobi.h"
#include "precond/ilut.h"
#include "precond/coarse_solve.h"
#include "mesh/poisson.h"
#include "io/metrics_logger.h"
#include "io/matrix_market.h"
#include "analysis/condition_estimator.h"
#include "analysis/spectral_analysis.h"
#include "analysis/convergence_analysis.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

using namespace schwarz;

namespace {

// SplitMix64 (Steele, Lea, Flood 2014); u = (x >> 11) * 2^-53.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void print_vec(const char* key, const std::vector<double>& v, bool last = false) {
    std::printf("\"%s\":[", key);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? "," : "", v[i]);
    std::printf("]%s", last ? "" : ",");
}
void print_ivec(const char* key, const std::vector<int>& v, bool last = false) {
    std::printf("\"%s\":[", key);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s%d", i ? "," : "", v[i]);
    std::printf("]%s", last ? "" : ",");
}

CSRMatrix grid(const std::string& spec) {
    // "2d:NX:NY", "3d:NX:NY:NZ" or "rnd:N:SEED" (N x N, diagonal plus about
    // three random off-diagonals per row, values uniform in [-1, 1) and the
    // diagonal 4 + |u|; columns sorted)
    int a = 0, b = 0, c = 0;
    if (spec.rfind("rnd:", 0) == 0) {
        unsigned long long seed = 0;
        std::sscanf(spec.c_str() + 4, "%d:%llu", &a, &seed);
        SplitMix64 rng(seed);
        CSRMatrix A;
        A.nrows = A.ncols = a;
        A.row_ptr.assign(a + 1, 0);
        for (int i = 0; i < a; ++i) {
            std::vector<int> cols{i};
            for (int t = 0; t < 3; ++t) cols.push_back((int)(rng.next() % (uint64_t)a));
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            for (int j : cols) {
                A.col_idx.push_back(j);
                const double u = 2.0 * rng.uniform01() - 1.0;
                A.vals.push_back(j == i ? 4.0 + std::abs(u) : u);
            }
            A.row_ptr[i + 1] = (int)A.col_idx.size();
        }
        return A;
    }
    if (spec.rfind("2d:", 0) == 0) {
        std::sscanf(spec.c_str() + 3, "%d:%d", &a, &b);
        return make_poisson_2d(a, b);
    }
    std::sscanf(spec.c_str() + 3, "%d:%d:%d", &a, &b, &c);
    return make_poisson_3d(a, b, c);
}

int cmd_ilu0(const std::string& spec) {
    CSRMatrix A = grid(spec);
    ILU0Precond p;
    p.setup(A);
    std::vector<double> x(A.nrows), y(A.nrows);
    SplitMix64 rng(7);
    for (auto& t : x) t = 2.0 * rng.uniform01() - 1.0;
    p.apply(x.data(), y.data(), A.nrows);
    std::printf("{");
    print_vec("lu", p.lu_vals());
    print_ivec("diag", p.diag_ptr());
    print_vec("x", x);
    print_vec("y", y, true);
    std::printf("}\n");
    return 0;
}

// ilut GRID DROP FILL: l_nnz, u_nnz and one apply on a SplitMix x.
int cmd_ilut(const std::string& spec, double drop, int fill) {
    CSRMatrix A = grid(spec);
    ILUTPrecond p(drop, fill);
    p.setup(A);
    std::vector<double> x(A.nrows), y(A.nrows);
    SplitMix64 rng(11);
    for (auto& t : x) t = 2.0 * rng.uniform01() - 1.0;
    p.apply(x.data(), y.data(), A.nrows);
    std::printf("{\"l_nnz\":%d,\"u_nnz\":%d,", p.l_nnz(), p.u_nnz());
    print_vec("x", x);
    print_vec("y", y, true);
    std::printf("}\n");
    return 0;
}

int cmd_solve(int argc, char** argv) {
    // solve METHOD PC GRID RHS RESTART MAXITER TOL  (PC "ilut" uses ILUTPrecond())
    if (argc < 9) return 2;
    std::string method = argv[2], pc = argv[3], rhs = argv[5];
    CSRMatrix A = grid(argv[4]);
    GMRESParams prm;
    prm.restart = std::atoi(argv[6]);
    prm.max_iter = std::atoi(argv[7]);
    prm.tol = std::atof(argv[8]);
    int n = A.nrows;
    std::vector<double> xs(n, 1.0), b(n), x(n, 0.0);
    if (rhs.rfind("uniform:", 0) == 0) {
        SplitMix64 rng(std::strtoull(rhs.c_str() + 8, nullptr, 10));
        for (auto& t : xs) t = 2.0 * rng.uniform01() - 1.0;
    }
    A.spmv(xs, b);
    std::unique_ptr<Preconditioner> M;
    if (pc == "jacobi") M.reset(new JacobiPrecond());
    else if (pc == "ilu0") M.reset(new ILU0Precond());
    else if (pc == "ilut") M.reset(new ILUTPrecond());
    if (M) M->setup(A);
    // A negative TOL means relative: |TOL| * ||M^-1 b|| for GMRES (the quantity
    // its absolute test measures at x0 = 0) and |TOL| * ||b|| for FGMRES.
    if (prm.tol < 0) {
        std::vector<double> mb(n);
        if (M && method != "fgmres") M->apply(b.data(), mb.data(), n);
        else mb = b;
        prm.tol = -prm.tol * norm2(mb);
    }
    GMRESResult r = method == "fgmres"     ? fgmres(A, b, x, M.get(), prm)
                    : method == "bicgstab" ? bicgstab(A, b, x, M.get(), prm)
                                           : gmres(A, b, x, M.get(), prm);
    std::vector<double> ax;
    A.spmv(x, ax);
    double rn = 0, bn = 0, en = 0, xn = 0;
    for (int i = 0; i < n; ++i) {
        rn += (b[i] - ax[i]) * (b[i] - ax[i]);
        bn += b[i] * b[i];
        en += (x[i] - xs[i]) * (x[i] - xs[i]);
        xn += xs[i] * xs[i];
    }
    std::printf("{\"iterations\":%d,\"residual_norm\":%.17g,\"converged\":%d,\"tol\":%.17g,",
                r.iterations, r.residual_norm, r.converged ? 1 : 0, prm.tol);
    std::printf("\"rel_true_residual\":%.17g,\"x_rel_error\":%.17g,", std::sqrt(rn / bn),
                std::sqrt(en / xn));
    if (n <= 4096) print_vec("x", x);
    print_vec("hist", r.residual_history, true);
    std::printf("}\n");
    return 0;
}

// Reads from stdin: GRID nsub; per subdomain: m, then m sorted global dofs,
// then m owned flags (0/1); then n input values.  Writes the RAS and ASM
// outputs computed with the reference extract_submatrix + ILU0Precond, with
// gather / restricted scatter / additive scatter as src/precond/ras.h:45-70 and
// src/precond/asm.h:35-62 do it (ascending subdomain order).
int cmd_blocks() {
    char spec[64];
    int nsub;
    if (std::scanf("%63s %d", spec, &nsub) != 2) return 2;
    CSRMatrix A = grid(spec);
    int n = A.nrows;
    std::vector<std::vector<int>> dofs(nsub), owned(nsub);
    for (int s = 0; s < nsub; ++s) {
        int m;
        if (std::scanf("%d", &m) != 1) return 2;
        dofs[s].resize(m);
        owned[s].resize(m);
        for (int i = 0; i < m; ++i) if (std::scanf("%d", &dofs[s][i]) != 1) return 2;
        for (int i = 0; i < m; ++i) if (std::scanf("%d", &owned[s][i]) != 1) return 2;
    }
    std::vector<double> x(n), yr(n, 0.0), ya(n, 0.0);
    for (int i = 0; i < n; ++i) if (std::scanf("%lf", &x[i]) != 1) return 2;
    for (int s = 0; s < nsub; ++s) {
        CSRMatrix Al = A.extract_submatrix(dofs[s]);
        ILU0Precond p;
        p.setup(Al);
        int m = (int)dofs[s].size();
        std::vector<double> rl(m), zl(m);
        for (int i = 0; i < m; ++i) rl[i] = x[dofs[s][i]];
        p.apply(rl.data(), zl.data(), m);
        for (int i = 0; i < m; ++i) {
            if (owned[s][i]) yr[dofs[s][i]] = zl[i];
            ya[dofs[s][i]] += zl[i];
        }
    }
    std::printf("{");
    print_vec("ras", yr);
    print_vec("asm", ya, true);
    std::printf("}\n");
    return 0;
}

int cmd_lu(int n, uint64_t seed, int singular) {
    SplitMix64 rng(seed);
    std::vector<double> A((size_t)n * n);
    for (auto& t : A) t = 2.0 * rng.uniform01() - 1.0;
    if (singular) for (int j = 0; j < n; ++j) A[(size_t)(n - 1) * n + j] = A[j];
    std::vector<double> b(n);
    for (auto& t : b) t = 2.0 * rng.uniform01() - 1.0;
    CoarseSolver cs;
    try {
        cs.setup(A, n);
    } catch (const std::exception& e) {
        std::printf("{\"error\":\"%s\"}\n", e.what());
        return 0;
    }
    std::vector<double> x = cs.solve(b);
    std::printf("{");
    print_vec("A", A);
    print_vec("b", b);
    print_vec("x", x, true);
    std::printf("}\n");
    return 0;
}

// analyze PC GRID K POWER_ITERS: estimate_condition on A and on M^{-1}A and
// estimate_spectral_radius, as src/main.cpp:211-220 calls them.
int cmd_analyze(int argc, char** argv) {
    if (argc < 6) return 2;
    std::string pc = argv[2];
    CSRMatrix A = grid(argv[3]);
    const int k = std::atoi(argv[4]), pit = std::atoi(argv[5]);
    std::unique_ptr<Preconditioner> M;
    if (pc == "jacobi") M.reset(new JacobiPrecond());
    else if (pc == "ilu0") M.reset(new ILU0Precond());
    if (M) M->setup(A);
    auto pe = [](const char* key, const ConditionEstimate& e) {
        std::printf("\"%s\":{\"lambda_min\":%.17g,\"lambda_max\":%.17g,\"kappa\":%.17g,"
                    "\"steps\":%d,\"valid\":%d},",
                    key, e.lambda_min, e.lambda_max, e.kappa, e.lanczos_steps, e.valid ? 1 : 0);
    };
    std::printf("{");
    pe("cond_A", estimate_condition(A, nullptr, k));
    pe("cond_MA", estimate_condition(A, M.get(), k));
    SpectralResult sr = estimate_spectral_radius(A, M.get(), pit);
    std::printf("\"rho\":%.17g,\"power_iters\":%d,\"ok\":%d}\n", sr.spectral_radius,
                sr.power_iters, sr.convergence_ok ? 1 : 0);
    return 0;
}

// conv OUT CONVERGED: residual history on stdin (count, then values);
// analyze_convergence + write_jsonl(OUT, "m.mtx", "ras", 8, 2), then
// print_summary on stdout.
int cmd_conv(int argc, char** argv) {
    if (argc < 4) return 2;
    int cnt;
    if (std::scanf("%d", &cnt) != 1) return 2;
    std::vector<double> h(cnt);
    for (auto& t : h) if (std::scanf("%lf", &t) != 1) return 2;
    ConvergenceMetrics m = analyze_convergence(h, std::atoi(argv[3]) != 0);
    m.write_jsonl(argv[2], "m.mtx", "ras", 8, 2);
    m.print_summary();
    return 0;
}

// mmtime FILE: read_matrix_market wall time (s), sizes and a checksum only.
int cmd_mmtime(const char* path) {
    auto t0 = std::chrono::steady_clock::now();
    CSRMatrix A = read_matrix_market(path);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double sum = 0.0;
    for (double v : A.vals) sum += v;
    std::printf("{\"seconds\":%.6f,\"nrows\":%d,\"nnz\":%d,\"sum\":%.17g}\n", s, A.nrows,
                A.nnz(), sum);
    return 0;
}

// mm FILE: read_matrix_market, printed as JSON (or the exception message).
int cmd_mm(const char* path) {
    CSRMatrix A;
    try {
        A = read_matrix_market(path);
    } catch (const std::exception& e) {
        std::printf("{\"error\":\"%s\"}\n", e.what());
        return 0;
    }
    std::printf("{\"nrows\":%d,\"ncols\":%d,", A.nrows, A.ncols);
    print_ivec("row_ptr", A.row_ptr);
    print_ivec("col_idx", A.col_idx);
    print_vec("vals", A.vals, true);
    std::printf("}\n");
    return 0;
}

int cmd_metrics() {
    Metrics m;
    m.matrix = "poisson2d_256";
    m.method = "GMRES+ilu0";
    m.n = 65536;
    m.nnz = 326656;
    m.threads = 8;
    m.overlap = 1;
    m.iterations = 412;
    m.tol = 1e-8;
    m.residual_norm = 9.87654321e-9;
    m.time_setup_s = 0.0123456789;
    m.time_solve_s = 1.5;
    std::printf("%s\n", m.to_json().c_str());
    Metrics d;  // defaults
    std::printf("%s\n", d.to_json().c_str());
    Metrics g;
    g.matrix = "poisson3d_128";
    g.method = "GPU_GMRES+ras";
    g.n = 2097152;
    g.nnz = 14581760;
    g.gpus = 1;
    g.restart = 50;
    g.tol = 0.000123;
    g.residual_norm = 0.0;
    g.time_solve_s = 123456.5;
    g.speedup = 3.25;
    g.efficiency = 1e-7;
    std::printf("%s\n", g.to_json().c_str());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    std::string c = argv[1];
    if (c == "ilu0" && argc >= 3) return cmd_ilu0(argv[2]);
    if (c == "solve") return cmd_solve(argc, argv);
    if (c == "blocks") return cmd_blocks();
    if (c == "lu" && argc >= 5) return cmd_lu(std::atoi(argv[2]), std::strtoull(argv[3], nullptr, 10), std::atoi(argv[4]));
    if (c == "metrics") return cmd_metrics();
    if (c == "analyze") return cmd_analyze(argc, argv);
    if (c == "ilut" && argc >= 5) return cmd_ilut(argv[2], std::atof(argv[3]), std::atoi(argv[4]));
    if (c == "conv") return cmd_conv(argc, argv);
    if (c == "mm" && argc >= 3) return cmd_mm(argv[2]);
    if (c == "mmtime" && argc >= 3) return cmd_mmtime(argv[2]);
    if (c == "mmw" && argc >= 4) {  // mmw GRID OUT: write_matrix_market; the matrix on stdout
        CSRMatrix A = grid(argv[2]);
        write_matrix_market(argv[3], A);
        std::printf("{\"nrows\":%d,\"ncols\":%d,", A.nrows, A.ncols);
        print_ivec("row_ptr", A.row_ptr);
        print_ivec("col_idx", A.col_idx);
        print_vec("vals", A.vals, true);
        std::printf("}\n");
        return 0;
    }
    return 2;
}
