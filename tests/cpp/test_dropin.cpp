// Drop-in API checks, re-expressing the reference's own test programs
// (tests/test_gmres.cpp, tests/test_schwarz.cpp, tests/test_two_level.cpp) against
// include/schwarz/schwarz.h.  Every solve runs on the GPU.  Prints one line per
// check and returns the number of failures.
#include "schwarz/schwarz.h"

#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

using namespace schwarz;

static int fails = 0, checks = 0;
#define CHECK(cond, msg)                                                          \
    do {                                                                          \
        ++checks;                                                                 \
        if (cond) std::printf("PASS: %s\n", msg);                                 \
        else {                                                                    \
            ++fails;                                                              \
            std::printf("FAIL: %s (line %d)\n", msg, __LINE__);                   \
        }                                                                         \
    } while (0)

static double l1(const std::vector<double>& a, const std::vector<double>& b) {
    double e = 0;
    for (size_t i = 0; i < a.size(); ++i) e += std::fabs(a[i] - b[i]);
    return e;
}

static void gmres_identity() {  // tests/test_gmres.cpp:27-44
    CSRMatrix I = make_identity(10);
    std::vector<double> b(10), x(10, 0.0);
    for (int i = 0; i < 10; ++i) b[i] = i + 1.0;
    GMRESResult r = gmres(I, b, x, nullptr);
    CHECK(r.converged, "identity converges");
    CHECK(r.iterations <= 2, "identity within 2 iterations");
    CHECK(l1(x, b) < 1e-12, "identity solution x = b");
}

static void gmres_spd3() {  // tests/test_gmres.cpp:46-92
    CSRMatrix A;
    A.nrows = A.ncols = 3;
    A.row_ptr = {0, 2, 5, 7};
    A.col_idx = {0, 1, 0, 1, 2, 1, 2};
    A.vals = {4, 1, 1, 3, 1, 1, 2};
    std::vector<double> xs = {1, 2, 3}, b;
    A.spmv(xs, b);
    {
        std::vector<double> x(3, 0.0);
        gmres(A, b, x, nullptr);
        CHECK(l1(x, xs) < 1e-8, "small SPD, no preconditioner");
    }
    {
        JacobiPrecond J;
        J.setup(A);
        std::vector<double> x(3, 0.0);
        gmres(A, b, x, &J);
        CHECK(l1(x, xs) < 1e-8, "small SPD, Jacobi");
    }
    {
        ILU0Precond P;
        P.setup(A);
        std::vector<double> x(3, 0.0);
        gmres(A, b, x, &P);
        CHECK(l1(x, xs) < 1e-8, "small SPD, ILU0");
    }
}

static void gmres_poisson5() {  // tests/test_gmres.cpp:94-142
    CSRMatrix A = make_poisson_2d(5, 5);
    std::vector<double> ones(25, 1.0), b, x(25, 0.0);
    A.spmv(ones, b);
    ILU0Precond P;
    P.setup(A);
    GMRESParams p;
    p.restart = 25;
    p.tol = 1e-10;
    GMRESResult r = gmres(A, b, x, &P, p);
    CHECK(r.converged, "Poisson 5x5 converges");
    CHECK(r.residual_norm < 1e-10, "Poisson 5x5 residual < 1e-10");
    CHECK(l1(x, ones) < 1e-6, "Poisson 5x5 solution");
}

static GMRESResult solve_with(Preconditioner& M, int nx, int restart, int max_iter,
                              bool flexible = false) {
    CSRMatrix A = make_poisson_2d(nx, nx);
    std::vector<double> ones(A.nrows, 1.0), b, x(A.nrows, 0.0);
    A.spmv(ones, b);
    M.setup(A);
    GMRESParams p;
    p.restart = restart;
    p.tol = 1e-10;
    p.max_iter = max_iter;
    return flexible ? fgmres(A, b, x, &M, p) : gmres(A, b, x, &M, p);
}

static void schwarz_tests() {  // tests/test_schwarz.cpp:26-149
    ASMPrecond a(4, 1);
    GMRESResult ra = solve_with(a, 20, 30, 500);
    CHECK(ra.converged && ra.residual_norm < 1e-9, "ASM(4,1) on 20x20 converges");
    RASPrecond r(4, 1);
    GMRESResult rr = solve_with(r, 20, 30, 500);
    CHECK(rr.converged && rr.residual_norm < 1e-9, "RAS(4,1) on 20x20 converges");
    ASMPrecond a8(8, 2);
    RASPrecond r8(8, 2);
    GMRESResult x1 = solve_with(a8, 30, 30, 1000), x2 = solve_with(r8, 30, 30, 1000);
    CHECK(x2.iterations <= x1.iterations + 5, "RAS(8,2) iterations <= ASM(8,2) + 5");
    RASPrecond r0(8, 0), r2(8, 2);
    GMRESResult o0 = solve_with(r0, 30, 30, 2000), o2 = solve_with(r2, 30, 30, 2000);
    CHECK(o2.iterations <= o0.iterations, "overlap 2 needs no more iterations than overlap 0");
}

static void two_level_tests() {  // tests/test_two_level.cpp:27-143
    TwoLevelSchwarzPrecond ta(8, 1, FineLevelType::ASM), tr(8, 1, FineLevelType::RAS);
    GMRESResult a = solve_with(ta, 30, 30, 1000), r = solve_with(tr, 30, 30, 1000);
    CHECK(a.converged, "two-level ASM(8,1) converges");
    CHECK(r.converged, "two-level RAS(8,1) converges");
    TwoLevelSchwarzPrecond t16(16, 1);
    RASPrecond r16(16, 1);
    GMRESResult b2 = solve_with(t16, 40, 30, 2000), b1 = solve_with(r16, 40, 30, 2000);
    CHECK(b2.iterations <= b1.iterations + 5, "two-level <= one-level + 5 (40x40, 16 parts)");
    for (int np : {4, 8, 16}) {
        TwoLevelSchwarzPrecond t(np, 1);
        GMRESResult q = solve_with(t, 30, 30, 1000, true);
        CHECK(q.converged, ("FGMRES two-level nparts=" + std::to_string(np)).c_str());
    }
}

static void extensions() {
    // a host-side user preconditioner through the generic Preconditioner interface
    struct Scaled : Preconditioner {
        void setup(const CSRMatrix&) override {}
        void apply(const double* x, double* y, int n) const override {
            for (int i = 0; i < n; ++i) y[i] = 0.25 * x[i];
        }
    } s;
    GMRESResult r = solve_with(s, 12, 30, 1000);
    CHECK(r.converged, "user-defined host preconditioner");
    // fgmres_flex with a FlexApply
    CSRMatrix A = make_poisson_2d(16, 16);
    ILU0Precond P;
    P.setup(A);
    std::vector<double> ones(A.nrows, 1.0), b, x(A.nrows, 0.0);
    A.spmv(ones, b);
    int calls = 0;
    FlexApply f = [&](const double* in, double* out, int n) {
        ++calls;
        P.apply(in, out, n);
    };
    GMRESParams p;
    GMRESResult q = fgmres_flex(A, b, x, f, p);
    CHECK(q.converged && calls == q.iterations, "fgmres_flex calls the functor once per step");
    // rtol mode + CGS2 + FP32 two-level
    TwoLevelSchwarzPrecond t(8, 1);
    t.set_precision(SG_FP32);
    t.setup(A);
    std::vector<double> x2(A.nrows, 0.0);
    GMRESParams p2;
    p2.rtol = 1e-8;
    p2.ortho = SG_ORTHO_CGS2;
    GMRESResult m = fgmres(A, b, x2, &t, p2);
    CHECK(m.converged && m.rel_true_residual <= 1e-8, "mixed FGMRES reaches an FP64-verified 1e-8");
    // errors keep the reference's messages
    bool threw = false;
    try {
        CSRMatrix Z = make_identity(8);
        for (int i = 0; i < 4; ++i) Z.vals[i] = 0.0;
        TwoLevelSchwarzPrecond z(2, 0);
        z.set_partition({0, 0, 0, 0, 1, 1, 1, 1});
        z.setup(Z);
    } catch (const std::exception& e) {
        threw = std::string(e.what()).find("Coarse matrix is singular") != std::string::npos;
    }
    CHECK(threw, "singular coarse matrix throws the reference's message");
    Metrics mt;
    mt.matrix = "poisson2d_16";
    mt.method = "GMRES+ilu0";
    std::string js = mt.to_json();
    CHECK(js.rfind("{\"matrix\":\"poisson2d_16\",\"method\":\"GMRES+ilu0\",\"n\":0", 0) == 0,
          "Metrics::to_json field order");
}

static void next_rows() {
    // BiCGSTAB (src/core/bicgstab.h): relative tolerance, right preconditioning
    CSRMatrix A = make_poisson_2d(30, 30);
    std::vector<double> ones(A.nrows, 1.0), b;
    A.spmv(ones, b);
    ILU0Precond ilu;
    ilu.setup(A);
    std::vector<double> x(A.nrows, 0.0);
    GMRESResult r = bicgstab(A, b, x, &ilu);
    CHECK(r.converged && std::abs(r.iterations - 23) <= 1,
          "BiCGSTAB + ILU0 on 30x30: the reference's 23 iterations");
    CHECK(l1(x, ones) < 1e-6 * A.nrows, "BiCGSTAB solution");
    CHECK((int)r.residual_history.size() == r.iterations, "one history entry per iteration");
    // Matrix Market round trip (src/io/matrix_market.h)
    const std::string path = "/tmp/sgamd_dropin_" + std::to_string(::getpid()) + ".mtx";
    write_matrix_market(path, A, 17);
    CSRMatrix B = read_matrix_market(path);
    std::remove(path.c_str());
    CHECK(B.nrows == A.nrows && B.row_ptr == A.row_ptr && B.col_idx == A.col_idx &&
              B.vals == A.vals,
          "Matrix Market write/read round trip");
    bool threw = false;
    try {
        read_matrix_market(path);
    } catch (const std::runtime_error& e) {
        threw = std::string(e.what()) == "Cannot open file: " + path;
    }
    CHECK(threw, "missing file throws the reference's message");
    // analysis (src/analysis/*.h): Poisson 2D 12x12 closed-form extremes
    CSRMatrix P = make_poisson_2d(12, 12);
    ConditionEstimate ce = estimate_condition(P, nullptr, 80);
    const double lo = 8.0 * std::pow(std::sin(M_PI / 26.0), 2);
    CHECK(ce.valid && std::abs(ce.lambda_min - lo) < 1e-10 && std::abs(ce.lambda_max - (8 - lo)) < 1e-9,
          "Lanczos extremes of Poisson 12x12");
    JacobiPrecond jac;
    jac.setup(P);
    SpectralResult sr = estimate_spectral_radius(P, &jac, 300);
    CHECK(sr.power_iters > 0 && sr.spectral_radius < 1.0 && (int)sr.eigenvector.size() == P.nrows,
          "power iteration with Jacobi");
    ConvergenceMetrics cm = analyze_convergence(r.residual_history, r.converged);
    CHECK(cm.iterations == r.iterations && cm.converged && cm.stagnation_start == -1,
          "convergence metrics of the BiCGSTAB run");
}

int main() {
    next_rows();
    gmres_identity();
    gmres_spd3();
    gmres_poisson5();
    schwarz_tests();
    two_level_tests();
    extensions();
    std::printf("%d/%d checks passed\n", checks - fails, checks);
    return fails;
}
