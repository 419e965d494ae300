// schwarz_solve -- command-line driver, drop-in for the reference's single-rank
// path (src/main.cpp:63-275): same flags (--matrix / --poisson, --precond,
// --restart, --tol, --maxiter, --fgmres, --bicgstab, --nparts, --overlap,
// --gpu, --output, --analyze), stdout lines and JSONL records, every solve and
// analysis on the GPU.  Extensions: --poisson3d N, --jump3d N (3D diffusion, N^3,
// coefficient jumps between 16^3 blocks),
// --rhs ones|uniform|normal with --seed, --rtol, --precision fp64|fp32,
// --ortho mgs|cgs2, --partition FILE, --ufactor U, --weighted-partition,
// --coarse-inverse, --batch K (K systems with right-hand-side seeds seed..,
// solved together), --learned WEIGHTS.npy (learned FGMRES stage on the 2D
// grid), --mm-sum-duplicates / --mm-strict (Matrix Market reader fixes).
// Exit code 0 if converged, else 1 (src/main.cpp:274).
#include "../../include/schwarz/schwarz.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <thread>

using namespace schwarz;

namespace {

void usage() {
    std::cerr
        << "Usage: schwarz_solve [options]\n"
           "\nMatrix source (one required):\n"
           "  --matrix <file.mtx>      Matrix Market file\n"
           "  --mm-sum-duplicates      merge duplicate entries of the file by summing\n"
           "  --mm-strict              header flags from the banner only; reject malformed lines\n"
           "  --poisson <N>            Poisson 2D, N x N grid (n = N^2)\n"
           "  --poisson3d <N>          Poisson 3D, N^3 grid\n"
           "  --jump3d <N>             3D finite-volume diffusion, kappa = 10^{-3,0,3} on 16^3 blocks\n"
           "\nPreconditioner:\n"
           "  --precond <type>         none|jacobi|ilu0|ilut|asm|ras|twolevel_asm|twolevel_ras\n"
           "  --drop-tol <t>           ILUT drop tolerance (default 1e-4)\n"
           "  --fill-factor <f>        ILUT fill factor (default 10)\n"
           "  --precision <p>          fp64 (default) | fp32 (factors and applies)\n"
           "\nSolver:\n"
           "  --restart <m>            restart length (default 30)\n"
           "  --tol <val>              absolute tolerance on the residual estimate (default 1e-10)\n"
           "  --rtol <val>             stop on an FP64-verified ||b-Ax||/||b|| <= val\n"
           "  --maxiter <n>            max iterations (default 1000)\n"
           "  --fgmres                 flexible GMRES (right preconditioning)\n"
           "  --bicgstab               BiCGSTAB (right preconditioning; --tol relative to ||r0||)\n"
           "  --ortho <o>              mgs (default) | cgs2\n"
           "\nDomain decomposition:\n"
           "  --nparts <n>             subdomains (default 4)\n"
           "  --overlap <n>            overlap layers (default 1)\n"
           "  --partition <file>       one subdomain id per row (text)\n"
           "  --ufactor <u>            allowed imbalance in 1/1000 (default 30)\n"
           "  --weighted-partition     weigh partition edges by |a_ij|\n"
           "  --coarse-inverse         two-level: apply the coarse solve as one mat-vec\n"
           "\nRight-hand side:\n"
           "  --rhs <r>                ones (b = A*1, default) | uniform (b = A*x*, x* ~ U[-1,1]) | normal (b ~ N(0,1))\n"
           "  --seed <s>               seed for --rhs uniform|normal (default 1)\n"
           "\nOutput:\n"
           "  --batch <K>              solve K systems together (right-hand sides from seeds seed..seed+K-1)\n"
           "  --learned <w.npy>        learned FGMRES stage on top of --precond (2D --poisson grids)\n"
           "  --learned-every <k>      run the network on every k-th application (default 1)\n"
           "  --learned-coarsen <c>    the network sees c x c block averages: 1, 2 or 4 (default 1)\n"
           "  --gpu                    accepted; every solve runs on the GPU\n"
           "  --analyze                Lanczos condition estimates, spectral radius and\n"
           "                           convergence metrics (<output>.conv.jsonl)\n"
           "  --output <file.jsonl>    append a metrics record\n";
}

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
};

std::vector<double> normals(uint64_t seed, int n) {
    SplitMix64 r(seed);
    std::vector<double> out(n);
    for (int i = 0; i < n; i += 2) {
        double u1 = ((double)(r.next() >> 11) + 1.0) * 0x1.0p-53;
        double u2 = r.u01();
        double rad = std::sqrt(-2.0 * std::log(u1));
        out[i] = rad * std::cos(2.0 * M_PI * u2);
        if (i + 1 < n) out[i + 1] = rad * std::sin(2.0 * M_PI * u2);
    }
    return out;
}

// 7-point finite-volume diffusion, harmonic-mean faces, Dirichlet boundary faces
// add the cell's own kappa; kappa = 10^u, u uniform on {-3, 0, 3} per 16^3 block.
CSRMatrix make_jump3d(int N, uint64_t seed) {
    const int B = 16, nb = (N + B - 1) / B;
    SplitMix64 r(seed);
    std::vector<double> kb((size_t)nb * nb * nb);
    for (auto& k : kb) k = std::pow(10.0, -3.0 + 3.0 * std::floor(3.0 * r.u01()));
    auto kap = [&](int x, int y, int z) { return kb[((size_t)(z / B) * nb + y / B) * nb + x / B]; };
    CSRMatrix A;
    const int nxy = N * N, n = nxy * N;
    A.nrows = A.ncols = n;
    A.row_ptr.assign(n + 1, 0);
    A.col_idx.reserve((size_t)7 * n);
    A.vals.reserve((size_t)7 * n);
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                const int i = z * nxy + y * N + x;
                const double k = kap(x, y, z);
                A.row_ptr[i] = (int)A.col_idx.size();
                double diag = 0.0;
                auto face = [&](bool in, int j, double kj) {
                    double f = in ? 2.0 * k * kj / (k + kj) : k;
                    diag += f;
                    return f;
                };
                const int nb6[6][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                double f[6];
                bool in[6];
                for (int d = 0; d < 6; ++d) {
                    int xx = x + nb6[d][0], yy = y + nb6[d][1], zz = z + nb6[d][2];
                    in[d] = xx >= 0 && xx < N && yy >= 0 && yy < N && zz >= 0 && zz < N;
                    f[d] = face(in[d], 0, in[d] ? kap(xx, yy, zz) : k);
                }
                const int off[6] = {-nxy, -N, -1, 1, N, nxy};
                for (int d = 0; d < 3; ++d)
                    if (in[d]) { A.col_idx.push_back(i + off[d]); A.vals.push_back(-f[d]); }
                A.col_idx.push_back(i);
                A.vals.push_back(diag);
                for (int d = 3; d < 6; ++d)
                    if (in[d]) { A.col_idx.push_back(i + off[d]); A.vals.push_back(-f[d]); }
            }
    A.row_ptr[n] = (int)A.col_idx.size();
    return A;
}

}  // namespace

int main(int argc, char** argv) {
    std::string precond = "ilu0", output, matrix, rhs = "ones", part_file, ortho = "mgs",
                precision = "fp64";
    GMRESParams prm;
    int overlap = 1, nparts = 4, poisson = 0, poisson3d = 0, jump3d = 0, ufactor = 0;
    bool use_fgmres = false, use_bicgstab = false, weighted = false, coarse_inv = false;
    bool analyze = false;
    double drop_tol = 1e-4;  // src/main.cpp:81-82
    int fill_factor = 10;
    MatrixMarketOptions mm;
    std::string matrix_file;
    uint64_t seed = 1;
    int batch = 0, learned_every = 1, learned_coarsen = 1;
    std::string learned;
    for (int i = 1; i < argc; ++i) {
        auto is = [&](const char* f) { return !std::strcmp(argv[i], f); };
        auto val = [&](const char* f) { return !std::strcmp(argv[i], f) && i + 1 < argc; };
        if (val("--precond")) precond = argv[++i];
        else if (val("--restart")) prm.restart = std::atoi(argv[++i]);
        else if (val("--tol")) prm.tol = std::atof(argv[++i]);
        else if (val("--rtol")) prm.rtol = std::atof(argv[++i]);
        else if (val("--maxiter")) prm.max_iter = std::atoi(argv[++i]);
        else if (val("--overlap")) overlap = std::atoi(argv[++i]);
        else if (val("--nparts")) nparts = std::atoi(argv[++i]);
        else if (val("--output")) output = argv[++i];
        else if (val("--poisson")) poisson = std::atoi(argv[++i]);
        else if (val("--poisson3d")) poisson3d = std::atoi(argv[++i]);
        else if (val("--jump3d")) jump3d = std::atoi(argv[++i]);
        else if (val("--rhs")) rhs = argv[++i];
        else if (val("--seed")) seed = std::strtoull(argv[++i], nullptr, 10);
        else if (val("--partition")) part_file = argv[++i];
        else if (val("--ufactor")) ufactor = std::atoi(argv[++i]);
        else if (val("--ortho")) ortho = argv[++i];
        else if (val("--precision")) precision = argv[++i];
        else if (val("--batch")) batch = std::atoi(argv[++i]);
        else if (val("--learned")) learned = argv[++i];
        else if (val("--learned-every")) learned_every = std::atoi(argv[++i]);
        else if (val("--learned-coarsen")) learned_coarsen = std::atoi(argv[++i]);
        else if (val("--matrix")) matrix_file = argv[++i];
        else if (is("--fgmres")) use_fgmres = true;
        else if (is("--bicgstab")) use_bicgstab = true;
        else if (is("--analyze")) analyze = true;
        else if (is("--mm-sum-duplicates")) mm.sum_duplicates = true;
        else if (is("--mm-strict")) mm.strict = true;
        else if (is("--gpu")) {}
        else if (is("--weighted-partition")) weighted = true;
        else if (is("--coarse-inverse")) coarse_inv = true;
        else if (val("--drop-tol")) drop_tol = std::atof(argv[++i]);
        else if (val("--fill-factor")) fill_factor = std::atoi(argv[++i]);
        else {
            usage();
            return 1;
        }
    }
    if (poisson <= 0 && poisson3d <= 0 && jump3d <= 0 && matrix_file.empty()) {
        usage();  // src/main.cpp:122-126
        return 1;
    }
    try {
        CSRMatrix A;
        if (poisson <= 0 && poisson3d <= 0 && jump3d <= 0) {
            std::cout << "Loading matrix: " << matrix_file << "\n";  // src/main.cpp:136-139
            A = read_matrix_market(matrix_file, mm);
            matrix = matrix_file;
            if (A.nrows != A.ncols) throw std::runtime_error("the matrix is not square");
        } else if (poisson > 0) {
            std::cout << "Generating Poisson2D " << poisson << "x" << poisson << "\n";
            A = make_poisson_2d(poisson, poisson);
            matrix = "poisson2d_" + std::to_string(poisson);
        } else if (poisson3d > 0) {
            std::cout << "Generating Poisson3D " << poisson3d << "^3\n";
            A = make_poisson_3d(poisson3d, poisson3d, poisson3d);
            matrix = "poisson3d_" + std::to_string(poisson3d);
        } else {
            std::cout << "Generating jump-coefficient diffusion 3D " << jump3d << "^3\n";
            A = make_jump3d(jump3d, 3);
            matrix = "jump3d_" + std::to_string(jump3d);
        }
        if (!matrix_file.empty()) matrix = matrix_file;  // src/main.cpp:133-135
        std::cout << "  n=" << A.nrows << "  nnz=" << A.nnz() << "\n";
        const int n = A.nrows;
        if (use_bicgstab && (use_fgmres || !learned.empty()))
            throw std::runtime_error("--bicgstab runs alone (no --fgmres / --learned)");
        auto make_rhs = [&](uint64_t sd) {
            std::vector<double> bb(n), xs(n, 1.0);
            if (rhs == "normal") {
                bb = normals(sd, n);
            } else {
                if (rhs == "uniform") {
                    SplitMix64 r(sd);
                    for (auto& v : xs) v = 2.0 * r.u01() - 1.0;
                }
                A.spmv(xs, bb);  // src/main.cpp:153-155
            }
            return bb;
        };
        std::vector<double> b = make_rhs(seed);
        if (batch > 0) {
            sg_precond_kind kind = SG_PC_NONE;
            if (precond == "jacobi") kind = SG_PC_JACOBI;
            else if (precond == "ilu0") kind = SG_PC_ILU0;
            else if (precond == "asm") kind = SG_PC_ASM;
            else if (precond == "ras") kind = SG_PC_RAS;
            else if (precond == "twolevel_asm") kind = SG_PC_TWOLEVEL_ASM;
            else if (precond == "twolevel_ras") kind = SG_PC_TWOLEVEL_RAS;
            else if (precond == "ilut") throw std::runtime_error("--batch does not take ILUT");
            std::vector<CSRMatrix> sys(batch, A);
            std::vector<std::vector<double>> bs, xs(batch);
            for (int q = 0; q < batch; ++q) bs.push_back(make_rhs(seed + q));
            prm.ortho = ortho == "cgs2" ? SG_ORTHO_CGS2 : SG_ORTHO_MGS;
            const sg_method bm = use_bicgstab ? SG_BICGSTAB : (use_fgmres ? SG_FGMRES : SG_GMRES_LEFT);
            std::cout << "Solving " << batch << " systems together with "
                      << (use_bicgstab ? std::string("BiCGSTAB")
                                       : (use_fgmres ? "FGMRES(" : "GMRES(") + std::to_string(prm.restart) + ")")
                      << " + " << precond << "\n";
            auto t0b = std::chrono::steady_clock::now();
            std::vector<GMRESResult> rs =
                solve_batch(sys, bs, xs, kind, nparts, overlap, precision == "fp32" ? SG_FP32 : SG_FP64,
                            prm, bm);
            const double tb = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0b).count();
            int conv = 0, itmin = 1 << 30, itmax = 0;
            double worst = 0.0;
            for (const auto& r : rs) {
                conv += r.converged ? 1 : 0;
                itmin = std::min(itmin, r.iterations);
                itmax = std::max(itmax, r.iterations);
                worst = std::max(worst, r.rel_true_residual);
            }
            std::cout << "  Converged: " << (conv == batch ? "yes" : "no") << " (" << conv << "/" << batch << ")\n"
                      << "  Iterations: " << itmin << ".." << itmax << "\n"
                      << "  Residual: " << rs.back().residual_norm << "\n"
                      << "  Solve time: " << tb << " s\n"
                      << "  Relative true residual: " << worst << " (worst)\n";
            if (!output.empty()) {
                Metrics m;
                m.matrix = matrix;
                m.method = std::string("BATCH_") +
                           (use_bicgstab ? "BiCGSTAB+" : use_fgmres ? "FGMRES+" : "GMRES+") + precond;
                m.n = n * batch;
                m.nnz = A.nnz() * batch;
                m.gpus = 1;
                m.overlap = overlap;
                m.restart = prm.restart;
                m.tol = prm.tol;
                m.iterations = itmax;
                m.residual_norm = rs.back().residual_norm;
                m.time_solve_s = tb;
                m.extra.emplace_back("systems", std::to_string(batch));
                m.write_jsonl(output);
            }
            return conv == batch ? 0 : 1;
        }

        std::unique_ptr<Preconditioner> M;
        auto t0 = std::chrono::steady_clock::now();
        DevicePreconditioner* D = nullptr;
        if (precond == "jacobi") M.reset(D = new JacobiPrecond());
        else if (precond == "ilu0") M.reset(D = new ILU0Precond());
        else if (precond == "asm") M.reset(D = new ASMPrecond(nparts, overlap));
        else if (precond == "ras") M.reset(D = new RASPrecond(nparts, overlap));
        else if (precond == "twolevel_asm")
            M.reset(D = new TwoLevelSchwarzPrecond(nparts, overlap, FineLevelType::ASM));
        else if (precond == "twolevel_ras")
            M.reset(D = new TwoLevelSchwarzPrecond(nparts, overlap, FineLevelType::RAS));
        else if (precond == "ilut") M.reset(D = new ILUTPrecond(drop_tol, fill_factor));
        if (D) {
            D->set_precision(precision == "fp32" ? SG_FP32 : SG_FP64);
            D->set_ufactor(ufactor);
            D->set_weighted_partition(weighted);
            D->set_coarse_inverse(coarse_inv);
            if (!part_file.empty()) {
                std::ifstream in(part_file);
                std::vector<int> p;
                int v;
                while (in >> v) p.push_back(v);
                if ((int)p.size() != n) throw std::runtime_error("partition file has the wrong length");
                D->set_partition(std::move(p));
            }
            D->setup(A);
            if (precond == "ilut")  // src/main.cpp:194-197
                std::cout << "  ILUT: drop_tol=" << drop_tol << " fill_factor=" << fill_factor << "\n";
        }
        auto t1 = std::chrono::steady_clock::now();
        const double t_setup = std::chrono::duration<double>(t1 - t0).count();

        if (analyze && M) {  // src/main.cpp:210-221
            std::cout << "\n--- Pre-solve analysis ---\n";
            print_condition_estimate(estimate_condition(A, nullptr, 80), "A (unpreconditioned)");
            print_condition_estimate(estimate_condition(A, M.get(), 80), "M^{-1}A (preconditioned)");
            print_spectral_result(estimate_spectral_radius(A, M.get(), 300), precond);
            std::cout << "--------------------------\n\n";
        }

        prm.ortho = ortho == "cgs2" ? SG_ORTHO_CGS2 : SG_ORTHO_MGS;
        const std::string label = use_fgmres ? "FGMRES(" + std::to_string(prm.restart) + ")"
                                             : "GMRES(" + std::to_string(prm.restart) + ")";
        std::unique_ptr<LearnedStage> stage;
        if (!learned.empty()) {
            if (poisson <= 0 || !D)
                throw std::runtime_error("--learned needs a 2D --poisson grid and a device --precond");
            stage.reset(new LearnedStage(*D, load_npy_f32(learned), poisson, poisson,
                                         std::vector<double>((size_t)n, 0.0), 0.01, learned_every,
                                         learned_coarsen));
            stage->setup(A);
            use_fgmres = true;  // the stage changes per application
        }
        const std::string label2 = use_bicgstab ? std::string("BiCGSTAB")
                                   : use_fgmres ? "FGMRES(" + std::to_string(prm.restart) + ")"
                                                : label;
        std::cout << "Solving with " << label2 << " + " << precond << (stage ? " + learned stage" : "") << "\n";
        std::vector<double> x(n, 0.0);
        const Preconditioner* P = stage ? static_cast<const Preconditioner*>(stage.get()) : M.get();
        GMRESResult res = use_bicgstab ? bicgstab(A, b, x, P, prm)
                          : use_fgmres ? fgmres(A, b, x, P, prm)
                                       : gmres(A, b, x, P, prm);
        if (stage) {
            auto c = stage->counts();
            std::cout << "  Learned stage: " << c.second << "/" << c.first << " corrections kept\n";
        }
        std::cout << "  Converged: " << (res.converged ? "yes" : "no") << "\n"
                  << "  Iterations: " << res.iterations << "\n"
                  << "  Residual: " << res.residual_norm << "\n"
                  << "  Setup time: " << t_setup << " s\n"
                  << "  Solve time: " << res.time_solve_s << " s\n"
                  << "  Relative true residual: " << res.rel_true_residual << "\n";
        if (analyze) {  // src/main.cpp:254-261
            ConvergenceMetrics ca = analyze_convergence(res.residual_history, res.converged);
            ca.print_summary();
            if (!output.empty())
                ca.write_jsonl(output + ".conv.jsonl", matrix, precond, nparts, overlap);
        }
        if (!output.empty()) {
            Metrics m;
            m.matrix = matrix;
            m.method = (use_bicgstab ? "BiCGSTAB+" : use_fgmres ? "FGMRES+" : "GMRES+") + precond +
                       (stage ? "+learned" : "");
            m.n = n;
            m.nnz = A.nnz();
            m.ranks = 1;
            m.threads = (int)std::thread::hardware_concurrency();
            m.gpus = 1;
            m.overlap = overlap;
            m.restart = prm.restart;
            m.tol = prm.tol;
            m.iterations = res.iterations;
            m.residual_norm = res.residual_norm;
            m.time_setup_s = t_setup;
            m.time_solve_s = res.time_solve_s;
            std::ostringstream rt;
            rt << std::scientific << std::setprecision(6) << res.rel_true_residual;
            m.extra.emplace_back("rel_true_residual", rt.str());
            m.extra.emplace_back("precision", "\"" + precision + "\"");
            m.extra.emplace_back("ortho", "\"" + ortho + "\"");
            m.write_jsonl(output);
        }
        return res.converged ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << "schwarz_solve: " << e.what() << "\n";
        return 3;
    }
}
