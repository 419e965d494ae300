// schwarz.h -- drop-in C++ interface of the MI355X Schwarz-(F)GMRES solver.
//
// Same names, signatures and semantics as the reference's host API
// (izbassarn7/schwarz-gmres-dashboard): CSRMatrix (src/core/sparse_matrix.h:13-96),
// Preconditioner and its subclasses (src/precond/preconditioner.h:7-20,
// jacobi.h, ilu0.h, asm.h, ras.h, two_level_schwarz.h), gmres / fgmres /
// fgmres_flex with GMRESParams / GMRESResult (src/core/gmres.h:12-33,
// src/core/fgmres.h:37-55), make_poisson_2d/3d (src/mesh/poisson.h:10-74) and
// Metrics (src/io/metrics_logger.h:10-57).  Every numerical operation runs on
// the GPU through the C ABI in sg_capi.h; host pointers are staged to and from
// the device.  The library's own preconditioners keep their factors on the
// device, so gmres/fgmres with them never leave the GPU inside a solve.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../sg_capi.h"

namespace schwarz {

struct CSRMatrix {
    int nrows = 0;
    int ncols = 0;
    std::vector<int> row_ptr;
    std::vector<int> col_idx;
    std::vector<double> vals;

    int nnz() const { return static_cast<int>(vals.size()); }
    void resize(int nr, int nc, int nz);
    // y = A x on the device (row sums in CSR order: bit-equal to the reference)
    void spmv(const double* x, double* y) const;
    void spmv(const std::vector<double>& x, std::vector<double>& y) const;
    double diagonal(int i) const;
    CSRMatrix extract_submatrix(const std::vector<int>& indices) const;
};

CSRMatrix make_identity(int n);
CSRMatrix make_poisson_2d(int nx, int ny);
CSRMatrix make_poisson_3d(int nx, int ny, int nz);

// The process-wide device context (device SG_DEVICE or 0), created on first use.
sg_context* default_context();

class Preconditioner {
public:
    virtual ~Preconditioner();
    virtual void setup(const CSRMatrix& A) = 0;
    // y = M^{-1} x (host pointers)
    virtual void apply(const double* x, double* y, int n) const = 0;
    void apply(const std::vector<double>& x, std::vector<double>& y) const;
    // Device handle of the library's own preconditioners; null for user classes.
    virtual sg_precond* device_handle() const { return nullptr; }
};

// Base of the device-backed preconditioners below.
class DevicePreconditioner : public Preconditioner {
public:
    void setup(const CSRMatrix& A) override;
    void apply(const double* x, double* y, int n) const override;
    sg_precond* device_handle() const override { return M_; }
    ~DevicePreconditioner() override;
    // Mixed precision (FP32 factors and applies) -- an extension.
    void set_precision(sg_precision p) { opts_.precision = p; }
    // Caller-supplied partition (host int[n]) instead of the built-in partitioner.
    void set_partition(std::vector<int> part) { part_ = std::move(part); }
    void set_ufactor(int u) { opts_.ufactor = u; }
    void set_weighted_partition(bool w) { opts_.weighted_partition = w ? 1 : 0; }
    void set_coarse_inverse(bool on) { opts_.coarse_solve = on ? 1 : 0; }
    std::vector<int> partition() const;

protected:
    DevicePreconditioner(sg_precond_kind kind, int nparts, int overlap);
    sg_precond_kind kind_;
    sg_precond_opts opts_{};
    std::vector<int> part_;
    sg_precond* M_ = nullptr;
    int n_ = 0;
};

class JacobiPrecond : public DevicePreconditioner {
public:
    JacobiPrecond() : DevicePreconditioner(SG_PC_JACOBI, 1, 0) {}
};

class ILU0Precond : public DevicePreconditioner {
public:
    ILU0Precond() : DevicePreconditioner(SG_PC_ILU0, 1, 0) {}
    // factor values in CSR order, as ILU0Precond::lu_vals (src/precond/ilu0.h:83)
    std::vector<double> lu_vals() const;
};

// Dual-threshold ILU (src/precond/ilut.h:21-193): factorisation on the host as
// the reference's, triangular sweeps on the device.
class ILUTPrecond : public DevicePreconditioner {
public:
    explicit ILUTPrecond(double drop_tol = 1e-4, int fill_factor = 10)
        : DevicePreconditioner(SG_PC_ILUT, 1, 0) {
        opts_.drop_tol = drop_tol;
        opts_.fill_factor = fill_factor;
    }
    int n() const { return n_; }
    int l_nnz() const;
    int u_nnz() const;
};

class ASMPrecond : public DevicePreconditioner {
public:
    ASMPrecond(int nparts, int overlap) : DevicePreconditioner(SG_PC_ASM, nparts, overlap) {}
};

class RASPrecond : public DevicePreconditioner {
public:
    RASPrecond(int nparts, int overlap) : DevicePreconditioner(SG_PC_RAS, nparts, overlap) {}
};

enum class FineLevelType { ASM, RAS };

class TwoLevelSchwarzPrecond : public DevicePreconditioner {
public:
    TwoLevelSchwarzPrecond(int nparts, int overlap, FineLevelType fine = FineLevelType::RAS)
        : DevicePreconditioner(fine == FineLevelType::ASM ? SG_PC_TWOLEVEL_ASM
                                                          : SG_PC_TWOLEVEL_RAS,
                               nparts, overlap) {}
};

struct GMRESResult {
    int iterations = 0;
    double residual_norm = 0.0;
    bool converged = false;
    double time_solve_s = 0.0;
    std::vector<double> residual_history;
    // extensions
    double rel_true_residual = 0.0;
};

struct GMRESParams {
    int restart = 30;
    int max_iter = 1000;
    double tol = 1e-10;
    // extensions: rtol > 0 stops on an FP64-verified ||b-Ax||/||b|| <= rtol;
    // ortho selects MGS (reference order) or CGS2.
    double rtol = 0.0;
    sg_ortho ortho = SG_ORTHO_MGS;
};

// Restarted GMRES(m) with left preconditioning; x is the initial guess and the result.
GMRESResult gmres(const CSRMatrix& A, const std::vector<double>& b, std::vector<double>& x,
                  const Preconditioner* M, const GMRESParams& params = {});

// FGMRES(m) with right (flexible) preconditioning.
GMRESResult fgmres(const CSRMatrix& A, const std::vector<double>& b, std::vector<double>& x,
                   const Preconditioner* M, const GMRESParams& params = {});

// Right-preconditioned BiCGSTAB (src/core/bicgstab.h:57-204): params.tol is
// RELATIVE to ||r0||, max_iter bounds iterations, restart is unused.
GMRESResult bicgstab(const CSRMatrix& A, const std::vector<double>& b, std::vector<double>& x,
                     const Preconditioner* M, const GMRESParams& params = {});

using FlexApply = std::function<void(const double*, double*, int)>;
GMRESResult fgmres_flex(const CSRMatrix& A, const std::vector<double>& b,
                        std::vector<double>& x, FlexApply prec_apply,
                        const GMRESParams& params = {});

// Matrix Market coordinate files (src/io/matrix_market.h:14-103).  The
// one-argument reader behaves as the reference's (header flags from any '%'
// line, symmetric mirroring, entries sorted by (row, col), duplicates kept);
// the options fix its duplicate handling and header parsing (sg_mm_flags).
// Errors throw std::runtime_error ("Cannot open file: ...", "Premature end of
// file", or a message naming the file and line).
struct MatrixMarketOptions {
    bool sum_duplicates = false;  // merge entries with the same (row, col)
    bool strict = false;          // banner-only flags, skew sign, blank lines, clean errors
};
CSRMatrix read_matrix_market(const std::string& filename);
CSRMatrix read_matrix_market(const std::string& filename, const MatrixMarketOptions& opts);
// digits: significant digits of the values (6 = the reference's output; 17 round-trips)
void write_matrix_market(const std::string& filename, const CSRMatrix& A, int digits = 6);

// ---- convergence analytics (src/analysis/*.h; called under --analyze) ----------

// Per-run residual-history metrics (convergence_analysis.h:15-106).
struct ConvergenceMetrics {
    std::vector<double> residuals;
    std::vector<double> rates;  // (r_k / r_0)^(1/k)
    double initial_residual = 0.0;
    double final_residual = 0.0;
    double asymptotic_rate = 0.0;  // mean of the last 20 rates
    int stagnation_start = -1;     // -1 = none
    bool converged = false;
    int iterations = 0;
    static constexpr int STAGNATION_WINDOW = 20;
    static constexpr double STAGNATION_TOL = 1e-3;

    void compute(const std::vector<double>& res_history, bool conv);
    // appends one "type":"convergence" JSON line
    void write_jsonl(const std::string& path, const std::string& matrix,
                     const std::string& precond, int nparts, int overlap) const;
    void print_summary() const;
};
ConvergenceMetrics analyze_convergence(const std::vector<double>& residual_history,
                                       bool converged);

// Lanczos estimate of kappa(M^{-1}A) (condition_estimator.h:31-211), on the device.
struct ConditionEstimate {
    double lambda_min = 0.0;
    double lambda_max = 0.0;
    double kappa = 0.0;
    int lanczos_steps = 0;
    bool valid = false;
};
ConditionEstimate estimate_condition(const CSRMatrix& A, const Preconditioner* M, int k = 80,
                                     unsigned int seed = 42);
void print_condition_estimate(const ConditionEstimate& est, const std::string& label = "");

// Power-iteration estimate of rho(I - M^{-1}A) (spectral_analysis.h:19-107), on the device.
struct SpectralResult {
    double spectral_radius = 0.0;
    bool convergence_ok = false;
    int power_iters = 0;
    std::vector<double> eigenvector;
};
SpectralResult estimate_spectral_radius(const CSRMatrix& A, const Preconditioner* M,
                                        int max_iter = 200, double tol = 1e-6,
                                        unsigned int seed = 123);
void print_spectral_result(const SpectralResult& sr, const std::string& label = "");

// ---- extensions without a reference counterpart ----------------------------------

// Learned FGMRES stage on an nx x ny grid (row y*nx + x): base(x) plus the
// network correction of its residual, kept only when it reduces the residual
// (sg_precond_create_learned).  Valid with fgmres only.  setup(A) also sets up
// base; base must outlive the stage.
class LearnedStage : public Preconditioner {
public:
    // coarsen 1, 2 or 4: the network runs on c x c block averages (sg_learned_opts)
    LearnedStage(DevicePreconditioner& base, std::vector<float> params, int nx, int ny,
                 std::vector<double> log_kappa, double min_gain = 0.01, int every = 1,
                 int coarsen = 1);
    ~LearnedStage() override;
    void setup(const CSRMatrix& A) override;
    void apply(const double* x, double* y, int n) const override;
    sg_precond* device_handle() const override { return M_; }
    // applications so far and how many kept the network's correction
    std::pair<long long, long long> counts() const;

private:
    DevicePreconditioner& base_;
    std::vector<float> params_;
    std::vector<double> logk_;
    int nx_, ny_, every_, coarsen_;
    double min_gain_;
    sg_unet* net_ = nullptr;
    sg_precond* M_ = nullptr;
    sg_csr* A_ = nullptr;
};

// A 1-D little-endian float32 .npy file (the learned stage's weights).
std::vector<float> load_npy_f32(const std::string& path);

// Many independent systems solved together (sg_solve_batch): one block-diagonal
// device matrix, `kind` with nparts subdomains per system (built-in
// partitioner per system), each system with its own GMRES/FGMRES state.
// x[i] is system i's initial guess and result.
std::vector<GMRESResult> solve_batch(const std::vector<CSRMatrix>& systems,
                                     const std::vector<std::vector<double>>& b,
                                     std::vector<std::vector<double>>& x, sg_precond_kind kind,
                                     int nparts, int overlap, sg_precision precision,
                                     const GMRESParams& params, bool flexible);
// The same with any sg_method (SG_BICGSTAB included).
std::vector<GMRESResult> solve_batch(const std::vector<CSRMatrix>& systems,
                                     const std::vector<std::vector<double>>& b,
                                     std::vector<std::vector<double>>& x, sg_precond_kind kind,
                                     int nparts, int overlap, sg_precision precision,
                                     const GMRESParams& params, sg_method method);

// One JSON object per line, appended (field order and number formats as the
// reference's logger; extra fields may only be appended after "efficiency").
struct Metrics {
    std::string matrix;
    std::string method;
    int n = 0;
    int nnz = 0;
    int ranks = 1;
    int threads = 1;
    int gpus = 0;
    int overlap = 0;
    int restart = 30;
    int iterations = 0;
    double tol = 1e-10;
    double residual_norm = 0.0;
    double time_setup_s = 0.0;
    double time_solve_s = 0.0;
    double speedup = 1.0;
    double efficiency = 1.0;
    // appended fields: written as "key":value in insertion order when non-empty
    std::vector<std::pair<std::string, std::string>> extra;

    std::string to_json() const;
    void write_jsonl(const std::string& path) const;
};

}  // namespace schwarz
