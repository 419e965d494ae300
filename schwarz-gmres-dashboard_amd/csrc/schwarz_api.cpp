// Drop-in C++ interface (include/schwarz/schwarz.h) over the C ABI.
#include "../../include/schwarz/schwarz.h"

#include <chrono>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <mutex>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace schwarz {

namespace {

void check(sg_status st, sg_context* ctx) {
    if (st == SG_OK) return;
    throw std::runtime_error(sg_last_error(ctx));
}

// Device buffer owned by the default context.
struct Dev {
    void* p = nullptr;
    explicit Dev(size_t bytes) { check(sg_malloc(default_context(), bytes, &p), default_context()); }
    ~Dev() {
        if (p) sg_free(default_context(), p);
    }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    double* d() const { return static_cast<double*>(p); }
};

struct DevCsr {
    sg_csr* h = nullptr;
    explicit DevCsr(const CSRMatrix& A) {
        check(sg_csr_create(default_context(), A.nrows, A.ncols, A.nnz(), A.row_ptr.data(),
                            A.col_idx.data(), A.vals.data(), &h),
              default_context());
    }
    ~DevCsr() {
        if (h) sg_csr_destroy(h);
    }
};

void host_callback(void* user, const double* x, double* y, int n) {
    static_cast<const Preconditioner*>(user)->apply(x, y, n);
}

void flex_callback(void* user, const double* x, double* y, int n) {
    (*static_cast<FlexApply*>(user))(x, y, n);
}

GMRESResult run_solve(const CSRMatrix& A, const std::vector<double>& b, std::vector<double>& x,
                      sg_precond* M, int method, const GMRESParams& p) {
    sg_context* ctx = default_context();
    const int n = A.nrows;
    if ((int)b.size() < n) throw std::runtime_error("b is shorter than the matrix");
    if ((int)x.size() < n) x.resize(n, 0.0);
    DevCsr dA(A);
    Dev db(sizeof(double) * std::max(n, 1)), dx(sizeof(double) * std::max(n, 1));
    check(sg_memcpy_h2d(ctx, db.p, b.data(), sizeof(double) * n), ctx);
    check(sg_memcpy_h2d(ctx, dx.p, x.data(), sizeof(double) * n), ctx);
    sg_solve_params sp{};
    sp.method = method;
    sp.restart = p.restart;
    sp.max_iter = p.max_iter;
    sp.tol_abs = p.tol;
    sp.rtol = p.rtol;
    sp.ortho = p.ortho;
    sg_solve_result r{};
    std::vector<double> hist((size_t)std::max(0, p.max_iter) + p.restart + 2);
    check(sg_solve(dA.h, M, db.d(), dx.d(), &sp, &r, hist.data(), (int)hist.size()), ctx);
    check(sg_memcpy_d2h(ctx, x.data(), dx.p, sizeof(double) * n), ctx);
    GMRESResult out;
    out.iterations = r.iterations;
    out.converged = r.converged != 0;
    out.residual_norm = r.residual_norm;
    out.time_solve_s = r.time_solve_s;
    out.rel_true_residual = r.rel_true_residual;
    hist.resize(std::min<size_t>(hist.size(), (size_t)r.iterations));
    out.residual_history = std::move(hist);
    return out;
}

}  // namespace

sg_context* default_context() {
    static std::once_flag once;
    static sg_context* ctx = nullptr;
    std::call_once(once, [] {
        const char* dev = std::getenv("SG_DEVICE");
        sg_status st = sg_context_create(dev ? std::atoi(dev) : 0, &ctx);
        if (st != SG_OK)
            throw std::runtime_error(std::string("no usable HIP device: ") + sg_last_error(nullptr));
    });
    if (!ctx) throw std::runtime_error("no usable HIP device");
    return ctx;
}

// ---- CSRMatrix -------------------------------------------------------------

void CSRMatrix::resize(int nr, int nc, int nz) {
    nrows = nr;
    ncols = nc;
    row_ptr.resize(nr + 1);
    col_idx.resize(nz);
    vals.resize(nz);
}

void CSRMatrix::spmv(const double* x, double* y) const {
    sg_context* ctx = default_context();
    DevCsr dA(*this);
    Dev dx(sizeof(double) * std::max(ncols, 1)), dy(sizeof(double) * std::max(nrows, 1));
    check(sg_memcpy_h2d(ctx, dx.p, x, sizeof(double) * ncols), ctx);
    check(sg_spmv(dA.h, 1.0, dx.d(), 0.0, dy.d()), ctx);
    check(sg_memcpy_d2h(ctx, y, dy.p, sizeof(double) * nrows), ctx);
}

void CSRMatrix::spmv(const std::vector<double>& x, std::vector<double>& y) const {
    if ((int)x.size() < ncols) throw std::runtime_error("x is shorter than ncols");
    y.resize(nrows);
    spmv(x.data(), y.data());
}

double CSRMatrix::diagonal(int i) const {
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
        if (col_idx[k] == i) return vals[k];
    return 0.0;
}

CSRMatrix CSRMatrix::extract_submatrix(const std::vector<int>& idx) const {
    const int m = (int)idx.size();
    std::vector<int> inv(std::max(nrows, ncols), -1);
    for (int i = 0; i < m; ++i) inv[idx[i]] = i;
    CSRMatrix s;
    s.nrows = s.ncols = m;
    s.row_ptr.assign(m + 1, 0);
    for (int li = 0; li < m; ++li) {
        const int gi = idx[li];
        s.row_ptr[li] = (int)s.col_idx.size();
        for (int k = row_ptr[gi]; k < row_ptr[gi + 1]; ++k) {
            const int lj = inv[col_idx[k]];
            if (lj >= 0) {
                s.col_idx.push_back(lj);
                s.vals.push_back(vals[k]);
            }
        }
    }
    s.row_ptr[m] = (int)s.col_idx.size();
    return s;
}

CSRMatrix make_identity(int n) {
    CSRMatrix I;
    I.nrows = I.ncols = n;
    I.row_ptr.resize(n + 1);
    I.col_idx.resize(n);
    I.vals.assign(n, 1.0);
    std::iota(I.row_ptr.begin(), I.row_ptr.end(), 0);
    std::iota(I.col_idx.begin(), I.col_idx.end(), 0);
    return I;
}

// Dirichlet Laplacians, h = 1 scaling; rows z-major, then y, then x; columns
// ascending within a row (the neighbour order of src/mesh/poisson.h).
CSRMatrix make_poisson_2d(int nx, int ny) {
    CSRMatrix A;
    const int n = nx * ny;
    A.nrows = A.ncols = n;
    A.row_ptr.assign(n + 1, 0);
    A.col_idx.reserve((size_t)5 * n);
    A.vals.reserve((size_t)5 * n);
    auto put = [&](int c, double v) {
        A.col_idx.push_back(c);
        A.vals.push_back(v);
    };
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const int i = iy * nx + ix;
            A.row_ptr[i] = (int)A.col_idx.size();
            if (iy > 0) put(i - nx, -1.0);
            if (ix > 0) put(i - 1, -1.0);
            put(i, 4.0);
            if (ix + 1 < nx) put(i + 1, -1.0);
            if (iy + 1 < ny) put(i + nx, -1.0);
        }
    A.row_ptr[n] = (int)A.col_idx.size();
    return A;
}

CSRMatrix make_poisson_3d(int nx, int ny, int nz) {
    CSRMatrix A;
    const int nxy = nx * ny, n = nxy * nz;
    A.nrows = A.ncols = n;
    A.row_ptr.assign(n + 1, 0);
    A.col_idx.reserve((size_t)7 * n);
    A.vals.reserve((size_t)7 * n);
    auto put = [&](int c, double v) {
        A.col_idx.push_back(c);
        A.vals.push_back(v);
    };
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int i = iz * nxy + iy * nx + ix;
                A.row_ptr[i] = (int)A.col_idx.size();
                if (iz > 0) put(i - nxy, -1.0);
                if (iy > 0) put(i - nx, -1.0);
                if (ix > 0) put(i - 1, -1.0);
                put(i, 6.0);
                if (ix + 1 < nx) put(i + 1, -1.0);
                if (iy + 1 < ny) put(i + nx, -1.0);
                if (iz + 1 < nz) put(i + nxy, -1.0);
            }
    A.row_ptr[n] = (int)A.col_idx.size();
    return A;
}

// ---- preconditioners ---------------------------------------------------------

Preconditioner::~Preconditioner() = default;

void Preconditioner::apply(const std::vector<double>& x, std::vector<double>& y) const {
    y.resize(x.size());
    apply(x.data(), y.data(), static_cast<int>(x.size()));
}

DevicePreconditioner::DevicePreconditioner(sg_precond_kind kind, int nparts, int overlap)
    : kind_(kind) {
    opts_.nparts = nparts;
    opts_.overlap = overlap;
    opts_.precision = SG_FP64;
}

DevicePreconditioner::~DevicePreconditioner() {
    if (M_) sg_precond_destroy(M_);
}

void DevicePreconditioner::setup(const CSRMatrix& A) {
    sg_context* ctx = default_context();
    if (M_) {
        sg_precond_destroy(M_);
        M_ = nullptr;
    }
    sg_precond_opts o = opts_;
    o.partition = part_.empty() ? nullptr : part_.data();
    check(sg_precond_create(ctx, kind_, &o, &M_), ctx);
    DevCsr dA(A);
    check(sg_precond_setup(M_, dA.h), ctx);
    n_ = A.nrows;
}

void DevicePreconditioner::apply(const double* x, double* y, int n) const {
    sg_context* ctx = default_context();
    if (!M_) throw std::runtime_error("preconditioner used before setup");
    if (n != n_) throw std::runtime_error("preconditioner size mismatch");
    Dev dx(sizeof(double) * std::max(n, 1)), dy(sizeof(double) * std::max(n, 1));
    check(sg_memcpy_h2d(ctx, dx.p, x, sizeof(double) * n), ctx);
    check(sg_precond_apply(M_, dx.d(), dy.d()), ctx);
    check(sg_memcpy_d2h(ctx, y, dy.p, sizeof(double) * n), ctx);
}

std::vector<int> DevicePreconditioner::partition() const {
    std::vector<int> p(n_);
    check(sg_precond_partition(M_, p.data()), default_context());
    return p;
}

std::vector<double> ILU0Precond::lu_vals() const {
    int nnz = 0;
    check(sg_precond_ilu_values(M_, 0, &nnz, nullptr), default_context());
    std::vector<double> v(nnz);
    check(sg_precond_ilu_values(M_, 0, &nnz, v.data()), default_context());
    return v;
}

// ---- solvers -------------------------------------------------------------------

static GMRESResult solve_with(const CSRMatrix& A, const std::vector<double>& b,
                              std::vector<double>& x, const Preconditioner* M, int method,
                              const GMRESParams& p) {
    if (!M) return run_solve(A, b, x, nullptr, method, p);
    if (sg_precond* h = M->device_handle()) return run_solve(A, b, x, h, method, p);
    // a user-defined host preconditioner: staged through host memory per apply
    sg_context* ctx = default_context();
    sg_precond* cb = nullptr;
    check(sg_precond_create_callback(ctx, A.nrows, host_callback,
                                     const_cast<Preconditioner*>(M), 0, &cb),
          ctx);
    struct Drop {
        sg_precond* p;
        ~Drop() { sg_precond_destroy(p); }
    } drop{cb};
    return run_solve(A, b, x, cb, method, p);
}

GMRESResult gmres(const CSRMatrix& A, const std::vector<double>& b, std::vector<double>& x,
                  const Preconditioner* M, const GMRESParams& params) {
    return solve_with(A, b, x, M, SG_GMRES_LEFT, params);
}

GMRESResult fgmres(const CSRMatrix& A, const std::vector<double>& b, std::vector<double>& x,
                   const Preconditioner* M, const GMRESParams& params) {
    return solve_with(A, b, x, M, SG_FGMRES, params);
}

GMRESResult fgmres_flex(const CSRMatrix& A, const std::vector<double>& b,
                        std::vector<double>& x, FlexApply prec_apply,
                        const GMRESParams& params) {
    if (!prec_apply) return run_solve(A, b, x, nullptr, SG_FGMRES, params);
    sg_context* ctx = default_context();
    sg_precond* cb = nullptr;
    check(sg_precond_create_callback(ctx, A.nrows, flex_callback, &prec_apply, 0, &cb), ctx);
    struct Drop {
        sg_precond* p;
        ~Drop() { sg_precond_destroy(p); }
    } drop{cb};
    return run_solve(A, b, x, cb, SG_FGMRES, params);
}

// ---- Metrics ---------------------------------------------------------------------

// setprecision(6); tol in default notation, residual_norm in scientific, then
// fixed for the remaining doubles (the reference's logger format).
std::string Metrics::to_json() const {
    std::ostringstream os;
    os << std::setprecision(6) << "{"
       << "\"matrix\":\"" << matrix << "\","
       << "\"method\":\"" << method << "\","
       << "\"n\":" << n << ","
       << "\"nnz\":" << nnz << ","
       << "\"ranks\":" << ranks << ","
       << "\"threads\":" << threads << ","
       << "\"gpus\":" << gpus << ","
       << "\"overlap\":" << overlap << ","
       << "\"restart\":" << restart << ","
       << "\"iterations\":" << iterations << ","
       << "\"tol\":" << tol << ","
       << "\"residual_norm\":" << std::scientific << residual_norm << "," << std::fixed
       << "\"time_setup_s\":" << time_setup_s << ","
       << "\"time_solve_s\":" << time_solve_s << ","
       << "\"speedup\":" << speedup << ","
       << "\"efficiency\":" << efficiency;
    for (const auto& kv : extra) os << ",\"" << kv.first << "\":" << kv.second;
    os << "}";
    return os.str();
}

void Metrics::write_jsonl(const std::string& path) const {
    std::ofstream out(path, std::ios::app);
    out << to_json() << "\n";
}

}  // namespace schwarz
