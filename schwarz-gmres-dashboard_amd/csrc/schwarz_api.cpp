// Drop-in C++ interface (include/schwarz/schwarz.h) over the C ABI.
#include "../../include/schwarz/schwarz.h"

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
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
    hist.resize(std::min<size_t>(hist.size(), (size_t)r.hist_len));
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

int ILUTPrecond::l_nnz() const {
    long long l = 0;
    if (M_) check(sg_precond_ilut_nnz(M_, &l, nullptr), default_context());
    return (int)l;
}

int ILUTPrecond::u_nnz() const {
    long long u = 0;
    if (M_) check(sg_precond_ilut_nnz(M_, nullptr, &u), default_context());
    return (int)u;
}

std::vector<double> ILU0Precond::lu_vals() const {
    int nnz = 0;
    check(sg_precond_ilu_values(M_, 0, &nnz, nullptr), default_context());
    std::vector<double> v(nnz);
    check(sg_precond_ilu_values(M_, 0, &nnz, v.data()), default_context());
    return v;
}

// ---- learned stage ----------------------------------------------------------------

LearnedStage::LearnedStage(DevicePreconditioner& base, std::vector<float> params, int nx, int ny,
                           std::vector<double> log_kappa, double min_gain, int every, int coarsen)
    : base_(base), params_(std::move(params)), logk_(std::move(log_kappa)), nx_(nx), ny_(ny),
      every_(every), coarsen_(coarsen), min_gain_(min_gain) {
    if ((int)params_.size() != sg_unet_nparams())
        throw std::runtime_error("learned stage: wrong number of network parameters");
    if ((long long)logk_.size() != (long long)nx * ny)
        throw std::runtime_error("learned stage: log kappa must have nx * ny values");
}

LearnedStage::~LearnedStage() {
    if (M_) sg_precond_destroy(M_);
    if (net_) sg_unet_destroy(net_);
    if (A_) sg_csr_destroy(A_);
}

void LearnedStage::setup(const CSRMatrix& A) {
    sg_context* ctx = default_context();
    base_.setup(A);
    if (!net_) {
        check(sg_unet_create(ctx, &net_), ctx);
        check(sg_unet_set_params(net_, params_.data()), ctx);
    }
    if (M_) sg_precond_destroy(M_), M_ = nullptr;
    if (A_) sg_csr_destroy(A_), A_ = nullptr;
    sg_learned_opts o{nx_, ny_, min_gain_, every_, coarsen_};
    check(sg_precond_create_learned(ctx, base_.device_handle(), net_, logk_.data(), &o, &M_), ctx);
    check(sg_csr_create(ctx, A.nrows, A.ncols, A.nnz(), A.row_ptr.data(), A.col_idx.data(),
                        A.vals.data(), &A_),
          ctx);
    check(sg_precond_setup(M_, A_), ctx);
}

void LearnedStage::apply(const double* x, double* y, int n) const {
    sg_context* ctx = default_context();
    if (!M_) throw std::runtime_error("preconditioner used before setup");
    Dev dx(sizeof(double) * std::max(n, 1)), dy(sizeof(double) * std::max(n, 1));
    check(sg_memcpy_h2d(ctx, dx.p, x, sizeof(double) * n), ctx);
    check(sg_precond_apply(M_, dx.d(), dy.d()), ctx);
    check(sg_memcpy_d2h(ctx, y, dy.p, sizeof(double) * n), ctx);
}

std::pair<long long, long long> LearnedStage::counts() const {
    long long c = 0, a = 0;
    if (M_) check(sg_precond_learned_stats(M_, &c, &a), default_context());
    return {c, a};
}

std::vector<float> load_npy_f32(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    char magic[6];
    in.read(magic, 6);
    if (std::string(magic, 6) != "\x93NUMPY") throw std::runtime_error(path + ": not a .npy file");
    unsigned char ver[2];
    in.read(reinterpret_cast<char*>(ver), 2);
    uint32_t hlen = 0;
    if (ver[0] == 1) {
        unsigned char h[2];
        in.read(reinterpret_cast<char*>(h), 2);
        hlen = h[0] | (h[1] << 8);
    } else {
        unsigned char h[4];
        in.read(reinterpret_cast<char*>(h), 4);
        hlen = h[0] | (h[1] << 8) | (h[2] << 16) | ((uint32_t)h[3] << 24);
    }
    std::string header(hlen, ' ');
    in.read(&header[0], hlen);
    if (header.find("'<f4'") == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
        throw std::runtime_error(path + ": expected a C-ordered little-endian float32 array");
    const size_t l = header.find('('), r = header.find(')');
    long long count = 1;
    std::string shape = header.substr(l + 1, r - l - 1);
    std::stringstream ss(shape);
    std::string tok;
    while (std::getline(ss, tok, ','))
        if (tok.find_first_not_of(' ') != std::string::npos) count *= std::stoll(tok);
    std::vector<float> v((size_t)count);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(sizeof(float) * count));
    if (!in) throw std::runtime_error(path + ": truncated");
    return v;
}

// ---- batched solve -------------------------------------------------------------------

std::vector<GMRESResult> solve_batch(const std::vector<CSRMatrix>& systems,
                                     const std::vector<std::vector<double>>& b,
                                     std::vector<std::vector<double>>& x, sg_precond_kind kind,
                                     int nparts, int overlap, sg_precision precision,
                                     const GMRESParams& p, bool flexible) {
    return solve_batch(systems, b, x, kind, nparts, overlap, precision, p,
                       flexible ? SG_FGMRES : SG_GMRES_LEFT);
}

std::vector<GMRESResult> solve_batch(const std::vector<CSRMatrix>& systems,
                                     const std::vector<std::vector<double>>& b,
                                     std::vector<std::vector<double>>& x, sg_precond_kind kind,
                                     int nparts, int overlap, sg_precision precision,
                                     const GMRESParams& p, sg_method method) {
    sg_context* ctx = default_context();
    const int ns = (int)systems.size();
    if (ns < 1 || (int)b.size() != ns) throw std::runtime_error("solve_batch: one b per system");
    x.resize(ns);
    std::vector<int> off(1, 0), rp(1, 0), ci, part;
    std::vector<double> v, hb, hx;
    for (int q = 0; q < ns; ++q) {
        const CSRMatrix& A = systems[q];
        if (A.nrows != A.ncols || (int)b[q].size() < A.nrows)
            throw std::runtime_error("solve_batch: square systems with matching b");
        x[q].resize(A.nrows, 0.0);
        const int o = off.back(), nz0 = rp.back();
        for (int i = 0; i < A.nrows; ++i) rp.push_back(nz0 + A.row_ptr[i + 1]);
        for (int c : A.col_idx) ci.push_back(c + o);
        v.insert(v.end(), A.vals.begin(), A.vals.end());
        hb.insert(hb.end(), b[q].begin(), b[q].begin() + A.nrows);
        hx.insert(hx.end(), x[q].begin(), x[q].end());
        std::vector<int> pq(A.nrows, 0);
        if (nparts > 1)
            check(sg_partition(A.nrows, A.row_ptr.data(), A.col_idx.data(), A.vals.data(), nparts,
                               0, 0, pq.data()),
                  ctx);
        for (int e : pq) part.push_back(e + q * std::max(1, nparts));
        off.push_back(o + A.nrows);
    }
    const int n = off.back();
    sg_csr* B = nullptr;
    check(sg_csr_create_batch(ctx, ns, off.data(), rp.back(), rp.data(), ci.data(), v.data(), &B),
          ctx);
    struct DropB {
        sg_csr* b;
        ~DropB() { sg_csr_destroy(b); }
    } dropb{B};
    sg_precond* M = nullptr;
    if (kind != SG_PC_NONE) {
        sg_precond_opts o{};
        o.nparts = ns * std::max(1, nparts);
        o.overlap = overlap;
        o.precision = precision;
        o.partition = part.data();
        check(sg_precond_create(ctx, kind, &o, &M), ctx);
    }
    struct DropM {
        sg_precond* m;
        ~DropM() {
            if (m) sg_precond_destroy(m);
        }
    } dropm{M};
    if (M) check(sg_precond_setup(M, B), ctx);
    Dev db(sizeof(double) * n), dx(sizeof(double) * n);
    check(sg_memcpy_h2d(ctx, db.p, hb.data(), sizeof(double) * n), ctx);
    check(sg_memcpy_h2d(ctx, dx.p, hx.data(), sizeof(double) * n), ctx);
    sg_solve_params sp{};
    sp.method = method;
    sp.restart = p.restart;
    sp.max_iter = p.max_iter;
    sp.tol_abs = p.tol;
    sp.rtol = p.rtol;
    sp.ortho = p.ortho;
    std::vector<sg_solve_result> r(ns);
    check(sg_solve_batch(B, M, db.d(), dx.d(), &sp, r.data(), nullptr, 0), ctx);
    check(sg_memcpy_d2h(ctx, hx.data(), dx.p, sizeof(double) * n), ctx);
    std::vector<GMRESResult> out(ns);
    for (int q = 0; q < ns; ++q) {
        std::copy(hx.begin() + off[q], hx.begin() + off[q + 1], x[q].begin());
        out[q].iterations = r[q].iterations;
        out[q].converged = r[q].converged != 0;
        out[q].residual_norm = r[q].residual_norm;
        out[q].time_solve_s = r[q].time_solve_s;
        out[q].rel_true_residual = r[q].rel_true_residual;
    }
    return out;
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

GMRESResult bicgstab(const CSRMatrix& A, const std::vector<double>& b, std::vector<double>& x,
                     const Preconditioner* M, const GMRESParams& params) {
    return solve_with(A, b, x, M, SG_BICGSTAB, params);
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

// ---- Matrix Market ----------------------------------------------------------------

CSRMatrix read_matrix_market(const std::string& filename) {
    return read_matrix_market(filename, MatrixMarketOptions{});
}

CSRMatrix read_matrix_market(const std::string& filename, const MatrixMarketOptions& o) {
    int nr = 0, nc = 0, nz = 0;
    int *rp = nullptr, *ci = nullptr;
    double* v = nullptr;
    const int flags = (o.sum_duplicates ? SG_MM_SUM_DUPLICATES : 0) | (o.strict ? SG_MM_STRICT : 0);
    if (sg_mm_read(filename.c_str(), flags, &nr, &nc, &nz, &rp, &ci, &v) != SG_OK)
        throw std::runtime_error(sg_last_error(nullptr));
    CSRMatrix A;
    A.nrows = nr;
    A.ncols = nc;
    A.row_ptr.assign(rp, rp + nr + 1);
    A.col_idx.assign(ci, ci + nz);
    A.vals.assign(v, v + nz);
    sg_mm_free(rp);
    sg_mm_free(ci);
    sg_mm_free(v);
    return A;
}

void write_matrix_market(const std::string& filename, const CSRMatrix& A, int digits) {
    std::vector<int> rp = A.row_ptr;
    if (rp.empty()) rp.assign(1, 0);
    if (sg_mm_write(filename.c_str(), A.nrows, A.ncols, rp.data(), A.col_idx.data(),
                    A.vals.data(), digits) != SG_OK)
        throw std::runtime_error(sg_last_error(nullptr));
}

// ---- convergence analytics -------------------------------------------------------

void ConvergenceMetrics::compute(const std::vector<double>& h, bool conv) {
    residuals = h;
    sg_convergence c{};
    rates.assign(h.size(), 0.0);
    check(sg_convergence_analyze(h.data(), (int)h.size(), conv ? 1 : 0, &c, rates.data()),
          nullptr);
    if (h.empty()) rates.clear();
    initial_residual = c.initial_residual;
    final_residual = c.final_residual;
    asymptotic_rate = c.asymptotic_rate;
    stagnation_start = c.stagnation_start;
    converged = c.converged != 0;
    iterations = c.iterations;
}

void ConvergenceMetrics::write_jsonl(const std::string& path, const std::string& matrix,
                                     const std::string& precond, int nparts,
                                     int overlap) const {
    check(sg_convergence_write_jsonl(path.c_str(), matrix.c_str(), precond.c_str(), nparts,
                                     overlap, residuals.data(), (int)residuals.size(),
                                     converged ? 1 : 0),
          nullptr);
}

void ConvergenceMetrics::print_summary() const {
    std::cout << "  Convergence analysis:\n"
              << "    Initial residual : " << initial_residual << "\n"
              << "    Final residual   : " << final_residual << "\n"
              << "    Asymptotic rate  : " << asymptotic_rate << "\n";
    if (stagnation_start >= 0)
        std::cout << "    *** Stagnation detected at iter " << stagnation_start << " ***\n";
    else
        std::cout << "    No stagnation detected\n";
}

ConvergenceMetrics analyze_convergence(const std::vector<double>& h, bool converged) {
    ConvergenceMetrics m;
    m.compute(h, converged);
    return m;
}

namespace {

// Device handle of M for the analysis calls: the library's own preconditioner,
// or a host-callback wrapper around a user-defined one (destroyed by the guard).
struct PrecondHandle {
    sg_precond* h = nullptr;
    sg_precond* owned = nullptr;
    PrecondHandle(const Preconditioner* M, int n) {
        if (!M) return;
        h = M->device_handle();
        if (h) return;
        sg_context* ctx = default_context();
        check(sg_precond_create_callback(ctx, n, host_callback, const_cast<Preconditioner*>(M),
                                         0, &owned),
              ctx);
        h = owned;
    }
    ~PrecondHandle() {
        if (owned) sg_precond_destroy(owned);
    }
};

}  // namespace

ConditionEstimate estimate_condition(const CSRMatrix& A, const Preconditioner* M, int k,
                                     unsigned int seed) {
    sg_context* ctx = default_context();
    DevCsr dA(A);
    PrecondHandle ph(M, A.nrows);
    sg_condition_estimate e{};
    check(sg_estimate_condition(dA.h, ph.h, k, seed, &e, nullptr, nullptr), ctx);
    ConditionEstimate out;
    out.lambda_min = e.lambda_min;
    out.lambda_max = e.lambda_max;
    out.kappa = e.kappa;
    out.lanczos_steps = e.lanczos_steps;
    out.valid = e.valid != 0;
    return out;
}

void print_condition_estimate(const ConditionEstimate& est, const std::string& label) {
    std::cout << "  Condition estimate" << (label.empty() ? "" : " (" + label + ")")
              << " [Lanczos " << est.lanczos_steps << " steps]:\n";
    if (!est.valid || est.lambda_min <= 0.0) {
        std::cout << "    Ritz(min) = " << est.lambda_min << "\n"
                  << "    Ritz(max) = " << est.lambda_max << "\n"
                  << "    kappa_est = " << est.kappa << "\n"
                  << "    (Note: M^{-1}A is non-symmetric; bounds are approximate)\n";
    } else {
        std::cout << "    lambda_min = " << est.lambda_min << "\n"
                  << "    lambda_max = " << est.lambda_max << "\n"
                  << "    kappa      = " << est.kappa << "\n";
    }
}

SpectralResult estimate_spectral_radius(const CSRMatrix& A, const Preconditioner* M,
                                        int max_iter, double tol, unsigned int seed) {
    sg_context* ctx = default_context();
    DevCsr dA(A);
    PrecondHandle ph(M, A.nrows);
    const int n = A.nrows;
    Dev dv(sizeof(double) * std::max(n, 1));
    sg_spectral_result r{};
    check(sg_estimate_spectral_radius(dA.h, ph.h, max_iter, tol, seed, &r, dv.d()), ctx);
    SpectralResult out;
    out.spectral_radius = r.spectral_radius;
    out.convergence_ok = r.convergence_ok != 0;
    out.power_iters = r.power_iters;
    out.eigenvector.resize(n);
    check(sg_memcpy_d2h(ctx, out.eigenvector.data(), dv.p, sizeof(double) * n), ctx);
    return out;
}

void print_spectral_result(const SpectralResult& sr, const std::string& label) {
    std::cout << "  Spectral radius" << (label.empty() ? "" : " (" + label + ")")
              << " [power iter " << sr.power_iters << " steps]:\n"
              << "    rho(I - M^{-1}A) = " << sr.spectral_radius << "\n"
              << "    Convergence cond : "
              << (sr.convergence_ok ? "SATISFIED (rho < 1)" : "VIOLATED (rho >= 1)") << "\n";
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
