// CPU oracle: restatement of the reference Schwarz-(F)GMRES path.
// TEST INFRASTRUCTURE ONLY -- see oracle.h for the rules and pinning status.
// Build: g++ -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <set>
#include <sstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct Csr {
    int n = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

Csr make_csr(int n, const int* rp, const int* ci, const double* v) {
    Csr a;
    a.n = n;
    a.rp.assign(rp, rp + n + 1);
    a.ci.assign(ci, ci + rp[n]);
    a.v.assign(v, v + rp[n]);
    return a;
}

// src/core/sparse_matrix.h:30-37 -- row-sequential sum in CSR order.
void spmv(const Csr& a, const double* x, double* y) {
    for (int i = 0; i < a.n; ++i) {
        double s = 0.0;
        for (int j = a.rp[i]; j < a.rp[i + 1]; ++j) s += a.v[j] * x[a.ci[j]];
        y[i] = s;
    }
}

// src/core/vector.h:11-29
double dot(const double* a, const double* b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
double norm2(const double* a, int n) { return std::sqrt(dot(a, a, n)); }
// src/core/vector.h:31-35
void axpy(double alpha, const double* x, double* y, int n) {
    for (int i = 0; i < n; ++i) y[i] += alpha * x[i];
}

// src/core/sparse_matrix.h:53-83 -- local CSR of A[idx, idx], CSR order kept.
Csr extract(const Csr& a, const std::vector<int>& idx) {
    int m = (int)idx.size();
    std::vector<int> inv(a.n, -1);
    for (int i = 0; i < m; ++i) inv[idx[i]] = i;
    Csr s;
    s.n = m;
    s.rp.assign(m + 1, 0);
    for (int li = 0; li < m; ++li) {
        int gi = idx[li];
        s.rp[li] = (int)s.ci.size();
        for (int j = a.rp[gi]; j < a.rp[gi + 1]; ++j) {
            int lj = inv[a.ci[j]];
            if (lj >= 0) {
                s.ci.push_back(lj);
                s.v.push_back(a.v[j]);
            }
        }
    }
    s.rp[m] = (int)s.ci.size();
    return s;
}

// src/precond/ilu0.h:13-54 -- IKJ ILU(0) in place on A's pattern.
struct Ilu0 {
    Csr lu;
    std::vector<int> diag;
    std::vector<float> lu32;  // FP32 copy for the mixed-precision mode

    void factor(const Csr& a) {
        lu = a;
        int n = a.n;
        diag.assign(n, -1);
        for (int i = 0; i < n; ++i)
            for (int j = lu.rp[i]; j < lu.rp[i + 1]; ++j)
                if (lu.ci[j] == i) { diag[i] = j; break; }
        for (int i = 1; i < n; ++i) {
            for (int jp = lu.rp[i]; jp < lu.rp[i + 1]; ++jp) {
                int k = lu.ci[jp];
                if (k >= i) break;
                int dk = diag[k];
                if (dk < 0 || std::abs(lu.v[dk]) < 1e-30) continue;  // ilu0.h:36-37
                double aik = lu.v[jp] / lu.v[dk];
                lu.v[jp] = aik;
                for (int kp = dk + 1; kp < lu.rp[k + 1]; ++kp) {
                    int colk = lu.ci[kp];
                    for (int ip = lu.rp[i]; ip < lu.rp[i + 1]; ++ip) {
                        if (lu.ci[ip] == colk) { lu.v[ip] -= aik * lu.v[kp]; break; }
                    }
                }
            }
        }
        lu32.resize(lu.v.size());
        for (size_t k = 0; k < lu.v.size(); ++k) lu32[k] = (float)lu.v[k];
    }

    // src/precond/ilu0.h:57-79 -- unit-L forward, then U backward with division.
    void solve(const double* x, double* y) const {
        int n = lu.n;
        std::vector<double> t(x, x + n);
        for (int i = 0; i < n; ++i)
            for (int jp = lu.rp[i]; jp < lu.rp[i + 1]; ++jp) {
                int c = lu.ci[jp];
                if (c >= i) break;
                t[i] -= lu.v[jp] * t[c];
            }
        for (int i = n - 1; i >= 0; --i) {
            for (int jp = diag[i] + 1; jp < lu.rp[i + 1]; ++jp) t[i] -= lu.v[jp] * t[lu.ci[jp]];
            t[i] /= lu.v[diag[i]];
        }
        std::copy(t.begin(), t.end(), y);
    }

    // Same sweep order in FP32 arithmetic on FP32-rounded factors (mixed mode).
    void solve32(float* t) const {
        int n = lu.n;
        for (int i = 0; i < n; ++i)
            for (int jp = lu.rp[i]; jp < lu.rp[i + 1]; ++jp) {
                int c = lu.ci[jp];
                if (c >= i) break;
                t[i] -= lu32[jp] * t[c];
            }
        for (int i = n - 1; i >= 0; --i) {
            for (int jp = diag[i] + 1; jp < lu.rp[i + 1]; ++jp) t[i] -= lu32[jp] * t[lu.ci[jp]];
            t[i] /= lu32[diag[i]];
        }
    }
};

// src/decomp/domain_decomp.h:61-93 given part[] (the METIS call :48-52 is
// replaced by the caller's partition).  Overlap: delta BFS layers over CSR
// columns, explicit zeros included; sets are sorted ascending.
struct Decomp {
    int nparts = 0;
    std::vector<int> part;
    std::vector<std::vector<int>> owned, dofs;
};

Decomp decompose(const Csr& a, const int* part, int nparts, int overlap) {
    Decomp dd;
    dd.nparts = nparts;
    if (nparts <= 1 || part == nullptr) dd.part.assign(a.n, 0);  // domain_decomp.h:24-26
    else dd.part.assign(part, part + a.n);
    dd.owned.resize(nparts);
    for (int i = 0; i < a.n; ++i) dd.owned[dd.part[i]].push_back(i);
    dd.dofs.resize(nparts);
    for (int s = 0; s < nparts; ++s) {
        std::set<int> cur(dd.owned[s].begin(), dd.owned[s].end());
        for (int layer = 0; layer < overlap; ++layer) {
            std::set<int> add;
            for (int idx : cur)
                for (int j = a.rp[idx]; j < a.rp[idx + 1]; ++j)
                    if (!cur.count(a.ci[j])) add.insert(a.ci[j]);
            cur.insert(add.begin(), add.end());
        }
        dd.dofs[s].assign(cur.begin(), cur.end());
    }
    return dd;
}

// src/precond/coarse_solve.h:15-47 -- right-looking LU, partial pivoting on the
// first maximum |.|, whole-row swaps, singular if max < 1e-30.
bool lu_factor(int n, std::vector<double>& LU, std::vector<int>& piv) {
    piv.assign(n, 0);
    for (int k = 0; k < n; ++k) {
        int mr = k;
        double mv = std::abs(LU[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double t = std::abs(LU[(size_t)i * n + k]);
            if (t > mv) { mv = t; mr = i; }
        }
        piv[k] = mr;
        if (mv < 1e-30) return false;
        if (mr != k)
            for (int j = 0; j < n; ++j) std::swap(LU[(size_t)k * n + j], LU[(size_t)mr * n + j]);
        for (int i = k + 1; i < n; ++i) {
            LU[(size_t)i * n + k] /= LU[(size_t)k * n + k];
            for (int j = k + 1; j < n; ++j)
                LU[(size_t)i * n + j] -= LU[(size_t)i * n + k] * LU[(size_t)k * n + j];
        }
    }
    return true;
}

// src/precond/coarse_solve.h:50-67
template <typename T>
void lu_solve(int n, const T* LU, const int* piv, const T* b, T* x) {
    std::copy(b, b + n, x);
    for (int k = 0; k < n; ++k) {
        if (piv[k] != k) std::swap(x[k], x[piv[k]]);
        for (int i = k + 1; i < n; ++i) x[i] -= LU[(size_t)i * n + k] * x[k];
    }
    for (int k = n - 1; k >= 0; --k) {
        for (int j = k + 1; j < n; ++j) x[k] -= LU[(size_t)k * n + j] * x[j];
        x[k] /= LU[(size_t)k * n + k];
    }
}

struct Precond {
    virtual ~Precond() = default;
    virtual void apply(const double* x, double* y) const = 0;
    virtual int coarse(double*, double*, int*) const { return 0; }
};

// src/precond/jacobi.h:12-25
struct Jacobi : Precond {
    std::vector<double> inv;
    explicit Jacobi(const Csr& a) {
        inv.resize(a.n);
        for (int i = 0; i < a.n; ++i) {
            double d = 0.0;
            for (int j = a.rp[i]; j < a.rp[i + 1]; ++j)
                if (a.ci[j] == i) { d = a.v[j]; break; }  // sparse_matrix.h:45-51
            inv[i] = (std::abs(d) > 1e-15) ? 1.0 / d : 1.0;
        }
    }
    void apply(const double* x, double* y) const override {
        for (size_t i = 0; i < inv.size(); ++i) y[i] = inv[i] * x[i];
    }
};

struct Ilu0Pc : Precond {
    Ilu0 f;
    int precision;
    Ilu0Pc(const Csr& a, int prec) : precision(prec) { f.factor(a); }
    void apply(const double* x, double* y) const override {
        if (precision == 0) { f.solve(x, y); return; }
        std::vector<float> t(f.lu.n);
        for (int i = 0; i < f.lu.n; ++i) t[i] = (float)x[i];
        f.solve32(t.data());
        for (int i = 0; i < f.lu.n; ++i) y[i] = (double)t[i];
    }
};

// ILUT: src/precond/ilut.h:28-172, L and U kept as separate row lists as the
// reference stores them.  Elimination visits the row's sorted pattern as it
// stood before elimination (the reference's range-for, ilut.h:62); fill joins
// the list only for the later split.  precision 1 runs both sweeps in FP32.
struct IlutPc : Precond {
    int n = 0, precision = 0;
    std::vector<std::vector<std::pair<int, double>>> L, U;
    std::vector<double> ud;
    long long lnz = 0, unz = 0;
    size_t max_list = 0;  // longest index list of a row (the reference reserves 2 * lim)
    IlutPc(const Csr& a, double drop, int fill, int prec) : n(a.n), precision(prec) {
        if (n == 0) return;
        const int lim = fill * std::max(1, a.rp[n] / n);
        L.resize(n);
        U.resize(n);
        std::vector<double> w(n, 0.0);
        std::vector<int> idx;
        using P = std::pair<int, double>;
        auto trim = [lim](std::vector<P>& v) {
            if ((int)v.size() <= lim) return;
            std::partial_sort(v.begin(), v.begin() + lim, v.end(), [](const P& x, const P& y) {
                return std::abs(x.second) > std::abs(y.second);
            });
            v.resize(lim);
            std::sort(v.begin(), v.end(), [](const P& x, const P& y) { return x.first < y.first; });
        };
        for (int i = 0; i < n; ++i) {
            for (int p = a.rp[i]; p < a.rp[i + 1]; ++p) {
                if (w[a.ci[p]] == 0.0) idx.push_back(a.ci[p]);
                w[a.ci[p]] += a.v[p];
            }
            double nrm = 0.0;
            for (int j : idx) nrm = std::max(nrm, std::abs(w[j]));
            const double t = drop * nrm;
            std::sort(idx.begin(), idx.end());
            const std::vector<int> pattern(idx);
            for (int k : pattern) {
                if (k >= i) break;
                if (std::abs(w[k]) < t || std::abs(ud[k]) < 1e-30) {
                    w[k] = 0.0;
                    continue;
                }
                w[k] = w[k] / ud[k];
                for (const P& e : U[k]) {
                    if (e.first <= k) continue;
                    if (w[e.first] == 0.0) idx.push_back(e.first);
                    w[e.first] -= w[k] * e.second;
                }
            }
            std::sort(idx.begin(), idx.end());
            max_list = std::max(max_list, idx.size());
            std::vector<P> lo, up;
            for (int j : idx) {
                const double x = w[j];
                if (std::abs(x) == 0.0) continue;
                if (j < i) {
                    if (std::abs(x) >= t) lo.emplace_back(j, x);
                } else if (j == i || std::abs(x) >= t) {
                    up.emplace_back(j, x);
                }
            }
            trim(lo);
            trim(up);
            double d = 0.0;
            for (const P& e : up)
                if (e.first == i) { d = e.second; break; }
            ud.push_back(std::abs(d) < 1e-30 ? 1e-30 : d);
            lnz += (long long)lo.size();
            unz += (long long)up.size();
            L[i] = std::move(lo);
            U[i] = std::move(up);
            for (int j : idx) w[j] = 0.0;
            idx.clear();
        }
    }
    template <typename T>
    void sweep(const T* x, T* y) const {
        std::vector<T> z(n);
        for (int i = 0; i < n; ++i) {  // ilut.h:156-161
            z[i] = x[i];
            for (const auto& e : L[i]) z[i] -= (T)e.second * z[e.first];
        }
        for (int i = n - 1; i >= 0; --i) {  // ilut.h:164-171: the first U entry is skipped
            y[i] = z[i];
            for (size_t p = 1; p < U[i].size(); ++p) y[i] -= (T)U[i][p].second * y[U[i][p].first];
            y[i] /= (T)ud[i];
        }
    }
    void apply(const double* x, double* y) const override {
        if (precision == 0) { sweep(x, y); return; }
        std::vector<float> a(x, x + n), b(n);
        sweep(a.data(), b.data());
        for (int i = 0; i < n; ++i) y[i] = b[i];
    }
};

// One-level Schwarz: src/precond/ras.h:21-70 (restricted) and
// src/precond/asm.h:20-62 (additive).  ASM accumulates in ascending subdomain
// order, which is the single-thread order of the reference's critical section.
struct Schwarz : Precond {
    int n, precision;
    bool restricted;
    Decomp dd;
    std::vector<Ilu0> local;
    std::vector<std::vector<int>> owned_local;  // ras.h:34-41

    Schwarz(const Csr& a, int nparts, int overlap, const int* part, bool ras, int prec)
        : n(a.n), precision(prec), restricted(ras) {
        dd = decompose(a, part, nparts, overlap);
        local.resize(nparts);
        owned_local.resize(nparts);
        for (int s = 0; s < nparts; ++s) {
            local[s].factor(extract(a, dd.dofs[s]));
            std::set<int> os(dd.owned[s].begin(), dd.owned[s].end());
            for (int i = 0; i < (int)dd.dofs[s].size(); ++i)
                if (os.count(dd.dofs[s][i])) owned_local[s].push_back(i);
        }
    }

    void local_solve(int s, const double* x, std::vector<double>& z) const {
        const auto& d = dd.dofs[s];
        int ns = (int)d.size();
        z.resize(ns);
        if (precision == 0) {
            std::vector<double> r(ns);
            for (int i = 0; i < ns; ++i) r[i] = x[d[i]];
            local[s].solve(r.data(), z.data());
        } else {
            std::vector<float> t(ns);
            for (int i = 0; i < ns; ++i) t[i] = (float)x[d[i]];
            local[s].solve32(t.data());
            for (int i = 0; i < ns; ++i) z[i] = (double)t[i];
        }
    }

    void apply(const double* x, double* y) const override {
        std::fill(y, y + n, 0.0);
        if (restricted) {
#pragma omp parallel for schedule(dynamic)
            for (int s = 0; s < dd.nparts; ++s) {
                std::vector<double> z;
                local_solve(s, x, z);
                for (int li : owned_local[s]) y[dd.dofs[s][li]] = z[li];
            }
        } else {
            std::vector<std::vector<double>> zs(dd.nparts);
#pragma omp parallel for schedule(dynamic)
            for (int s = 0; s < dd.nparts; ++s) local_solve(s, x, zs[s]);
            for (int s = 0; s < dd.nparts; ++s)
                for (size_t i = 0; i < dd.dofs[s].size(); ++i) y[dd.dofs[s][i]] += zs[s][i];
        }
    }
};

// Two-level additive Schwarz: src/precond/two_level_schwarz.h:29-64 with the
// piecewise-constant coarse space of src/precond/coarse_space.h:26-78.
// R0 is held sparsely: row s is c_s = 1/sqrt(|owned_s|) on owned_s.  Every sum
// below visits the nonzero terms in the same ascending order as the dense
// loops of coarse_space.h:49-59,63-70,73-78, so results equal the dense form.
struct TwoLevel : Precond {
    std::unique_ptr<Schwarz> fine;
    int n, nc, precision;
    std::vector<double> c;       // c_s
    std::vector<double> A0, LU;  // row-major nc x nc
    std::vector<int> piv;
    std::vector<float> LU32;

    TwoLevel(const Csr& a, int nparts, int overlap, const int* part, bool ras, int prec)
        : n(a.n), precision(prec) {
        fine.reset(new Schwarz(a, nparts, overlap, part, ras, prec));
        const Decomp& dd = fine->dd;
        nc = dd.nparts;
        c.resize(nc);
        for (int s = 0; s < nc; ++s) c[s] = 1.0 / std::sqrt((double)dd.owned[s].size());
        // A0[s][t] = sum_{k in owned_s asc} c_s * (A phi_t)[k],
        // (A phi_t)[k] = sum over row k in CSR order of v * c_t restricted to
        // columns owned by t (coarse_space.h:49,52-59).
        A0.assign((size_t)nc * nc, 0.0);
        std::vector<double> aphi(nc, 0.0);
        std::vector<int> touched;
        std::vector<char> seen(nc, 0);
        for (int s = 0; s < nc; ++s) {
            for (int k : dd.owned[s]) {
                touched.clear();
                for (int j = a.rp[k]; j < a.rp[k + 1]; ++j) {
                    int t = dd.part[a.ci[j]];
                    if (!seen[t]) { seen[t] = 1; touched.push_back(t); aphi[t] = 0.0; }
                    aphi[t] += a.v[j] * c[t];
                }
                for (int t : touched) {
                    A0[(size_t)s * nc + t] += c[s] * aphi[t];
                    seen[t] = 0;
                }
            }
        }
        LU = A0;
        if (!lu_factor(nc, LU, piv)) throw std::runtime_error("Coarse matrix is singular");
        LU32.resize(LU.size());
        for (size_t k = 0; k < LU.size(); ++k) LU32[k] = (float)LU[k];
    }

    void apply(const double* x, double* y) const override {
        const Decomp& dd = fine->dd;
        std::vector<double> zf(n);
        fine->apply(x, zf.data());
        if (precision == 0) {
            std::vector<double> r0(nc, 0.0), z0(nc);
            for (int s = 0; s < nc; ++s)
                for (int k : dd.owned[s]) r0[s] += c[s] * x[k];  // coarse_space.h:63-70
            lu_solve<double>(nc, LU.data(), piv.data(), r0.data(), z0.data());
            for (int i = 0; i < n; ++i) {
                int s = dd.part[i];
                double zc = 0.0 + c[s] * z0[s];  // coarse_space.h:73-78
                y[i] = zf[i] + zc;               // two_level_schwarz.h:62-63
            }
        } else {
            std::vector<float> r0(nc, 0.0f), z0(nc);
            for (int s = 0; s < nc; ++s) {
                float cs = (float)c[s];
                for (int k : dd.owned[s]) r0[s] += cs * (float)x[k];
            }
            lu_solve<float>(nc, LU32.data(), piv.data(), r0.data(), z0.data());
            for (int i = 0; i < n; ++i) {
                int s = dd.part[i];
                y[i] = (double)((float)zf[i] + (float)c[s] * z0[s]);
            }
        }
    }

    int coarse(double* a0, double* lu, int* pv) const override {
        if (a0) std::copy(A0.begin(), A0.end(), a0);
        if (lu) std::copy(LU.begin(), LU.end(), lu);
        if (pv) std::copy(piv.begin(), piv.end(), pv);
        return nc;
    }
};

// src/core/gmres.cpp:11-23 (tangent form) and :24-29 (application).
struct Givens {
    double c = 1.0, s = 0.0;
    void compute(double a, double b) {
        if (std::abs(b) < 1e-30) { c = 1.0; s = 0.0; }
        else if (std::abs(b) > std::abs(a)) {
            double t = -a / b;
            s = 1.0 / std::sqrt(1.0 + t * t);
            c = s * t;
        } else {
            double t = -b / a;
            c = 1.0 / std::sqrt(1.0 + t * t);
            s = c * t;
        }
    }
    void apply(double& h1, double& h2) const {
        double t1 = c * h1 - s * h2, t2 = s * h1 + c * h2;
        h1 = t1;
        h2 = t2;
    }
};

struct Result {
    int iterations = 0;
    double residual_norm = 0.0;
    bool converged = false;
    std::vector<double> hist;
};

// Left-preconditioned GMRES(m): src/core/gmres.cpp:34-160.
Result gmres_left(const Csr& A, const double* b, double* x, const Precond* M, int m,
                  int max_iter, double tol) {
    int n = A.n;
    Result res;
    std::vector<double> r(n), w(n), z(n), Av(n);
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(n));
    std::vector<double> H((size_t)(m + 1) * m), g(m + 1), y(m);
    std::vector<Givens> rot(m);
    int total = 0;
    for (int cycle = 0; cycle < max_iter; ++cycle) {
        spmv(A, x, r.data());
        for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
        if (M) M->apply(r.data(), z.data()); else z = r;
        double beta = norm2(z.data(), n);
        if (beta < tol) { res.converged = true; res.residual_norm = beta; break; }
        for (int i = 0; i < n; ++i) V[0][i] = z[i] / beta;
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j;
        for (j = 0; j < m; ++j) {
            spmv(A, V[j].data(), Av.data());
            if (M) M->apply(Av.data(), w.data()); else w = Av;
            for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt, gmres.cpp:93-97
                double h = dot(w.data(), V[i].data(), n);
                H[i + (size_t)j * (m + 1)] = h;
                axpy(-h, V[i].data(), w.data(), n);
            }
            double hn = norm2(w.data(), n);
            H[(j + 1) + (size_t)j * (m + 1)] = hn;
            if (hn < 1e-30) { j++; break; }  // gmres.cpp:102-105
            for (int i = 0; i < n; ++i) V[j + 1][i] = w[i] / hn;
            for (int i = 0; i < j; ++i) rot[i].apply(H[i + (size_t)j * (m + 1)], H[(i + 1) + (size_t)j * (m + 1)]);
            rot[j].compute(H[j + (size_t)j * (m + 1)], H[(j + 1) + (size_t)j * (m + 1)]);
            rot[j].apply(H[j + (size_t)j * (m + 1)], H[(j + 1) + (size_t)j * (m + 1)]);
            rot[j].apply(g[j], g[j + 1]);
            double rr = std::abs(g[j + 1]);
            res.hist.push_back(rr);
            total++;
            if (rr < tol) { j++; res.converged = true; break; }
        }
        int k = j;
        for (int i = k - 1; i >= 0; --i) {  // gmres.cpp:131-137
            y[i] = g[i];
            for (int l = i + 1; l < k; ++l) y[i] -= H[i + (size_t)l * (m + 1)] * y[l];
            y[i] /= H[i + (size_t)i * (m + 1)];
        }
        for (int i = 0; i < k; ++i) axpy(y[i], V[i].data(), x, n);
        if (res.converged) break;
        if (total >= max_iter) break;  // gmres.cpp:146-147
    }
    res.iterations = total;
    res.residual_norm = res.hist.empty() ? 0.0 : res.hist.back();  // gmres.cpp:150-155
    return res;
}

// Right-preconditioned flexible GMRES(m): src/core/fgmres.cpp:28-144.
Result fgmres(const Csr& A, const double* b, double* x, const Precond* M, int m, int max_iter,
              double tol) {
    int n = A.n;
    Result res;
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(n, 0.0));
    std::vector<std::vector<double>> Z(m, std::vector<double>(n, 0.0));
    std::vector<double> H((size_t)(m + 1) * m), g(m + 1), y(m);
    std::vector<Givens> rot(m);
    std::vector<double> r(n), w(n);
    int total = 0;
    for (int cycle = 0; cycle < max_iter; ++cycle) {
        spmv(A, x, r.data());
        for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
        double beta = norm2(r.data(), n);
        if (beta < tol) { res.converged = true; res.residual_norm = beta; break; }
        double ib = 1.0 / beta;  // fgmres.cpp:66-67 multiplies by the reciprocal
        for (int i = 0; i < n; ++i) V[0][i] = r[i] * ib;
        std::fill(H.begin(), H.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j;
        for (j = 0; j < m; ++j) {
            if (M) M->apply(V[j].data(), Z[j].data()); else Z[j] = V[j];
            spmv(A, Z[j].data(), w.data());
            for (int i = 0; i <= j; ++i) {
                double h = dot(w.data(), V[i].data(), n);
                H[i + (size_t)j * (m + 1)] = h;
                axpy(-h, V[i].data(), w.data(), n);
            }
            double hn = norm2(w.data(), n);
            H[(j + 1) + (size_t)j * (m + 1)] = hn;
            if (hn < 1e-30) { j++; break; }
            double ih = 1.0 / hn;
            for (int i = 0; i < n; ++i) V[j + 1][i] = w[i] * ih;
            for (int i = 0; i < j; ++i) rot[i].apply(H[i + (size_t)j * (m + 1)], H[(i + 1) + (size_t)j * (m + 1)]);
            rot[j].compute(H[j + (size_t)j * (m + 1)], H[(j + 1) + (size_t)j * (m + 1)]);
            rot[j].apply(H[j + (size_t)j * (m + 1)], H[(j + 1) + (size_t)j * (m + 1)]);
            rot[j].apply(g[j], g[j + 1]);
            double rr = std::abs(g[j + 1]);
            res.hist.push_back(rr);
            total++;
            if (rr < tol) { j++; res.converged = true; break; }
        }
        int k = j;
        for (int i = k - 1; i >= 0; --i) {
            y[i] = g[i];
            for (int l = i + 1; l < k; ++l) y[i] -= H[i + (size_t)l * (m + 1)] * y[l];
            y[i] /= H[i + (size_t)i * (m + 1)];
        }
        for (int i = 0; i < k; ++i) axpy(y[i], Z[i].data(), x, n);  // fgmres.cpp:130-131
        if (res.converged) break;
        if (total >= max_iter) break;
    }
    res.iterations = total;
    res.residual_norm = res.hist.empty() ? 0.0 : res.hist.back();
    return res;
}

// Right-preconditioned BiCGSTAB: src/core/bicgstab.h:57-204.  Relative tests
// against ||r0||; breakdown thresholds 1e-300; the half-step exit keeps
// x += alpha y and records ||s||.
Result bicgstab(const Csr& A, const double* b, double* x, const Precond* M, int max_iter,
                double tol) {
    const int n = A.n;
    Result res;
    std::vector<double> r(n), rh(n), p(n, 0.0), v(n, 0.0), y(n), z(n), s(n), t(n), Ax(n);
    spmv(A, x, Ax.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - Ax[i];  // bicgstab.h:78-79
    rh = r;
    const double r0 = norm2(r.data(), n);
    if (r0 < 1e-14) {  // bicgstab.h:85-90
        res.residual_norm = r0;
        res.converged = true;
        return res;
    }
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    for (int it = 1; it <= max_iter; ++it) {
        const double rho_prev = rho;
        rho = dot(rh.data(), r.data(), n);
        if (std::abs(rho) < 1e-300) {  // bicgstab.h:102-108
            res.iterations = it;
            res.residual_norm = norm2(r.data(), n);
            res.converged = false;
            break;
        }
        if (it == 1) {
            p = r;
        } else {
            const double beta = (rho / rho_prev) * (alpha / omega);
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        if (M) M->apply(p.data(), y.data());
        else y = p;
        spmv(A, y.data(), v.data());
        const double rv = dot(rh.data(), v.data(), n);
        if (std::abs(rv) < 1e-300) {  // bicgstab.h:127-132
            res.iterations = it;
            res.residual_norm = norm2(r.data(), n);
            res.converged = false;
            break;
        }
        alpha = rho / rv;
        for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        const double sn = norm2(s.data(), n);
        if (sn < tol * r0) {  // bicgstab.h:140-148
            for (int i = 0; i < n; ++i) x[i] += alpha * y[i];
            res.iterations = it;
            res.residual_norm = sn;
            res.converged = true;
            res.hist.push_back(sn);
            break;
        }
        if (M) M->apply(s.data(), z.data());
        else z = s;
        spmv(A, z.data(), t.data());
        const double tt = dot(t.data(), t.data(), n);
        if (tt < 1e-300) {  // bicgstab.h:158-165
            for (int i = 0; i < n; ++i) x[i] += alpha * y[i];
            res.iterations = it;
            res.residual_norm = sn;
            res.converged = sn < tol * r0;
            break;
        }
        omega = dot(t.data(), s.data(), n) / tt;
        for (int i = 0; i < n; ++i) x[i] += alpha * y[i] + omega * z[i];
        for (int i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
        const double rn = norm2(r.data(), n);
        res.hist.push_back(rn);
        res.iterations = it;
        res.residual_norm = rn;
        if (rn < tol * r0) {
            res.converged = true;
            break;
        }
        if (std::abs(omega) < 1e-300) break;  // bicgstab.h:187-193
    }
    return res;
}

// Unit start vector of the analysis routines: a 32-bit LCG (multiplier
// 1664525, increment 1013904223), each state read as a signed int / 2^31,
// normalised (src/analysis/condition_estimator.h:99-112,
// spectral_analysis.h:42-53).
std::vector<double> lcg_unit(int n, unsigned seed, bool guard_zero) {
    std::vector<double> v(n);
    unsigned st = seed;
    double nn = 0.0;
    for (int i = 0; i < n; ++i) {
        st = st * 1664525u + 1013904223u;
        v[i] = static_cast<double>(static_cast<int>(st)) / 2147483648.0;
        nn += v[i] * v[i];
    }
    nn = std::sqrt(nn);
    if (guard_zero && nn < 1e-14) nn = 1.0;
    for (int i = 0; i < n; ++i) v[i] /= nn;
    return v;
}

}  // namespace

extern "C" {

// src/mesh/poisson.h:10-38 -- neighbour order -nx, -1, diag, +1, +nx.
int or_poisson2d(int nx, int ny, int* rp, int* ci, double* v) {
    int nnz = 0;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            int idx = iy * nx + ix;
            if (rp) rp[idx] = nnz;
            auto put = [&](int c, double val) { if (ci) { ci[nnz] = c; v[nnz] = val; } ++nnz; };
            if (iy > 0) put(idx - nx, -1.0);
            if (ix > 0) put(idx - 1, -1.0);
            put(idx, 4.0);
            if (ix < nx - 1) put(idx + 1, -1.0);
            if (iy < ny - 1) put(idx + nx, -1.0);
        }
    if (rp) rp[nx * ny] = nnz;
    return nnz;
}

// src/mesh/poisson.h:41-74 -- neighbour order -nxy, -nx, -1, diag, +1, +nx, +nxy.
int or_poisson3d(int nx, int ny, int nz, int* rp, int* ci, double* v) {
    int nnz = 0, nxy = nx * ny;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                int idx = iz * nxy + iy * nx + ix;
                if (rp) rp[idx] = nnz;
                auto put = [&](int c, double val) { if (ci) { ci[nnz] = c; v[nnz] = val; } ++nnz; };
                if (iz > 0) put(idx - nxy, -1.0);
                if (iy > 0) put(idx - nx, -1.0);
                if (ix > 0) put(idx - 1, -1.0);
                put(idx, 6.0);
                if (ix < nx - 1) put(idx + 1, -1.0);
                if (iy < ny - 1) put(idx + nx, -1.0);
                if (iz < nz - 1) put(idx + nxy, -1.0);
            }
    if (rp) rp[nxy * nz] = nnz;
    return nnz;
}

void or_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) s += v[j] * x[ci[j]];
        y[i] = s;
    }
}

double or_dot(int n, const double* a, const double* b) { return dot(a, b, n); }

int or_extract(int n, const int* rp, const int* ci, const double* v, int m, const int* idx,
               int* sub_rp, int* sub_ci, double* sub_v) {
    Csr a = make_csr(n, rp, ci, v);
    Csr s = extract(a, std::vector<int>(idx, idx + m));
    if (sub_rp) std::copy(s.rp.begin(), s.rp.end(), sub_rp);
    if (sub_ci) std::copy(s.ci.begin(), s.ci.end(), sub_ci);
    if (sub_v) std::copy(s.v.begin(), s.v.end(), sub_v);
    return s.rp[m];
}

void or_ilu0_factor(int n, const int* rp, const int* ci, const double* a, double* lu, int* diag) {
    Ilu0 f;
    f.factor(make_csr(n, rp, ci, a));
    std::copy(f.lu.v.begin(), f.lu.v.end(), lu);
    std::copy(f.diag.begin(), f.diag.end(), diag);
}

void or_ilu0_solve(int n, const int* rp, const int* ci, const double* lu, const int* diag,
                   const double* x, double* y) {
    Ilu0 f;
    f.lu = make_csr(n, rp, ci, lu);
    f.diag.assign(diag, diag + n);
    f.solve(x, y);
}

int or_overlap(int n, const int* rp, const int* ci, const int* part, int nparts, int overlap,
               int* sub_ptr, int* sub_idx) {
    Csr a;
    a.n = n;
    a.rp.assign(rp, rp + n + 1);
    a.ci.assign(ci, ci + rp[n]);
    Decomp dd = decompose(a, part, nparts, overlap);
    int tot = 0;
    if (sub_ptr) sub_ptr[0] = 0;
    for (int s = 0; s < nparts; ++s) {
        for (int d : dd.dofs[s]) {
            if (sub_idx) sub_idx[tot] = d;
            ++tot;
        }
        if (sub_ptr) sub_ptr[s + 1] = tot;
    }
    return tot;
}

void* or_precond_new(int kind, int n, const int* rp, const int* ci, const double* v, int nparts,
                     int overlap, const int* part, int precision) {
    Csr a = make_csr(n, rp, ci, v);
    try {
        switch (kind) {
            case OR_PC_JACOBI: return new Jacobi(a);
            case OR_PC_ILU0: return new Ilu0Pc(a, precision);
            case OR_PC_ASM: return new Schwarz(a, nparts, overlap, part, false, precision);
            case OR_PC_RAS: return new Schwarz(a, nparts, overlap, part, true, precision);
            case OR_PC_TWOLEVEL_ASM: return new TwoLevel(a, nparts, overlap, part, false, precision);
            case OR_PC_TWOLEVEL_RAS: return new TwoLevel(a, nparts, overlap, part, true, precision);
            default: return nullptr;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "oracle: %s\n", e.what());
        return nullptr;
    }
}

void* or_ilut_new(int n, const int* rp, const int* ci, const double* v, double drop_tol,
                  int fill_factor, int precision, long long* l_nnz, long long* u_nnz) {
    IlutPc* p = new IlutPc(make_csr(n, rp, ci, v), drop_tol, fill_factor, precision);
    if (l_nnz) *l_nnz = p->lnz;
    if (u_nnz) *u_nnz = p->unz;
    return p;
}

long long or_ilut_max_list(void* h) { return (long long)static_cast<IlutPc*>(h)->max_list; }

void or_precond_apply(void* h, const double* x, double* y) {
    static_cast<Precond*>(h)->apply(x, y);
}

int or_precond_coarse(void* h, double* A0, double* LU, int* piv) {
    return static_cast<Precond*>(h)->coarse(A0, LU, piv);
}

void or_precond_free(void* h) { delete static_cast<Precond*>(h); }

int or_lu_factor(int n, double* A, int* piv) {
    std::vector<double> LU(A, A + (size_t)n * n);
    std::vector<int> p;
    bool ok = lu_factor(n, LU, p);
    std::copy(LU.begin(), LU.end(), A);
    std::copy(p.begin(), p.end(), piv);
    return ok ? 0 : -1;
}

void or_lu_solve(int n, const double* LU, const int* piv, const double* b, double* x) {
    lu_solve<double>(n, LU, piv, b, x);
}

int or_solve(int method, int n, const int* rp, const int* ci, const double* v, const double* b,
             double* x, void* pc, int restart, int max_iter, double tol, double* residual_norm,
             int* converged, double* hist, int hist_cap, int* hist_len) {
    Csr a = make_csr(n, rp, ci, v);
    const Precond* M = static_cast<const Precond*>(pc);
    Result r = method == 1 ? fgmres(a, b, x, M, restart, max_iter, tol)
                           : gmres_left(a, b, x, M, restart, max_iter, tol);
    if (residual_norm) *residual_norm = r.residual_norm;
    if (converged) *converged = r.converged ? 1 : 0;
    int hl = (int)r.hist.size();
    if (hist) for (int i = 0; i < hl && i < hist_cap; ++i) hist[i] = r.hist[i];
    if (hist_len) *hist_len = hl;
    return r.iterations;
}

// src/io/metrics_logger.h:28-51 -- setprecision(6); tol default notation,
// residual_norm scientific, then fixed for the remaining doubles.
int or_metrics_json(const char* matrix, const char* method, int n, int nnz, int ranks,
                    int threads, int gpus, int overlap, int restart, int iterations, double tol,
                    double residual_norm, double time_setup_s, double time_solve_s,
                    double speedup, double efficiency, char* out, int cap) {
    std::ostringstream os;
    os << std::setprecision(6) << "{"
       << "\"matrix\":\"" << matrix << "\","
       << "\"method\":\"" << method << "\","
       << "\"n\":" << n << ",\"nnz\":" << nnz << ",\"ranks\":" << ranks
       << ",\"threads\":" << threads << ",\"gpus\":" << gpus << ",\"overlap\":" << overlap
       << ",\"restart\":" << restart << ",\"iterations\":" << iterations
       << ",\"tol\":" << tol
       << ",\"residual_norm\":" << std::scientific << residual_norm << "," << std::fixed
       << "\"time_setup_s\":" << time_setup_s << ",\"time_solve_s\":" << time_solve_s
       << ",\"speedup\":" << speedup << ",\"efficiency\":" << efficiency << "}";
    std::string s = os.str();
    if ((int)s.size() + 1 > cap) return -(int)s.size() - 1;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return (int)s.size();
}


// src/core/bicgstab.h:57-204; returns iterations.
int or_bicgstab(int n, const int* rp, const int* ci, const double* v, const double* b, double* x,
                void* pc, int max_iter, double tol, double* residual_norm, int* converged,
                double* hist, int hist_cap, int* hist_len) {
    Csr a = make_csr(n, rp, ci, v);
    Result r = bicgstab(a, b, x, static_cast<const Precond*>(pc), max_iter, tol);
    if (residual_norm) *residual_norm = r.residual_norm;
    if (converged) *converged = r.converged ? 1 : 0;
    const int hl = (int)r.hist.size();
    if (hist) for (int i = 0; i < hl && i < hist_cap; ++i) hist[i] = r.hist[i];
    if (hist_len) *hist_len = hl;
    return r.iterations;
}

// Lanczos recurrence of estimate_condition (src/analysis/condition_estimator.h:94-151):
// alpha[m], beta[m-1] of the tridiagonal; returns m (0 when k < 2).
int or_lanczos(int n, const int* rp, const int* ci, const double* v, void* pc, int k,
               unsigned seed, double* alpha, double* beta) {
    Csr a = make_csr(n, rp, ci, v);
    const Precond* M = static_cast<const Precond*>(pc);
    k = std::min(k, n);
    if (k < 2) return 0;
    std::vector<double> vc = lcg_unit(n, seed, true), vp(n, 0.0), w(n), z(n);
    double bprev = 0.0;
    int m = 0, nb = 0;
    for (int j = 0; j < k; ++j) {
        spmv(a, vc.data(), w.data());
        if (M) M->apply(w.data(), z.data());
        else z = w;
        const double al = dot(vc.data(), z.data(), n);
        alpha[m++] = al;
        for (int i = 0; i < n; ++i) z[i] -= al * vc[i] + bprev * vp[i];
        const double bn = norm2(z.data(), n);
        if (bn < 1e-10 * (1.0 + std::abs(al))) break;  // invariant subspace
        if (j < k - 1) {
            beta[nb++] = bn;
            vp = vc;
            bprev = bn;
            const double inv = 1.0 / bn;
            for (int i = 0; i < n; ++i) vc[i] = z[i] * inv;
        }
    }
    return m;
}

// Eigenvalues of the symmetric tridiagonal (d[n], e[n-1]) by implicit QL with
// Wilkinson shifts, sorted ascending (condition_estimator.h:43-88, the
// textbook tqli iteration with its at-most-100n sweeps per eigenvalue).
void or_tridiag_eigs(int n, const double* d_in, const double* e_in, double* out) {
    std::vector<double> d(d_in, d_in + n), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
    const int cap = 100 * n;
    for (int l = 0; l < n; ++l) {
        int m, sweeps = 0;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::abs(d[m]) + std::abs(d[m + 1]);
                if (std::abs(e[m]) <= 1e-14 * dd) break;
            }
            if (m == l) break;
            if (++sweeps > cap) break;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double sn = 1.0, cs = 1.0, p = 0.0;
            for (int i = m - 1; i >= l; --i) {
                const double f = sn * e[i], b = cs * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    break;
                }
                sn = f / r;
                cs = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * sn + 2.0 * cs * b;
                p = sn * r;
                d[i + 1] = g + p;
                g = cs * r - b;
            }
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        } while (m != l);
    }
    std::sort(d.begin(), d.end());
    for (int i = 0; i < n; ++i) out[i] = d[i];
}

// estimate_condition (condition_estimator.h:94-191): out = {lambda_min,
// lambda_max, kappa}; returns the Lanczos steps, and *valid.
int or_condition(int n, const int* rp, const int* ci, const double* v, void* pc, int k,
                 unsigned seed, double* out, int* valid) {
    out[0] = out[1] = out[2] = 0.0;
    *valid = 0;
    const int kk = std::max(2, std::min(k, n));
    std::vector<double> al(kk), be(kk);
    const int m = or_lanczos(n, rp, ci, v, pc, k, seed, al.data(), be.data());
    if (m < 2) return 0;
    for (int i = 0; i < m; ++i) if (!std::isfinite(al[i])) return 0;
    for (int i = 0; i + 1 < m; ++i) if (!std::isfinite(be[i])) return 0;
    std::vector<double> ev(m), fin;
    or_tridiag_eigs(m, al.data(), be.data(), ev.data());
    for (double x : ev) if (std::isfinite(x)) fin.push_back(x);
    if (fin.empty()) return 0;
    const double lmin = fin.front(), lmax = fin.back();
    out[0] = lmin;
    out[1] = lmax;
    out[2] = std::abs(lmin) > 1e-14 ? lmax / lmin : 1e30;
    *valid = lmin > 0.0;
    return m;
}

// estimate_spectral_radius (src/analysis/spectral_analysis.h:33-95): power
// iteration on E = I - M^{-1}A; reports |mu| of the step before the stopping
// one, as the reference does.  Returns the iterations; *rho the radius.
int or_spectral(int n, const int* rp, const int* ci, const double* v, void* pc, int max_iter,
                double tol, unsigned seed, double* rho) {
    Csr a = make_csr(n, rp, ci, v);
    const Precond* M = static_cast<const Precond*>(pc);
    std::vector<double> x = lcg_unit(n, seed, false), w(n), z(n);
    double mu_prev = 0.0;
    int iters = 0;
    for (int it = 0; it < max_iter; ++it) {
        spmv(a, x.data(), w.data());
        if (M) M->apply(w.data(), z.data());
        else z = w;
        for (int i = 0; i < n; ++i) w[i] = x[i] - z[i];
        const double mu = dot(x.data(), w.data(), n);
        const double en = norm2(w.data(), n);
        iters = it + 1;
        if (en < 1e-14) break;
        for (int i = 0; i < n; ++i) x[i] = w[i] / en;
        if (std::abs(mu - mu_prev) < tol * (1.0 + std::abs(mu))) break;
        mu_prev = mu;
    }
    *rho = std::abs(mu_prev);
    return iters;
}

}  // extern "C"
