// Matrix Market coordinate files <-> CSR (host code, plain C++).
//
// Default mode reproduces read_matrix_market (src/io/matrix_market.h:14-92)
// on every input the reference reads with defined behaviour:
//   * leading lines whose first character is '%' are comments; "symmetric" or
//     "pattern" anywhere in one of them sets that flag (:24-33), so a
//     skew-symmetric file is mirrored with the same sign;
//   * the first other line holds "nrows ncols nnz" (:35-39);
//   * each of the next nnz lines holds "row col [value]", 1-based; a missing
//     value reads as 1.0 and an unreadable one as 0.0, as the C++ stream
//     extraction leaves them (:48-57); further tokens are ignored;
//   * symmetric files add the mirror of every off-diagonal entry (:60-61);
//   * entries are ordered by (row, col) and duplicates are kept as separate
//     CSR entries -- the block labelled "Deduplicate" (:81-89) does not merge
//     them.  Among entries with the same (row, col) the reference's unstable
//     sort leaves the order unspecified; here they keep file order.
// Where the reference's behaviour is undefined (an unreadable size line or
// index, an index out of range, e.g. from a blank or comment line among the
// entries) this reader reports an error instead.
//
// SG_MM_SUM_DUPLICATES merges entries with the same (row, col) by summing them
// in file order.  SG_MM_STRICT reads the header flags from the
// %%MatrixMarket banner only (skew-symmetric mirrored with the opposite sign,
// hermitian as symmetric, array and complex files rejected), skips blank and
// comment lines anywhere, and rejects unreadable values and extra entries.
//
// Numbers are parsed as the C++ stream extraction parses them: integers as
// [sign]digits; values by taking the longest [sign]digits[.digits][(e|E)
// [sign]digits] prefix and converting it with strtod (correctly rounded), an
// overflow giving +-DBL_MAX.
#include "../../include/sg_capi.h"
#include "host_setup.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sg {

namespace {

struct Cursor {
    const char* p;
    const char* end;  // end of the line
};

void skip_ws(Cursor& c) {
    while (c.p < c.end && std::isspace((unsigned char)*c.p)) ++c.p;
}

// [sign]digits; false when no digit follows (or on overflow).
bool read_int(Cursor& c, long long& out) {
    skip_ws(c);
    const char* q = c.p;
    bool neg = false;
    if (q < c.end && (*q == '+' || *q == '-')) neg = *q++ == '-';
    if (q >= c.end || !std::isdigit((unsigned char)*q)) return false;
    long long v = 0;
    while (q < c.end && std::isdigit((unsigned char)*q)) {
        v = v * 10 + (*q++ - '0');
        if (v > (long long)INT_MAX + 1) return false;
    }
    c.p = q;
    out = neg ? -v : v;
    return true;
}

enum ValStatus { kValOk, kValMissing, kValBad };

ValStatus read_value(Cursor& c, double& out) {
    skip_ws(c);
    if (c.p >= c.end) return kValMissing;
    const char* q = c.p;
    char buf[128];
    int k = 0;
    auto put = [&](char ch) {
        if (k < (int)sizeof(buf) - 1) buf[k++] = ch;
    };
    if (q < c.end && (*q == '+' || *q == '-')) put(*q++);
    bool digits = false;
    while (q < c.end && std::isdigit((unsigned char)*q)) put(*q++), digits = true;
    if (q < c.end && *q == '.') {
        put(*q++);
        while (q < c.end && std::isdigit((unsigned char)*q)) put(*q++), digits = true;
    }
    if (digits && q < c.end && (*q == 'e' || *q == 'E')) {
        put(*q++);
        if (q < c.end && (*q == '+' || *q == '-')) put(*q++);
        while (q < c.end && std::isdigit((unsigned char)*q)) put(*q++);
    }
    buf[k] = '\0';
    c.p = q;
    char* stop = nullptr;
    const double v = std::strtod(buf, &stop);
    if (k == 0 || stop != buf + k) {  // no number: the stream stores 0
        out = 0.0;
        return kValBad;
    }
    if (std::isinf(v)) {  // overflow: the stream stores +-DBL_MAX
        out = v > 0 ? DBL_MAX : -DBL_MAX;
        return kValBad;
    }
    out = v;
    return kValOk;
}

bool blank(const Cursor& c) {
    for (const char* q = c.p; q < c.end; ++q)
        if (!std::isspace((unsigned char)*q)) return false;
    return true;
}

std::string lower(std::string s) {
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

struct Entry {
    int row, col;
    double val;
};

}  // namespace

void mm_read(const std::string& path, int flags, HostCsr& A) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("Cannot open file: " + path);
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const bool strict = flags & SG_MM_STRICT;
    const bool sum_dup = flags & SG_MM_SUM_DUPLICATES;
    const char* p = text.data();
    const char* const fend = p + text.size();
    long long lineno = 0;
    auto next_line = [&](Cursor& c) {
        if (p >= fend) return false;
        const char* e = static_cast<const char*>(std::memchr(p, '\n', (size_t)(fend - p)));
        if (!e) e = fend;
        c.p = p;
        c.end = e;
        p = e < fend ? e + 1 : fend;
        ++lineno;
        return true;
    };
    auto where = [&]() { return path + ":" + std::to_string(lineno) + ": "; };

    bool symmetric = false, skew = false, pattern = false;
    Cursor c{nullptr, nullptr};
    bool have = false, first = true;
    while ((have = next_line(c))) {
        const bool comment = c.p < c.end && *c.p == '%';
        if (strict && first) {
            std::string ln(c.p, c.end);
            if (ln.compare(0, 14, "%%MatrixMarket") != 0)
                throw std::runtime_error(where() + "missing %%MatrixMarket banner");
            char o[64] = "", f[64] = "", t[64] = "", sy[64] = "";
            std::sscanf(ln.c_str() + 14, "%63s %63s %63s %63s", o, f, t, sy);
            const std::string obj = lower(o), fmt = lower(f), fld = lower(t), sym = lower(sy);
            if (obj != "matrix") throw std::runtime_error(where() + "object '" + obj + "' is not a matrix");
            if (fmt != "coordinate")
                throw std::runtime_error(where() + "format '" + fmt + "' is not coordinate");
            if (fld == "complex") throw std::runtime_error(where() + "complex matrices are not supported");
            if (fld != "real" && fld != "integer" && fld != "pattern" && fld != "double")
                throw std::runtime_error(where() + "unknown field '" + fld + "'");
            if (sym != "general" && sym != "symmetric" && sym != "skew-symmetric" && sym != "hermitian")
                throw std::runtime_error(where() + "unknown symmetry '" + sym + "'");
            pattern = fld == "pattern";
            symmetric = sym != "general";
            skew = sym == "skew-symmetric";
            first = false;
            continue;
        }
        first = false;
        if (comment) {
            if (!strict) {
                std::string ln(c.p, c.end);
                if (ln.find("symmetric") != std::string::npos) symmetric = true;
                if (ln.find("pattern") != std::string::npos) pattern = true;
            }
            continue;
        }
        if (strict && blank(c)) continue;
        break;
    }
    if (!have) throw std::runtime_error(path + ": no size line");
    long long nr = 0, nc = 0, nz = 0;
    if (!read_int(c, nr) || !read_int(c, nc) || !read_int(c, nz) || nr < 0 || nc < 0 || nz < 0)
        throw std::runtime_error(where() + "cannot read 'nrows ncols nnz'");
    if (strict && !blank(c)) throw std::runtime_error(where() + "extra tokens after the sizes");
    if (symmetric && nr != nc && strict)
        throw std::runtime_error(where() + "symmetric matrix must be square");

    std::vector<Entry> ent;
    ent.reserve((size_t)(symmetric ? 2 * nz : nz));
    for (long long k = 0; k < nz; ++k) {
        if (!next_line(c)) throw std::runtime_error("Premature end of file");
        if (strict && (blank(c) || *c.p == '%')) {
            --k;
            continue;
        }
        long long r, col;
        if (!read_int(c, r) || !read_int(c, col))
            throw std::runtime_error(where() + "cannot read the row and column of entry " +
                                     std::to_string(k + 1));
        double v = 1.0;  // kept when the value is missing
        if (!pattern) {
            const ValStatus vs = read_value(c, v);
            if (vs == kValBad && strict) throw std::runtime_error(where() + "cannot read the value");
            if (vs == kValMissing && strict) throw std::runtime_error(where() + "missing value");
        }
        if (r < 1 || r > nr || col < 1 || col > nc)
            throw std::runtime_error(where() + "entry (" + std::to_string(r) + ", " +
                                     std::to_string(col) + ") outside the " + std::to_string(nr) +
                                     " x " + std::to_string(nc) + " matrix");
        const int ri = (int)r - 1, ci = (int)col - 1;
        ent.push_back(Entry{ri, ci, v});
        if (symmetric && ri != ci) {
            if (ci >= nr || ri >= nc)
                throw std::runtime_error(where() + "mirrored entry outside the matrix");
            ent.push_back(Entry{ci, ri, skew && strict ? -v : v});
        }
    }
    if (strict)
        while (next_line(c))
            if (!blank(c) && *c.p != '%')
                throw std::runtime_error(where() + "more entries than the size line declares");

    // stable order by (row, col): counting sort by row, then a stable sort of
    // each row by column
    A = HostCsr{};
    A.n = (int)nr;
    A.ncols = (int)nc;
    A.rp.assign((size_t)nr + 1, 0);
    for (const Entry& e : ent) A.rp[e.row + 1]++;
    for (long long i = 0; i < nr; ++i) A.rp[i + 1] += A.rp[i];
    std::vector<int> pos(A.rp.begin(), A.rp.end() - 1);
    std::vector<std::pair<int, double>> cv(ent.size());
    for (const Entry& e : ent) cv[pos[e.row]++] = {e.col, e.val};
    ent.clear();
    ent.shrink_to_fit();
    for (long long i = 0; i < nr; ++i)
        std::stable_sort(cv.begin() + A.rp[i], cv.begin() + A.rp[i + 1],
                         [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
                             return a.first < b.first;
                         });
    A.ci.reserve(cv.size());
    A.v.reserve(cv.size());
    int out = 0;
    for (long long i = 0; i < nr; ++i) {
        const int b = A.rp[i], e = A.rp[i + 1];
        A.rp[i] = out;
        for (int k = b; k < e; ++k) {
            if (sum_dup && out > A.rp[i] && A.ci.back() == cv[k].first) {
                A.v.back() += cv[k].second;
                continue;
            }
            A.ci.push_back(cv[k].first);
            A.v.push_back(cv[k].second);
            ++out;
        }
    }
    A.rp[nr] = out;
}

// write_matrix_market (src/io/matrix_market.h:94-103): "general" banner, the
// size line, one "row col value" line per stored entry, 1-based, values with
// `digits` significant digits in %g style (6 = the reference's default stream
// precision; 17 round-trips).
void mm_write(const std::string& path, int nrows, int ncols, const int* rp, const int* ci,
              const double* v, int digits) {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("Cannot open file for writing: " + path);
    std::unique_ptr<std::FILE, int (*)(std::FILE*)> guard(f, &std::fclose);
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", nrows, ncols,
                 nrows ? rp[nrows] : 0);
    for (int i = 0; i < nrows; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            std::fprintf(f, "%d %d %.*g\n", i + 1, ci[k] + 1, digits, v[k]);
    if (std::ferror(f)) throw std::runtime_error("write failed: " + path);
}

}  // namespace sg

using namespace sg;

extern "C" sg_status sg_mm_read(const char* path, int flags, int* nrows, int* ncols, int* nnz,
                                int** row_ptr, int** col_idx, double** vals) {
    if (!path || !nrows || !ncols || !nnz || !row_ptr || !col_idx || !vals) return SG_E_ARG;
    *row_ptr = nullptr;
    *col_idx = nullptr;
    *vals = nullptr;
    try {
        HostCsr A;
        mm_read(path, flags, A);
        const size_t nz = A.ci.size();
        int* rp = static_cast<int*>(std::malloc(sizeof(int) * (A.rp.size())));
        int* ci = static_cast<int*>(std::malloc(sizeof(int) * std::max<size_t>(1, nz)));
        double* v = static_cast<double*>(std::malloc(sizeof(double) * std::max<size_t>(1, nz)));
        if (!rp || !ci || !v) {
            std::free(rp);
            std::free(ci);
            std::free(v);
            set_global_error("host out of memory");
            return SG_E_ARG;
        }
        std::copy(A.rp.begin(), A.rp.end(), rp);
        std::copy(A.ci.begin(), A.ci.end(), ci);
        std::copy(A.v.begin(), A.v.end(), v);
        *nrows = A.n;
        *ncols = A.ncols;
        *nnz = (int)nz;
        *row_ptr = rp;
        *col_idx = ci;
        *vals = v;
        return SG_OK;
    } catch (const std::exception& e) {
        set_global_error(e.what());
        return SG_E_IO;
    }
}

extern "C" void sg_mm_free(void* p) { std::free(p); }

extern "C" sg_status sg_mm_write(const char* path, int nrows, int ncols, const int* row_ptr,
                                 const int* col_idx, const double* vals, int digits) {
    if (!path || nrows < 0 || ncols < 0 || (nrows && !row_ptr) || digits < 1 || digits > 17)
        return SG_E_ARG;
    if (nrows && row_ptr[nrows] > 0 && (!col_idx || !vals)) return SG_E_ARG;
    try {
        mm_write(path, nrows, ncols, row_ptr, col_idx, vals, digits);
        return SG_OK;
    } catch (const std::exception& e) {
        set_global_error(e.what());
        return SG_E_IO;
    }
}
