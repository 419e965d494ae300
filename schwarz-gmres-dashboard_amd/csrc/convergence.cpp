// Residual-history metrics of one solve (host code): ConvergenceMetrics of
// src/analysis/convergence_analysis.h:15-106.
//   rates[k-1] = (r_k / r_0)^(1/k) (0 when the ratio is not positive, r_0
//   shifted by 1e-300), asymptotic rate = mean of the last min(20, n) rates
//   (the last rate when fewer than two), stagnation = the first window of 20
//   steps whose relative reduction is below 1e-3; the JSONL line and number
//   formats (six significant digits, %g style) as the reference writes them.
#include "../../include/sg_capi.h"
#include "host_setup.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sg {

namespace {

constexpr int kWindow = 20;
constexpr double kStagTol = 1e-3;

void analyze(const double* h, int n, int converged, sg_convergence* out, double* rates) {
    *out = sg_convergence{};
    out->stagnation_start = -1;
    out->converged = converged ? 1 : 0;
    out->iterations = n;
    if (n <= 0) return;
    out->initial_residual = h[0];
    out->final_residual = h[n - 1];
    std::vector<double> r(n);
    for (int k = 1; k <= n; ++k) {
        const double ratio = h[k - 1] / (h[0] + 1e-300);
        r[k - 1] = ratio > 0.0 ? std::pow(ratio, 1.0 / k) : 0.0;
    }
    const int tail = std::min(kWindow, n);
    if (tail >= 2) {
        double s = 0.0;
        for (int k = n - tail; k < n; ++k) s += r[k];
        out->asymptotic_rate = s / tail;
    } else {
        out->asymptotic_rate = r[n - 1];
    }
    for (int i = kWindow; i < n; ++i) {
        const double a = h[i - kWindow], b = h[i];
        if ((a - b) / (a + 1e-300) < kStagTol) {
            out->stagnation_start = i - kWindow + 1;
            break;
        }
    }
    if (rates) std::copy(r.begin(), r.end(), rates);
}

}  // namespace

}  // namespace sg

using namespace sg;

extern "C" sg_status sg_convergence_analyze(const double* hist, int n, int converged,
                                            sg_convergence* out, double* rates) {
    if (!out || n < 0 || (n > 0 && !hist)) return SG_E_ARG;
    analyze(hist, n, converged, out, rates);
    return SG_OK;
}

extern "C" sg_status sg_convergence_write_jsonl(const char* path, const char* matrix,
                                                const char* precond, int nparts, int overlap,
                                                const double* hist, int n, int converged) {
    if (!path || !matrix || !precond || n < 0 || (n > 0 && !hist)) return SG_E_ARG;
    sg_convergence c;
    analyze(hist, n, converged, &c, nullptr);
    std::FILE* f = std::fopen(path, "a");
    if (!f) {
        set_global_error(std::string("Cannot open file for appending: ") + path);
        return SG_E_IO;
    }
    std::unique_ptr<std::FILE, int (*)(std::FILE*)> guard(f, &std::fclose);
    std::fprintf(f,
                 "{\"type\":\"convergence\",\"matrix\":\"%s\",\"precond\":\"%s\",\"nparts\":%d,"
                 "\"overlap\":%d,\"iterations\":%d,\"converged\":%s,\"initial_residual\":%.6g,"
                 "\"final_residual\":%.6g,\"asymptotic_rate\":%.6g,\"stagnation_start\":%d,"
                 "\"residual_history\":[",
                 matrix, precond, nparts, overlap, c.iterations, c.converged ? "true" : "false",
                 c.initial_residual, c.final_residual, c.asymptotic_rate, c.stagnation_start);
    for (int i = 0; i < n; ++i) std::fprintf(f, "%s%.6g", i ? "," : "", hist[i]);
    std::fprintf(f, "]}\n");
    if (std::ferror(f)) {
        set_global_error(std::string("write failed: ") + path);
        return SG_E_IO;
    }
    return SG_OK;
}
