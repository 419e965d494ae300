// Prints the product's schwarz::Metrics::to_json for the three records of
// tests/golden/metrics.json (written by the reference's own Metrics::to_json,
// src/io/metrics_logger.h:28-51), one per line, then appends the first record
// twice with write_jsonl to the file named by argv[1].  Host code only: it
// needs no GPU.
#include <cstdio>

#include "schwarz/schwarz.h"

using schwarz::Metrics;

int main(int argc, char** argv) {
    Metrics a;
    a.matrix = "poisson2d_256";
    a.method = "GMRES+ilu0";
    a.n = 65536;
    a.nnz = 326656;
    a.threads = 8;
    a.overlap = 1;
    a.iterations = 412;
    a.tol = 1e-8;
    a.residual_norm = 9.87654321e-9;
    a.time_setup_s = 0.0123456789;
    a.time_solve_s = 1.5;
    Metrics b;
    Metrics c;
    c.matrix = "poisson3d_128";
    c.method = "GPU_GMRES+ras";
    c.n = 2097152;
    c.nnz = 14581760;
    c.gpus = 1;
    c.restart = 50;
    c.tol = 0.000123;
    c.residual_norm = 0.0;
    c.time_solve_s = 123456.5;
    c.speedup = 3.25;
    c.efficiency = 1e-7;
    std::printf("%s\n%s\n%s\n", a.to_json().c_str(), b.to_json().c_str(), c.to_json().c_str());
    if (argc > 1) {
        a.write_jsonl(argv[1]);
        a.write_jsonl(argv[1]);
    }
    return 0;
}
