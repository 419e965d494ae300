// Host-side setup of the Schwarz preconditioners: graph partition, overlap
// sets, subdomain extraction, ILU(0) level schedules and the packed layouts
// the HIP kernels consume.  Plain C++ (built with g++), integer work only,
// except for copying matrix values.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sg {

struct HostCsr {
    int n = 0;      // rows
    int ncols = 0;  // columns
    std::vector<int> rp, ci;
    std::vector<double> v;
    int nnz() const { return rp.empty() ? 0 : rp[n]; }
};

// Balanced k-way partition of the graph of A (self loops ignored, pattern
// symmetrised).  Multilevel recursive bisection: heavy-edge coarsening, greedy
// graph-growing initial bisection and boundary Fiduccia-Mattheyses refinement.
// Deterministic.  Stands in for METIS_PartGraphKway
// (reference src/decomp/domain_decomp.h:23-58); nparts <= 1 gives all zeros as
// there (:24-26).  ufactor is the allowed imbalance in 1/1000 (30 = 1.03).
// weighted = true weighs edge (i,j) by max(|a_ij|, |a_ji|) so that cuts prefer
// weak couplings (coefficient-aware partitions for jump coefficients); the
// reference's METIS call passes no weights, which weighted = false reproduces.
std::vector<int> partition_graph(const HostCsr& A, int nparts, int ufactor, bool weighted = false);

// Owned and overlapped index sets exactly as build_decomposition
// (reference src/decomp/domain_decomp.h:61-93): owned_dofs[s] ascending,
// subdomain_dofs[s] = owned_dofs[s] grown by `overlap` breadth-first layers
// over the CSR columns of A (explicit zeros included), sorted ascending.
void build_sets(const HostCsr& A, const std::vector<int>& part, int nparts, int overlap,
                std::vector<std::vector<int>>& owned, std::vector<std::vector<int>>& dofs);

// Local CSR of A[idx, idx] in CSR order (reference
// src/core/sparse_matrix.h:53-83).
HostCsr extract_submatrix(const HostCsr& A, const std::vector<int>& idx);

// Position of the diagonal in each row, -1 when absent.
std::vector<int> diag_positions(const HostCsr& A);

// Level of each row for the forward (unit-L) sweep: 0 without strictly-lower
// entries, else 1 + max over them.  The IKJ ILU(0) factorisation has the same
// dependencies, so rows of one level can be factored in parallel.
std::vector<int> lower_levels(const HostCsr& A);
// Level of each row for the backward (U) sweep (dependencies on columns > row).
std::vector<int> upper_levels(const HostCsr& A);

// Dual-threshold incomplete LU as the reference's ILUTPrecond::setup
// (src/precond/ilut.h:28-149), returned as one factor matrix F whose row i is
// [L row i (unit-lower part, ascending columns)] [(i, u_ii)] [U row i without
// its first entry] with diag[i] the position of (i, u_ii): the reference's
// apply (ilut.h:152-172) is then a forward sweep over the entries left of
// diag[i] and a backward sweep over those right of it, divided by F[diag[i]]
// -- the shape the ILU(0) sweeps consume.  l_nnz / u_nnz are the sizes of the
// reference's L and U arrays.
void ilut_factor(const HostCsr& A, double drop_tol, int fill_factor, HostCsr& F,
                 std::vector<int>& diag, long long& l_nnz, long long& u_nnz);

// Matrix Market coordinate files (matrix_market.cpp); flags are SG_MM_*.
void mm_read(const std::string& path, int flags, HostCsr& A);
void mm_write(const std::string& path, int nrows, int ncols, const int* rp, const int* ci,
              const double* v, int digits);

// Message returned by sg_last_error(NULL) (context.hip).
void set_global_error(const std::string& msg);

}  // namespace sg
