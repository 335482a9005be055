// Minimal CSR sparse matrix for the oracle's assembled operators.
//
// TEST INFRASTRUCTURE ONLY (CPU oracle). Plays the role of Julia's
// SparseArrays in the reference (src/ElasticitySplitNode.jl:44-60,163,203-211):
// triplet assembly that sums duplicates, products that keep structural
// entries, and an explicit dropzeros.
#pragma once
#include <cstdint>
#include <vector>

struct Csr {
    int64_t nrows = 0, ncols = 0;
    std::vector<int64_t> rowptr;  // nrows+1
    std::vector<int64_t> col;
    std::vector<double> val;
    int64_t nnz() const { return (int64_t)col.size(); }
};

struct Triplets {
    std::vector<int64_t> i, j;
    std::vector<double> v;
    void push(int64_t r, int64_t c, double x) { i.push_back(r); j.push_back(c); v.push_back(x); }
};

// sparse(I, J, V, m, n): duplicates summed, entries kept even when they sum to 0.
Csr csr_from_triplets(const Triplets& t, int64_t m, int64_t n);
Csr csr_identity(int64_t n);
Csr csr_add(const Csr& a, const Csr& b, double alpha = 1.0, double beta = 1.0);
Csr csr_mul(const Csr& a, const Csr& b);  // structural product (no zero dropping)
Csr csr_scale(const Csr& a, double s);
Csr csr_transpose(const Csr& a);
void csr_dropzeros(Csr& a);
void csr_spmv(const Csr& a, const double* x, double* y);  // y = a*x
