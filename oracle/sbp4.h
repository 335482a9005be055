// SBP order-4 diagonal-norm operators on an equidistant 1D grid (oracle side).
//
// TEST INFRASTRUCTURE ONLY: this file belongs to the CPU oracle. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it, and only
// as the checker.
//
// The reference takes these stencils from Diffinitive's `standard_diagonal.toml`
// (order 4; call site src/BP8.jl:195, src/Elasticity.jl:46-47,198-203), which is an
// un-vendored git dependency (Diffinitive @ 92d842cf, v0.1.8) and is absent here.
// The values below restate the published Mattsson & Nordstrom (2004, JCP 199)
// order-4 diagonal-norm operators. They are pinned by the SBP identities
//   H D1 + D1^T H = diag(-1, 0, ..., 0, 1)
//   H D2 = -M - e_0 d_0^T + e_N d_N^T,  M = M^T >= 0
// and polynomial exactness (tests/test_oracle_sbp.py), and end to end by the
// reference's published stiffness scalars (tests/test_oracle_pins.py).
#pragma once
#include <cstddef>
#include <vector>

namespace sbp4 {

// Dense 1D operators, row-major n x n, already scaled by the spacing h.
struct Ops1D {
    int n = 0;
    double h = 0;
    std::vector<double> H;    // diagonal quadrature weights (n)
    std::vector<double> D1;   // first derivative (n*n)
    std::vector<double> D2;   // narrow second derivative (n*n)
    std::vector<double> dl;   // +x derivative at node 0, as a row (n)
    std::vector<double> dr;   // +x derivative at node n-1, as a row (n)
    double at(const std::vector<double>& M, int i, int j) const { return M[(size_t)i * n + j]; }
};

// Builds the operators; n must be >= 9 (two closure widths plus one point).
Ops1D build(int n, double h);

// Row-major dense product C = A*B of n x n matrices.
std::vector<double> matmul(const std::vector<double>& A, const std::vector<double>& B, int n);

}  // namespace sbp4
