// Internal declarations of libeqd: operator tables, device views and kernels.
//
// Vector space. Every Krylov vector lives on range(P) of the reference's split-node
// system (src/ElasticitySplitNode.jl:171-211): far-field DOFs are exactly zero and
// the two DOFs of a fault pair are equal. The library therefore stores one merged
// grid: x1 runs over both blocks with the fault face stored once, and only interior
// nodes are kept.
//   merged x index  m in [0, NX),  NX = 2*n1 - 3   (plane M = m + 1 of 2*n1 - 1)
//   y, z index      in [0, NY),    NY = n23 - 2
//   node            = m + NX*(y + NY*z)
// A full-space dot product equals the merged one with weight 2 on the face plane.
//
// Batched layout ("tile-interleaved"): columns are grouped in tiles of K; element
// (node, comp, column j of tile t) sits at ((t*NN + node)*3 + comp)*K + j, so the K
// columns of one node are contiguous and a wavefront with K = 64 handles one node
// for 64 right-hand sides with uniform control flow.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace eqd {

constexpr int kTabW = 16;  // max taps of one 1D table row (zero padded)

// One banded 1D operator, row r reads input columns start[r] .. start[r]+width[r]-1.
struct HostTab {
    int nrows = 0;
    std::vector<int> start, width;
    std::vector<double> c;  // nrows * kTabW
};

// All 1D data the kernels need, built once per operator on the host.
struct HostTables {
    int n1 = 0, n23 = 0, NX = 0, NY = 0, mface = 0;
    double h1 = 0, h2 = 0, lambda = 0, mu = 0;
    // reduced operator -H P (D+SAT) P on the merged interior grid
    HostTab AXs, AXo, MX;        // x: same-comp (c=0), other-comp (c=1,2), (lam+mu) D1x
    HostTab Ts, To, D1, D1s;     // y,z: same-comp, other-comp, D1, (lam+mu) D1
    std::vector<double> Hx, Hy;  // quadrature weights (merged x, interior y)
    // right-hand side H P (D+SAT) chi(s): jump coefficients along merged x
    std::vector<double> Jt, Jn;  // tangential comps (from s directly), normal comp (from D1 s)
    HostTab D1full;              // interior rows, full columns (slip may sit on the ring)
    // traction extraction 1/2 (T U- + T U+) on Omega_f
    std::vector<double> Dn;      // normal-derivative weights on merged x (times mu/2)
    double Jext = 0;             // coefficient of the slip itself
    HostTab D1ext;               // rows: full y index; cols: interior
};

HostTables build_tables(int n1, int n23, double L_normal, double L_fault, double lambda, double mu);

// diag(A) at the rows of the merged unknowns (node-major, component fastest), the Jacobi
// preconditioner's M = diag(A) on range(P).
std::vector<double> merged_diagonal(const HostTables& t);

// Fixed-reach band form of a table: band[r*(2R+1) + R + d] = coefficient of column r + d.
// The reaches below hold for every grid (they are set by the SBP-4 closures restricted
// to interior columns); build_band() throws if a table exceeds its reach.
constexpr int kRxS = 5;  // AXs: wide D1*D1 along x1 plus the SAT normal closure
constexpr int kRxO = 3;  // AXo, MX
constexpr int kRyS = 4;  // Ts
constexpr int kRyO = 2;  // To, D1, D1s
std::vector<double> build_band(const HostTab& t, int R);

// Pore-pressure operator factor along one Omega_f axis (D2 + Neumann SATs) and the
// SBP-4 quadrature weights, n x n row-major.
void sbp4_neumann_1d(int n, double h, std::vector<double>& L, std::vector<double>& H);
// SBP-4 first derivative on n points with spacing h, n x n row-major.
void sbp4_d1_1d(int n, double h, std::vector<double>& D1);

// Device view of the band tables used by the pencil kernel.
template <typename T>
struct PencilView {
    int NX, NY, mface, nseg;
    int by;  // x2 rows per block of the pencil order (NY: whole rows)
    const T* AXs;  // NX x (2*kRxS+1)
    const T* AXo;  // NX x (2*kRxO+1)
    const T* MX;   // NX x (2*kRxO+1)
    const T* Ts;   // NY x (2*kRyS+1)
    const T* To;   // NY x (2*kRyO+1)
    const T* D1;
    const T* D1s;
    const T* Hx;
    const T* Hy;
};

// Device view of one table (typed coefficients).
template <typename T>
struct Tab {
    const int* start;
    const int* width;
    const T* c;
};

template <typename T>
struct OpView {
    int NX, NY, mface;
    Tab<T> AXs, AXo, MX, Ts, To, D1, D1s;
    const T* Hx;
    const T* Hy;
};

struct ExtractView {
    int NX, NY, mface, n23, i2lo, n2f, nf;
    double mu, Jext;
    const double* Dn;
    Tab<double> D1ext;
};

struct RhsView {
    int NX, NY, n23, i2lo, n2f, nf;
    const double* Jt;
    const double* Jn;
    const double* Hx;
    const double* Hy;
    Tab<double> D1full;
};

}  // namespace eqd
