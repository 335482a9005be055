// Host construction of the 1D operator tables behind the matrix-free kernels.
//
// The stencils are the SBP order-4 diagonal-norm operators the reference reads from
// Diffinitive's standard_diagonal.toml (src/BP8.jl:195): the published Mattsson &
// Nordstrom (2004) coefficients. The elastic operator is
//   D[j][j] = mu*(D2_1 + D2_2 + D2_3) + mu*D2_j + lam*D1_j D1_j,
//   D[j][k] = (lam + mu) * D1_j D1_k                        (src/Elasticity.jl:80-89)
// the fault SAT uses the fixed-+x1 traction of src/Elasticity.jl:195-221 with the
// half-weighted penalty of src/ElasticitySplitNode.jl:139-163, and P, H are those of
// src/ElasticitySplitNode.jl:171-211. Every x1 operator is first written on the full
// two-block x line (2*n1 points), averaged over the fault pair as P does, and then
// restricted to range(P) (face columns merged, far-field rows/columns dropped).
#include <cmath>
#include <stdexcept>

#include "eqd_internal.h"

namespace eqd {
namespace {

struct Dense1D {
    int n;
    double h;
    std::vector<double> H, D1, D2, dl, dr;  // dl/dr: +x derivative rows at the two ends
    double& d1(int i, int j) { return D1[(size_t)i * n + j]; }
    double& d2(int i, int j) { return D2[(size_t)i * n + j]; }
};

Dense1D sbp4_dense(int n, double h) {
    static const double wH[4] = {17.0 / 48, 59.0 / 48, 43.0 / 48, 49.0 / 48};
    static const double cD1[4][6] = {{-24.0 / 17, 59.0 / 34, -4.0 / 17, -3.0 / 34, 0, 0},
                                     {-0.5, 0, 0.5, 0, 0, 0},
                                     {4.0 / 43, -59.0 / 86, 0, 59.0 / 86, -4.0 / 43, 0},
                                     {3.0 / 98, 0, -59.0 / 98, 0, 32.0 / 49, -4.0 / 49}};
    static const double cD2[4][6] = {{2, -5, 4, -1, 0, 0},
                                     {1, -2, 1, 0, 0, 0},
                                     {-4.0 / 43, 59.0 / 43, -110.0 / 43, 59.0 / 43, -4.0 / 43, 0},
                                     {-1.0 / 49, 0, 59.0 / 49, -118.0 / 49, 64.0 / 49, -4.0 / 49}};
    static const double iD1[5] = {1.0 / 12, -2.0 / 3, 0, 2.0 / 3, -1.0 / 12};
    static const double iD2[5] = {-1.0 / 12, 4.0 / 3, -2.5, 4.0 / 3, -1.0 / 12};
    static const double bd[4] = {-11.0 / 6, 3, -1.5, 1.0 / 3};
    if (n < 9) throw std::runtime_error("SBP order 4 needs at least 9 points per dimension");
    Dense1D o{n, h, std::vector<double>(n, h), std::vector<double>((size_t)n * n, 0.0),
              std::vector<double>((size_t)n * n, 0.0), std::vector<double>(n, 0.0), std::vector<double>(n, 0.0)};
    for (int i = 0; i < 4; ++i) o.H[i] = o.H[n - 1 - i] = h * wH[i];
    for (int i = 4; i < n - 4; ++i)
        for (int t = 0; t < 5; ++t) {
            o.d1(i, i - 2 + t) = iD1[t] / h;
            o.d2(i, i - 2 + t) = iD2[t] / (h * h);
        }
    for (int r = 0; r < 4; ++r)
        for (int t = 0; t < 6; ++t) {
            o.d1(r, t) = cD1[r][t] / h;
            o.d1(n - 1 - r, n - 1 - t) = -cD1[r][t] / h;
            o.d2(r, t) = cD2[r][t] / (h * h);
            o.d2(n - 1 - r, n - 1 - t) = cD2[r][t] / (h * h);
        }
    for (int t = 0; t < 4; ++t) {
        o.dl[t] = bd[t] / h;
        o.dr[n - 1 - t] = -bd[t] / h;
    }
    return o;
}

std::vector<double> square(const std::vector<double>& A, int n) {
    std::vector<double> C((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            double a = A[(size_t)i * n + k];
            if (a != 0.0)
                for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * A[(size_t)k * n + j];
        }
    return C;
}

// Dense rows -> banded table (rows x cols, row-major), zero taps trimmed at both ends.
HostTab to_tab(const std::vector<double>& M, int rows, int cols) {
    HostTab t;
    t.nrows = rows;
    t.start.assign(rows, 0);
    t.width.assign(rows, 0);
    t.c.assign((size_t)rows * kTabW, 0.0);
    for (int r = 0; r < rows; ++r) {
        int lo = -1, hi = -1;
        for (int j = 0; j < cols; ++j)
            if (M[(size_t)r * cols + j] != 0.0) {
                if (lo < 0) lo = j;
                hi = j;
            }
        if (lo < 0) continue;
        if (hi - lo + 1 > kTabW) throw std::runtime_error("operator row wider than the table width");
        t.start[r] = lo;
        t.width[r] = hi - lo + 1;
        for (int j = lo; j <= hi; ++j) t.c[(size_t)r * kTabW + (j - lo)] = M[(size_t)r * cols + j];
    }
    return t;
}

}  // namespace

std::vector<double> build_band(const HostTab& t, int R) {
    const int Wb = 2 * R + 1;
    std::vector<double> b((size_t)t.nrows * Wb, 0.0);
    for (int r = 0; r < t.nrows; ++r)
        for (int k = 0; k < t.width[r]; ++k) {
            const int d = t.start[r] + k - r;
            const double c = t.c[(size_t)r * kTabW + k];
            if (c == 0.0) continue;
            if (d < -R || d > R) throw std::runtime_error("operator row exceeds its band reach");
            b[(size_t)r * Wb + R + d] = c;
        }
    return b;
}

HostTables build_tables(int n1, int n23, double L_normal, double L_fault, double lambda, double mu) {
    HostTables T;
    T.n1 = n1;
    T.n23 = n23;
    T.NX = 2 * n1 - 3;
    T.NY = n23 - 2;
    T.mface = n1 - 2;
    T.h1 = L_normal / (n1 - 1);
    T.h2 = 2.0 * L_fault / (n23 - 1);
    T.lambda = lambda;
    T.mu = mu;
    Dense1D ox = sbp4_dense(n1, T.h1), oy = sbp4_dense(n23, T.h2);
    std::vector<double> Wx = square(ox.D1, n1), Wy = square(oy.D1, n23);
    const int NF = 2 * n1;           // full two-block x line: minus i1 -> i1, plus j1 -> n1 + j1
    const int NM = 2 * n1 - 1;       // merged line (face stored once)
    const int fm = n1 - 1, fp = n1;  // face positions on the full line
    const double Hxf = ox.H[n1 - 1];
    const double sat = -1.0 / (2.0 * Hxf);  // 1/2 * Prolong, Prolong = -H^-1 e^T H_gamma

    // full-line operator: block-diagonal copy of an n1 x n1 block op
    auto blockdiag = [&](const std::vector<double>& B, double s) {
        std::vector<double> F((size_t)NF * NF, 0.0);
        for (int i = 0; i < n1; ++i)
            for (int j = 0; j < n1; ++j) {
                F[(size_t)i * NF + j] = s * B[(size_t)i * n1 + j];
                F[(size_t)(n1 + i) * NF + n1 + j] = s * B[(size_t)i * n1 + j];
            }
        return F;
    };
    // SAT rows on both face rows: sat * (Nm on minus cols  -  Np on plus cols)
    auto add_sat = [&](std::vector<double>& F, const std::vector<double>& Nm, const std::vector<double>& Np) {
        for (int row : {fm, fp})
            for (int j = 0; j < n1; ++j) {
                F[(size_t)row * NF + j] += sat * Nm[j];
                F[(size_t)row * NF + n1 + j] -= sat * Np[j];
            }
    };
    auto lincomb = [&](const std::vector<double>& a, double sa, const std::vector<double>& b, double sb) {
        std::vector<double> r(a.size());
        for (size_t i = 0; i < a.size(); ++i) r[i] = sa * a[i] + sb * b[i];
        return r;
    };
    std::vector<double> eF_m(n1, 0.0), eF_p(n1, 0.0), D1_m(n1), D1_p(n1);
    eF_m[n1 - 1] = 1.0;
    eF_p[0] = 1.0;
    for (int j = 0; j < n1; ++j) {
        D1_m[j] = ox.D1[(size_t)(n1 - 1) * n1 + j];  // boundary row of D1 at the minus face
        D1_p[j] = ox.D1[j];                          // and at the plus face
    }
    // x-direction term of the same component (c = 0): (2mu D2 + lam W) + SAT of sigma_11's normal part
    std::vector<double> Fs = blockdiag(lincomb(ox.D2, 2 * mu, Wx, lambda), 1.0);
    add_sat(Fs, lincomb(D1_m, lambda, ox.dr, 2 * mu), lincomb(D1_p, lambda, ox.dl, 2 * mu));
    // other components (c = 1, 2): mu D2 + SAT of mu*dhat
    std::vector<double> Fo = blockdiag(ox.D2, mu);
    add_sat(Fo, lincomb(ox.dr, mu, ox.dr, 0.0), lincomb(ox.dl, mu, ox.dl, 0.0));
    // mixed x-factor into the normal component: (lam+mu) D1 + SAT tangential lam*(e- - e+)
    std::vector<double> Fmn = blockdiag(ox.D1, lambda + mu);
    add_sat(Fmn, lincomb(eF_m, lambda, eF_m, 0.0), lincomb(eF_p, lambda, eF_p, 0.0));
    // mixed x-factor into a tangential component: (lam+mu) D1 + SAT tangential mu*(e- - e+)
    std::vector<double> Fmt = blockdiag(ox.D1, lambda + mu);
    add_sat(Fmt, lincomb(eF_m, mu, eF_m, 0.0), lincomb(eF_p, mu, eF_p, 0.0));

    // P-average the rows onto the merged line
    auto xrow = [&](int M) { return M < n1 - 1 ? M : M + 1; };
    auto avg_rows = [&](const std::vector<double>& F) {
        std::vector<double> R((size_t)NM * NF, 0.0);
        for (int M = 0; M < NM; ++M)
            for (int j = 0; j < NF; ++j)
                R[(size_t)M * NF + j] = (M == n1 - 1) ? 0.5 * (F[(size_t)fm * NF + j] + F[(size_t)fp * NF + j])
                                                      : F[(size_t)xrow(M) * NF + j];
        return R;
    };
    // restrict to range(P): merge the face columns, drop far-field rows/columns
    auto reduce = [&](const std::vector<double>& R) {
        std::vector<double> out((size_t)T.NX * T.NX, 0.0);
        for (int m = 0; m < T.NX; ++m) {
            int M = m + 1;
            for (int j = 0; j < NF; ++j) {
                int Mc = j <= fm ? j : j - 1;  // full column -> merged column
                if (Mc < 1 || Mc > NM - 2) continue;
                out[(size_t)m * T.NX + (Mc - 1)] += R[(size_t)M * NF + j];
            }
        }
        return out;
    };
    auto jump = [&](const std::vector<double>& R) {  // chi = -s/2 at fm, +s/2 at fp
        std::vector<double> J(T.NX);
        for (int m = 0; m < T.NX; ++m) {
            int M = m + 1;
            J[m] = 0.5 * (R[(size_t)M * NF + fp] - R[(size_t)M * NF + fm]);
        }
        return J;
    };
    std::vector<double> Rs = avg_rows(Fs), Ro = avg_rows(Fo), Rmn = avg_rows(Fmn), Rmt = avg_rows(Fmt);
    T.AXs = to_tab(reduce(Rs), T.NX, T.NX);
    T.AXo = to_tab(reduce(Ro), T.NX, T.NX);
    T.MX = to_tab(reduce(Rmn), T.NX, T.NX);  // the SAT tangential part cancels on range(P)
    {
        std::vector<double> a = reduce(Rmn), b = reduce(Rmt);
        for (size_t i = 0; i < a.size(); ++i)
            if (a[i] != b[i]) throw std::runtime_error("mixed x-factors differ on range(P)");
    }
    T.Jt = jump(Ro);
    T.Jn = jump(Rmn);
    T.Hx.resize(T.NX);
    for (int m = 0; m < T.NX; ++m) {
        int M = m + 1;
        T.Hx[m] = M <= n1 - 1 ? ox.H[M] : ox.H[M - (n1 - 1)];
    }

    // y/z tables on the interior rows and columns
    const int NY = T.NY;
    auto interior = [&](const std::vector<double>& B, double s) {
        std::vector<double> out((size_t)NY * NY);
        for (int i = 0; i < NY; ++i)
            for (int j = 0; j < NY; ++j) out[(size_t)i * NY + j] = s * B[(size_t)(i + 1) * n23 + (j + 1)];
        return out;
    };
    T.Ts = to_tab(interior(lincomb(oy.D2, 2 * mu, Wy, lambda), 1.0), NY, NY);
    T.To = to_tab(interior(oy.D2, mu), NY, NY);
    T.D1 = to_tab(interior(oy.D1, 1.0), NY, NY);
    T.D1s = to_tab(interior(oy.D1, lambda + mu), NY, NY);
    T.Hy.resize(NY);
    for (int i = 0; i < NY; ++i) T.Hy[i] = oy.H[i + 1];
    {  // interior rows, full columns
        std::vector<double> out((size_t)NY * n23);
        for (int i = 0; i < NY; ++i)
            for (int j = 0; j < n23; ++j) out[(size_t)i * n23 + j] = oy.D1[(size_t)(i + 1) * n23 + j];
        T.D1full = to_tab(out, NY, n23);
    }
    {  // full rows, interior columns
        std::vector<double> out((size_t)n23 * NY);
        for (int i = 0; i < n23; ++i)
            for (int j = 0; j < NY; ++j) out[(size_t)i * NY + j] = oy.D1[(size_t)i * n23 + (j + 1)];
        T.D1ext = to_tab(out, n23, NY);
    }
    // extraction: 1/2 mu (dhat- (+) dhat+) on the full line, P-embedded columns
    {
        std::vector<double> Re((size_t)NF, 0.0);
        for (int j = 0; j < n1; ++j) {
            Re[j] = 0.5 * mu * ox.dr[j];
            Re[n1 + j] = 0.5 * mu * ox.dl[j];
        }
        T.Dn.assign(T.NX, 0.0);
        for (int j = 0; j < NF; ++j) {
            int Mc = j <= fm ? j : j - 1;
            if (Mc < 1 || Mc > NM - 2) continue;
            T.Dn[Mc - 1] += Re[j];
        }
        T.Jext = 0.5 * (Re[fp] - Re[fm]);
    }
    return T;
}

namespace {
double tab_entry(const HostTab& t, int row, int col) {
    const int j = col - t.start[row];
    return (j >= 0 && j < t.width[row]) ? t.c[(size_t)row * kTabW + j] : 0.0;
}
}  // namespace

// diag(A), A = -HP(D+SAT)P, at the rows the merged unknowns stand for, node-major with the
// component fastest. Only the same-axis, same-component tables reach the diagonal:
// component 0 meets AXs along x1 and To along x2 and x3, component c = 1, 2 meets AXo along
// x1 and Ts along its own axis; the mixed terms couple different components. A face unknown
// carries the fault pair (rm, rp): rows rm and rp of A are equal, so A[rm,rp] = A[rm,rm] and
// the merged operator's diagonal there is 2 A[rm,rm] (src/ElasticitySplitNode.jl:306-322).
std::vector<double> merged_diagonal(const HostTables& t) {
    const int NX = t.NX, NY = t.NY;
    std::vector<double> d((size_t)NX * NY * NY * 3);
    for (int z = 0; z < NY; ++z)
        for (int y = 0; y < NY; ++y)
            for (int m = 0; m < NX; ++m) {
                const double hw = -(t.Hx[m] * t.Hy[y] * t.Hy[z]) * (m == t.mface ? 0.5 : 1.0);
                const double tyo = tab_entry(t.To, y, y), tzo = tab_entry(t.To, z, z);
                const double a[3] = {tab_entry(t.AXs, m, m) + tyo + tzo,
                                     tab_entry(t.AXo, m, m) + tab_entry(t.Ts, y, y) + tzo,
                                     tab_entry(t.AXo, m, m) + tyo + tab_entry(t.Ts, z, z)};
                const size_t node = (size_t)m + (size_t)NX * (y + (size_t)NY * z);
                for (int c = 0; c < 3; ++c) d[node * 3 + c] = hw * a[c];
            }
    return d;
}

// One-dimensional factor of the pore-pressure operator on the Omega_f grid
// (src/PorePressure.jl:36-43): the narrow D2 plus the Neumann (no-flux) SATs,
// L = D2 + e_0 d_l / H_0 - e_N d_r / H_N, dense n x n row-major; H gets the weights.
void sbp4_neumann_1d(int n, double h, std::vector<double>& L, std::vector<double>& H) {
    Dense1D o = sbp4_dense(n, h);
    L = o.D2;
    for (int j = 0; j < n; ++j) {
        L[j] += o.dl[j] / o.H[0];
        L[(size_t)(n - 1) * n + j] -= o.dr[j] / o.H[n - 1];
    }
    H = o.H;
}

// SBP-4 first derivative D1 on n points (dense n x n row-major): the operator of the Darcy
// velocity q_j = -(k / viscosity) D1_j p on the Omega_f grid (src/PorePressure.jl:52-56).
void sbp4_d1_1d(int n, double h, std::vector<double>& D1) { D1 = sbp4_dense(n, h).D1; }

}  // namespace eqd
