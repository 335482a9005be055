// CPU FP64 oracle: assembled split-node system, Krylov-style CG, K build.
// TEST INFRASTRUCTURE ONLY -- see eqd_oracle.h for the rules and the reference map.
#include "eqd_oracle.h"
#include "csr.h"
#include "sbp4.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Model {
    eqo_params prm;
    int n1, n23;
    int64_t N;     // nodes per block
    int64_t Ntot;  // 6N
    double h1, h2;
    sbp4::Ops1D o1, o2;  // x1 ops (n1, h1); x2 = x3 ops (n23, h2)
    Csr A, HP_DSAT, P, T2, T3, DSAT;
    std::vector<double> Hdiag;
    std::vector<double> Minv;  // Jacobi M^{-1} (empty when precond = none)
    // Omega_f bookkeeping (src/FaultResponse.jl:112-141)
    int n2f, n3f, i2lo, i3lo;
    std::vector<int64_t> omega;  // face-linear index (i2 + n23*i3) per Omega_f node, x2 fastest
    std::vector<int64_t> chi_m[2], chi_p[2];
    eqo_stats st{0, 0, 0, 0};

    int64_t lin(int i1, int i2, int i3) const { return i1 + (int64_t)n1 * (i2 + (int64_t)n23 * i3); }
};

// 3D operator as the tensor product of 1D dense operators along x1, x2, x3
// (nullptr = identity); entries with a zero 1D factor are not stored.
Csr kron3(const Model& m, const std::vector<double>* M1, const std::vector<double>* M2,
          const std::vector<double>* M3, double scale) {
    const int n1 = m.n1, n2 = m.n23, n3 = m.n23;
    Triplets t;
    auto row_nz = [](const std::vector<double>* M, int n, int i, std::vector<std::pair<int, double>>& out) {
        out.clear();
        if (!M) {
            out.push_back({i, 1.0});
            return;
        }
        for (int j = 0; j < n; ++j) {
            double v = (*M)[(size_t)i * n + j];
            if (v != 0.0) out.push_back({j, v});
        }
    };
    std::vector<std::pair<int, double>> r1, r2, r3;
    for (int i3 = 0; i3 < n3; ++i3) {
        row_nz(M3, n3, i3, r3);
        for (int i2 = 0; i2 < n2; ++i2) {
            row_nz(M2, n2, i2, r2);
            for (int i1 = 0; i1 < n1; ++i1) {
                row_nz(M1, n1, i1, r1);
                int64_t row = m.lin(i1, i2, i3);
                for (auto& c3 : r3)
                    for (auto& c2 : r2)
                        for (auto& c1 : r1)
                            t.push(row, m.lin(c1.first, c2.first, c3.first), scale * c1.second * c2.second * c3.second);
            }
        }
    }
    return csr_from_triplets(t, m.N, m.N);
}

// elastic_blocks (src/Elasticity.jl:80-89): M[j][j] = mu*Lap + mu*D2_j + lam*D1_j D1_j,
// M[j][k] = mu*D1_j D1_k + lam*D1_j D1_k.
void elastic_blocks(const Model& m, Csr Mjk[3][3]) {
    const double lam = m.prm.lambda, mu = m.prm.mu;
    const std::vector<double>* D1[3] = {&m.o1.D1, &m.o2.D1, &m.o2.D1};
    const std::vector<double>* D2[3] = {&m.o1.D2, &m.o2.D2, &m.o2.D2};
    std::vector<double> W1 = sbp4::matmul(m.o1.D1, m.o1.D1, m.n1);
    std::vector<double> W2 = sbp4::matmul(m.o2.D1, m.o2.D1, m.n23);
    const std::vector<double>* W[3] = {&W1, &W2, &W2};
    auto axis_op = [&](int a, const std::vector<double>* M, double s) {
        return kron3(m, a == 0 ? M : nullptr, a == 1 ? M : nullptr, a == 2 ? M : nullptr, s);
    };
    Csr lap = csr_add(csr_add(axis_op(0, D2[0], mu), axis_op(1, D2[1], mu)), axis_op(2, D2[2], mu));
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
            if (j == k) {
                Mjk[j][j] = csr_add(csr_add(lap, axis_op(j, D2[j], mu)), axis_op(j, W[j], lam));
            } else {
                const std::vector<double>* f[3] = {nullptr, nullptr, nullptr};
                f[j] = D1[j];
                f[k] = D1[k];
                Csr mixed_mu = kron3(m, f[0], f[1], f[2], mu);
                Csr mixed_lam = kron3(m, f[0], f[1], f[2], lam);
                Mjk[j][k] = csr_add(mixed_mu, mixed_lam);
            }
        }
}

// traction_blocks (src/Elasticity.jl:195-221) on the x1 face at index `face`,
// fixed +x1 orientation; `dhat` is the +x1 boundary derivative row at that face.
void traction_blocks(const Model& m, int face, const std::vector<double>& dhat, Csr T[3][3]) {
    const double lam = m.prm.lambda, mu = m.prm.mu;
    const int n1 = m.n1, n2 = m.n23;
    const int64_t Nb = (int64_t)n2 * n2;
    auto eD1 = [&](int k, double s) {
        Triplets t;
        for (int i3 = 0; i3 < n2; ++i3)
            for (int i2 = 0; i2 < n2; ++i2) {
                int64_t b = i2 + (int64_t)n2 * i3;
                if (k == 0) {
                    for (int c = 0; c < n1; ++c) {
                        double v = m.o1.at(m.o1.D1, face, c);
                        if (v != 0.0) t.push(b, m.lin(c, i2, i3), s * v);
                    }
                } else {
                    int i = (k == 1) ? i2 : i3;
                    for (int c = 0; c < n2; ++c) {
                        double v = m.o2.at(m.o2.D1, i, c);
                        if (v == 0.0) continue;
                        t.push(b, k == 1 ? m.lin(face, c, i3) : m.lin(face, i2, c), s * v);
                    }
                }
            }
        return csr_from_triplets(t, Nb, m.N);
    };
    auto dh = [&](double s) {
        Triplets t;
        for (int i3 = 0; i3 < n2; ++i3)
            for (int i2 = 0; i2 < n2; ++i2)
                for (int c = 0; c < n1; ++c)
                    if (dhat[c] != 0.0) t.push(i2 + (int64_t)n2 * i3, m.lin(c, i2, i3), s * dhat[c]);
        return csr_from_triplets(t, Nb, m.N);
    };
    Triplets empty;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) T[i][k] = csr_from_triplets(empty, Nb, m.N);
    T[0][0] = csr_add(eD1(0, lam), dh(2.0 * mu));
    T[0][1] = eD1(1, lam);
    T[0][2] = eD1(2, lam);
    for (int i = 1; i < 3; ++i) {
        T[i][0] = eD1(i, mu);
        T[i][i] = dh(mu);
    }
}

void append_block(Triplets& t, const Csr& B, int64_t row0, int64_t col0, double s) {
    for (int64_t r = 0; r < B.nrows; ++r)
        for (int64_t p = B.rowptr[r]; p < B.rowptr[r + 1]; ++p) t.push(row0 + r, col0 + B.col[p], s * B.val[p]);
}

Csr hcat3(const Csr T[3], int64_t N) {
    Triplets t;
    for (int k = 0; k < 3; ++k) append_block(t, T[k], 0, k * N, 1.0);
    return csr_from_triplets(t, T[0].nrows, 3 * N);
}

std::unique_ptr<Model> build_model(const eqo_params& prm) {
    auto mp = std::make_unique<Model>();
    Model& m = *mp;
    m.prm = prm;
    m.n1 = prm.n1;
    m.n23 = prm.n23;
    if (m.n1 < 9 || m.n23 < 9) throw std::runtime_error("SBP order 4 needs at least 9 points per dimension");
    if (prm.precond != 0 && prm.precond != 1) throw std::runtime_error("precond must be :none or :jacobi");
    m.N = (int64_t)m.n1 * m.n23 * m.n23;
    m.Ntot = 6 * m.N;
    m.h1 = prm.L_normal / (m.n1 - 1);
    m.h2 = 2.0 * prm.L_fault / (m.n23 - 1);
    m.o1 = sbp4::build(m.n1, m.h1);
    m.o2 = sbp4::build(m.n23, m.h2);
    const int n1 = m.n1, n2 = m.n23;
    const int64_t N = m.N;
    const int fm = n1 - 1, fp = 0;  // fault face index on g- (upper) and g+ (lower)

    // Omega_f (src/FaultResponse.jl:112-126)
    {
        double tol = 1e-9 * std::max(prm.l_f, 1.0);
        std::vector<int> idx;
        for (int i = 0; i < n2; ++i) {
            double x = -prm.L_fault + i * m.h2;
            if (std::fabs(x) <= prm.l_f + tol) idx.push_back(i);
        }
        if (idx.empty()) throw std::runtime_error("no fault nodes inside Omega_f");
        double xlo = -prm.L_fault + idx.front() * m.h2, xhi = -prm.L_fault + idx.back() * m.h2;
        if (!(std::fabs(xlo + prm.l_f) <= tol && std::fabs(xhi - prm.l_f) <= tol))
            throw std::runtime_error("the elastic grid has no nodes exactly on +-l_f; pick a spacing that divides l_f");
        m.i2lo = m.i3lo = idx.front();
        m.n2f = m.n3f = (int)idx.size();
        for (int b = 0; b < m.n3f; ++b)
            for (int a = 0; a < m.n2f; ++a) m.omega.push_back((m.i2lo + a) + (int64_t)n2 * (m.i3lo + b));
        for (int ci = 0; ci < 2; ++ci) {
            int comp = ci + 1;
            for (int64_t b : m.omega) {
                int i2 = (int)(b % n2), i3 = (int)(b / n2);
                m.chi_m[ci].push_back(comp * N + m.lin(fm, i2, i3));
                m.chi_p[ci].push_back(3 * N + comp * N + m.lin(fp, i2, i3));
            }
        }
    }

    // D: block-diagonal elastic operator (src/ElasticitySplitNode.jl:117-119)
    Csr Mjk[3][3];
    elastic_blocks(m, Mjk);
    Triplets td;
    for (int side = 0; side < 2; ++side)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) append_block(td, Mjk[j][k], side * 3 * N + j * N, side * 3 * N + k * N, 1.0);
    Csr Dmat = csr_from_triplets(td, m.Ntot, m.Ntot);

    // SAT (src/ElasticitySplitNode.jl:139-163)
    Csr Tm[3][3], Tp[3][3];
    traction_blocks(m, fm, m.o1.dr, Tm);
    traction_blocks(m, fp, m.o1.dl, Tp);
    auto prolong = [&](int face) {
        Triplets t;
        for (int i3 = 0; i3 < n2; ++i3)
            for (int i2 = 0; i2 < n2; ++i2) {
                double Hb = m.o2.H[i2] * m.o2.H[i3];
                double Hv = m.o1.H[face] * m.o2.H[i2] * m.o2.H[i3];
                t.push(m.lin(face, i2, i3), i2 + (int64_t)n2 * i3, -(1.0 / Hv) * Hb);
            }
        return csr_from_triplets(t, N, (int64_t)n2 * n2);
    };
    Csr Pro_m = prolong(fm), Pro_p = prolong(fp);
    Triplets ts;
    for (int j = 0; j < 3; ++j) {
        int64_t row0_m = j * N, row0_p = 3 * N + j * N;
        Csr Tpj = hcat3(Tp[j], N), Tmj = hcat3(Tm[j], N);
        append_block(ts, csr_mul(Pro_m, Tmj), row0_m, 0, 0.5);
        append_block(ts, csr_mul(Pro_m, Tpj), row0_m, 3 * N, -0.5);
        append_block(ts, csr_mul(Pro_p, Tpj), row0_p, 3 * N, -0.5);
        append_block(ts, csr_mul(Pro_p, Tmj), row0_p, 0, 0.5);
    }
    Csr SATmat = csr_from_triplets(ts, m.Ntot, m.Ntot);
    m.DSAT = csr_add(Dmat, SATmat);

    // P (src/ElasticitySplitNode.jl:171-203), stored after dropzeros!
    {
        std::vector<int> kind(m.Ntot, 0);  // 0 identity, 1 far-field, 2 fault pair
        std::vector<int64_t> partner(m.Ntot, -1);
        for (int side = 0; side < 2; ++side)
            for (int i3 = 0; i3 < n2; ++i3)
                for (int i2 = 0; i2 < n2; ++i2)
                    for (int i1 = 0; i1 < n1; ++i1) {
                        bool ff = (i2 == 0 || i2 == n2 - 1 || i3 == 0 || i3 == n2 - 1) ||
                                  (side == 0 ? i1 == 0 : i1 == n1 - 1);
                        if (!ff) continue;
                        for (int c = 0; c < 3; ++c) kind[side * 3 * N + c * N + m.lin(i1, i2, i3)] = 1;
                    }
        for (int i3 = 1; i3 < n2 - 1; ++i3)
            for (int i2 = 1; i2 < n2 - 1; ++i2)
                for (int c = 0; c < 3; ++c) {
                    int64_t rm = c * N + m.lin(fm, i2, i3), rp = 3 * N + c * N + m.lin(fp, i2, i3);
                    kind[rm] = kind[rp] = 2;
                    partner[rm] = rp;
                    partner[rp] = rm;
                }
        Triplets tp;
        for (int64_t r = 0; r < m.Ntot; ++r) {
            if (kind[r] == 0) tp.push(r, r, 1.0);
            if (kind[r] == 2) {
                int64_t rm = std::min(r, partner[r]), rp = std::max(r, partner[r]);
                tp.push(r, rm, 0.5);
                tp.push(r, rp, 0.5);
            }
        }
        m.P = csr_from_triplets(tp, m.Ntot, m.Ntot);
    }

    // H and the assembled operators (src/ElasticitySplitNode.jl:205-211)
    m.Hdiag.resize(m.Ntot);
    {
        Triplets th;
        for (int side = 0; side < 2; ++side)
            for (int c = 0; c < 3; ++c)
                for (int i3 = 0; i3 < n2; ++i3)
                    for (int i2 = 0; i2 < n2; ++i2)
                        for (int i1 = 0; i1 < n1; ++i1) {
                            int64_t r = side * 3 * N + c * N + m.lin(i1, i2, i3);
                            double w = m.o1.H[i1] * m.o2.H[i2] * m.o2.H[i3];
                            m.Hdiag[r] = w;
                            th.push(r, r, w);
                        }
        Csr H = csr_from_triplets(th, m.Ntot, m.Ntot);
        Csr HP = csr_mul(H, m.P);
        m.HP_DSAT = csr_mul(HP, m.DSAT);
        m.A = csr_mul(csr_scale(m.HP_DSAT, -1.0), m.P);
    }

    // T2, T3 = 1/2 [Tm_j | Tp_j] (src/FaultResponse.jl:130-132)
    for (int j = 1; j < 3; ++j) {
        Triplets t;
        for (int k = 0; k < 3; ++k) {
            append_block(t, Tm[j][k], 0, k * N, 0.5);
            append_block(t, Tp[j][k], 0, 3 * N + k * N, 0.5);
        }
        (j == 1 ? m.T2 : m.T3) = csr_from_triplets(t, (int64_t)n2 * n2, m.Ntot);
    }

    // Jacobi preconditioner (src/ElasticitySplitNode.jl:352-357)
    if (prm.precond == 1) {
        m.Minv.resize(m.Ntot);
        for (int64_t r = 0; r < m.Ntot; ++r) {
            double d = 0.0;
            for (int64_t p = m.A.rowptr[r]; p < m.A.rowptr[r + 1]; ++p)
                if (m.A.col[p] == r) d += m.A.val[p];
            if (d < 0) throw std::runtime_error("diag(A) has negative entries; A is not PSD");
            m.Minv[r] = d == 0.0 ? 1.0 : 1.0 / d;
        }
    }
    return mp;
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// Krylov.cg! 0.10.9 semantics (src/ElasticitySplitNode.jl:378-396): x0 = 0,
// stop on the recursive residual rNorm <= atol + rtol*rNorm0, itmax 0 -> 2n,
// zero-curvature exit, rNorm + 1 <= 1 exit.
struct CgOut {
    int niter;
    bool solved;
};
CgOut cg_solve(const Model& m, const double* b, double* xout) {
    const int64_t n = m.Ntot;
    std::vector<double> x(n, 0.0), r(b, b + n), z, p, Ap(n);
    const bool pre = !m.Minv.empty();
    if (pre) {
        z.resize(n);
        for (int64_t i = 0; i < n; ++i) z[i] = m.Minv[i] * r[i];
    }
    const std::vector<double>& zr = pre ? z : r;
    p = zr;
    double gamma = dot(r, zr);
    double rNorm = std::sqrt(gamma);
    if (gamma == 0.0) {
        std::memset(xout, 0, sizeof(double) * n);
        return {0, true};
    }
    int64_t itmax = m.prm.itmax == 0 ? 2 * n : m.prm.itmax;
    double eps_ = m.prm.atol + m.prm.rtol * rNorm;
    double pNorm2 = gamma;
    bool solved = rNorm <= eps_, tired = 0 >= itmax, zero_curv = false;
    int iter = 0;
    const double meps = std::numeric_limits<double>::epsilon();
    while (!(solved || tired || zero_curv)) {
        csr_spmv(m.A, p.data(), Ap.data());
        double pAp = dot(p, Ap);
        if (pAp <= meps * pNorm2 && std::fabs(pAp) <= meps * pNorm2) {
            zero_curv = true;
            continue;
        }
        double alpha = gamma / pAp;
        for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
        for (int64_t i = 0; i < n; ++i) r[i] -= alpha * Ap[i];
        if (pre)
            for (int64_t i = 0; i < n; ++i) z[i] = m.Minv[i] * r[i];
        double gnext = dot(r, zr);
        rNorm = std::sqrt(gnext);
        bool mach = rNorm + 1.0 <= 1.0;
        solved = rNorm <= eps_ || mach;
        if (!solved) {
            double beta = gnext / gamma;
            pNorm2 = gnext + beta * beta * pNorm2;
            gamma = gnext;
            for (int64_t i = 0; i < n; ++i) p[i] = zr[i] + beta * p[i];
        }
        ++iter;
        tired = iter >= itmax;
    }
    std::memcpy(xout, x.data(), sizeof(double) * n);
    return {iter, solved};
}

void record(eqo_stats& st, const CgOut& o) {
    st.solves += 1;
    st.iterations += o.niter;
    st.max_iterations = std::max<int64_t>(st.max_iterations, o.niter);
    if (!o.solved) {
        st.unconverged += 1;
        if (st.unconverged == 1)
            std::fprintf(stderr, "warning: CG did not converge on a split-node solve after %d iterations\n", o.niter);
    }
}

void build_chi(const Model& m, const double* s2, const double* s3, double* chi) {
    std::memset(chi, 0, sizeof(double) * m.Ntot);
    for (size_t k = 0; k < m.omega.size(); ++k) {
        chi[m.chi_m[0][k]] = -s2[k] / 2;
        chi[m.chi_p[0][k]] = s2[k] / 2;
        chi[m.chi_m[1][k]] = -s3[k] / 2;
        chi[m.chi_p[1][k]] = s3[k] / 2;
    }
}

// shear_traction! (src/FaultResponse.jl:201-211) with buffers private to the caller.
struct Work {
    std::vector<double> chi, rhs, u, U, tau;
};
CgOut shear_traction(const Model& m, const double* s2, const double* s3, double* d2, double* d3, Work& w) {
    w.chi.resize(m.Ntot);
    w.rhs.resize(m.Ntot);
    w.u.resize(m.Ntot);
    w.U.resize(m.Ntot);
    w.tau.resize((size_t)m.n23 * m.n23);
    build_chi(m, s2, s3, w.chi.data());
    csr_spmv(m.HP_DSAT, w.chi.data(), w.rhs.data());
    CgOut o = cg_solve(m, w.rhs.data(), w.u.data());
    csr_spmv(m.P, w.u.data(), w.U.data());
    for (int64_t i = 0; i < m.Ntot; ++i) w.U[i] += w.chi[i];
    csr_spmv(m.T2, w.U.data(), w.tau.data());
    for (size_t k = 0; k < m.omega.size(); ++k) d2[k] = w.tau[m.omega[k]];
    csr_spmv(m.T3, w.U.data(), w.tau.data());
    for (size_t k = 0; k < m.omega.size(); ++k) d3[k] = w.tau[m.omega[k]];
    return o;
}

// Columns `cols` of K into out[(k)*2nf ...], strided over nthreads like
// src/FaultResponse.jl:265-281; each worker owns its buffers and counters.
void stiffness_cols(Model& m, const int* cols, int ncols, int nthreads, double* out) {
    const int nf = (int)m.omega.size();
    nthreads = std::max(1, std::min(nthreads, ncols));
    std::vector<eqo_stats> st(nthreads, eqo_stats{0, 0, 0, 0});
    auto worker = [&](int t) {
        std::vector<double> s2(nf), s3(nf);
        Work w;
        for (int q = t; q < ncols; q += nthreads) {
            int col = cols[q];
            std::fill(s2.begin(), s2.end(), 0.0);
            std::fill(s3.begin(), s3.end(), 0.0);
            if (col < nf) s2[col] = 1.0; else s3[col - nf] = 1.0;
            double* K = out + (size_t)q * 2 * nf;
            CgOut o = shear_traction(m, s2.data(), s3.data(), K, K + nf, w);
            st[t].solves += 1;
            st[t].iterations += o.niter;
            st[t].max_iterations = std::max<int64_t>(st[t].max_iterations, o.niter);
            st[t].unconverged += o.solved ? 0 : 1;
        }
    };
    if (nthreads == 1) worker(0);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) th.emplace_back(worker, t);
        for (auto& x : th) x.join();
    }
    for (auto& s : st) {
        m.st.solves += s.solves;
        m.st.iterations += s.iterations;
        m.st.max_iterations = std::max(m.st.max_iterations, s.max_iterations);
        m.st.unconverged += s.unconverged;
    }
}

const Csr* op_of(const Model& m, int op) {
    switch (op) {
        case EQO_OP_A: return &m.A;
        case EQO_OP_HPDSAT: return &m.HP_DSAT;
        case EQO_OP_P: return &m.P;
        case EQO_OP_T2: return &m.T2;
        case EQO_OP_T3: return &m.T3;
        case EQO_OP_DSAT: return &m.DSAT;
        default: return nullptr;
    }
}

}  // namespace

extern "C" {

void* eqo_create(const eqo_params* p, char* err, size_t errlen) {
    try {
        return build_model(*p).release();
    } catch (const std::exception& e) {
        if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
        return nullptr;
    }
}

void eqo_destroy(void* h) { delete static_cast<Model*>(h); }

int eqo_info(void* h, int64_t* out) {
    const Model& m = *static_cast<Model*>(h);
    out[EQO_NF] = (int64_t)m.omega.size();
    out[EQO_N2F] = m.n2f;
    out[EQO_N3F] = m.n3f;
    out[EQO_NTOT] = m.Ntot;
    out[EQO_NNZ_A] = m.A.nnz();
    out[EQO_NNZ_HPDSAT] = m.HP_DSAT.nnz();
    out[EQO_NNZ_P] = m.P.nnz();
    out[EQO_N_PER_BLOCK] = m.N;
    out[EQO_I2LO] = m.i2lo;
    out[EQO_I3LO] = m.i3lo;
    return 0;
}

int eqo_apply(void* h, int op, const double* x, double* y) {
    const Model& m = *static_cast<Model*>(h);
    if (op == EQO_OP_H) {
        for (int64_t i = 0; i < m.Ntot; ++i) y[i] = m.Hdiag[i] * x[i];
        return 0;
    }
    const Csr* a = op_of(m, op);
    if (!a) return 1;
    csr_spmv(*a, x, y);
    return 0;
}

int64_t eqo_export(void* h, int op, int64_t* rowptr, int64_t* col, double* val) {
    const Model& m = *static_cast<Model*>(h);
    const Csr* a = op_of(m, op);
    if (!a) return -1;
    if (rowptr) std::memcpy(rowptr, a->rowptr.data(), sizeof(int64_t) * (a->nrows + 1));
    if (col) std::memcpy(col, a->col.data(), sizeof(int64_t) * a->nnz());
    if (val) std::memcpy(val, a->val.data(), sizeof(double) * a->nnz());
    return a->nnz();
}

int eqo_build_chi(void* h, const double* s2, const double* s3, double* chi) {
    build_chi(*static_cast<Model*>(h), s2, s3, chi);
    return 0;
}

int eqo_cg(void* h, const double* b, double* x, int* niter, int* solved) {
    Model& m = *static_cast<Model*>(h);
    CgOut o = cg_solve(m, b, x);
    record(m.st, o);
    if (niter) *niter = o.niter;
    if (solved) *solved = o.solved ? 1 : 0;
    return 0;
}

int eqo_shear_traction(void* h, const double* s2, const double* s3, double* d2, double* d3) {
    Model& m = *static_cast<Model*>(h);
    Work w;
    CgOut o = shear_traction(m, s2, s3, d2, d3, w);
    record(m.st, o);
    return 0;
}

int eqo_fault_stiffness(void* h, int nthreads, double* K) {
    Model& m = *static_cast<Model*>(h);
    int ncols = 2 * (int)m.omega.size();
    std::vector<int> cols(ncols);
    for (int i = 0; i < ncols; ++i) cols[i] = i;
    stiffness_cols(m, cols.data(), ncols, nthreads, K);
    return 0;
}

int eqo_fault_stiffness_cols(void* h, const int* cols, int ncols, int nthreads, double* Kcols) {
    Model& m = *static_cast<Model*>(h);
    int nf2 = 2 * (int)m.omega.size();
    for (int i = 0; i < ncols; ++i)
        if (cols[i] < 0 || cols[i] >= nf2) return 2;
    stiffness_cols(m, cols, ncols, nthreads, Kcols);
    return 0;
}

// fault_stiffness_toeplitz (src/FaultResponse.jl:339-393): five sources in
// priority order (centre, then the four corners), first source to supply a
// separation wins; unreachable separations stay zero.
int eqo_fault_stiffness_toeplitz(void* h, double* K) {
    Model& m = *static_cast<Model*>(h);
    const int n2 = m.n2f, n3 = m.n3f, nf = (int)m.omega.size();
    if (nf != n2 * n3) return 3;
    const int ac = (n2 + 1) / 2 - 1, bc = (n3 + 1) / 2 - 1;  // 0-based
    const int src[5][2] = {{ac, bc}, {0, 0}, {n2 - 1, 0}, {0, n3 - 1}, {n2 - 1, n3 - 1}};
    const int W2 = 2 * n2 - 1, W3 = 2 * n3 - 1;
    std::vector<char> have((size_t)W2 * W3, 0);
    std::vector<double> ker((size_t)W2 * W3 * 4, 0.0);
    std::vector<double> s2(nf), s3(nf), c22(nf), c32(nf), d2(nf), d3(nf);
    Work w;
    for (int s = 0; s < 5; ++s) {
        int as = src[s][0], bs = src[s][1], j = as + n2 * bs;
        std::fill(s2.begin(), s2.end(), 0.0);
        std::fill(s3.begin(), s3.end(), 0.0);
        s2[j] = 1.0;
        record(m.st, shear_traction(m, s2.data(), s3.data(), c22.data(), c32.data(), w));
        s2[j] = 0.0;
        s3[j] = 1.0;
        record(m.st, shear_traction(m, s2.data(), s3.data(), d2.data(), d3.data(), w));
        for (int i = 0; i < nf; ++i) {
            int ai = i % n2, bi = i / n2;
            size_t key = (size_t)(ai - as + n2 - 1) + (size_t)W2 * (bi - bs + n3 - 1);
            if (have[key]) continue;
            have[key] = 1;
            ker[4 * key + 0] = c22[i];
            ker[4 * key + 1] = c32[i];
            ker[4 * key + 2] = d2[i];
            ker[4 * key + 3] = d3[i];
        }
    }
    const size_t ld = 2 * (size_t)nf;
    std::memset(K, 0, sizeof(double) * ld * ld);
    for (int jj = 0; jj < nf; ++jj) {
        int aj = jj % n2, bj = jj / n2;
        for (int i = 0; i < nf; ++i) {
            int ai = i % n2, bi = i / n2;
            size_t key = (size_t)(ai - aj + n2 - 1) + (size_t)W2 * (bi - bj + n3 - 1);
            if (!have[key]) continue;
            K[i + ld * jj] = ker[4 * key + 0];
            K[nf + i + ld * jj] = ker[4 * key + 1];
            K[i + ld * (nf + jj)] = ker[4 * key + 2];
            K[nf + i + ld * (nf + jj)] = ker[4 * key + 3];
        }
    }
    return 0;
}

int eqo_report(void* h, eqo_stats* st) {
    *st = static_cast<Model*>(h)->st;
    return 0;
}

void eqo_reset_stats(void* h) { static_cast<Model*>(h)->st = eqo_stats{0, 0, 0, 0}; }

int eqo_sbp1d(int n, double h, double* H, double* D1, double* D2, double* dl, double* dr) {
    try {
        sbp4::Ops1D o = sbp4::build(n, h);
        std::memcpy(H, o.H.data(), sizeof(double) * n);
        std::memcpy(D1, o.D1.data(), sizeof(double) * n * n);
        std::memcpy(D2, o.D2.data(), sizeof(double) * n * n);
        std::memcpy(dl, o.dl.data(), sizeof(double) * n);
        std::memcpy(dr, o.dr.data(), sizeof(double) * n);
        return 0;
    } catch (...) {
        return 1;
    }
}

}  // extern "C"
