// CPU FP64 oracle of the BP8 time loop. TEST INFRASTRUCTURE ONLY -- see bp8_oracle.h.
#include "bp8_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbp4.h"

namespace {

struct NewtonFail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// src/RateStateFriction.jl:77-96
double slip_rate(double T, double theta, double sig, double eta, double a, double b, double Dc, double Vs, double fs,
                 double V0, double tol = 1e-12, int maxiter = 50, double maxstep = 5.0, bool* ok = nullptr) {
    const double C = std::exp((fs + b * std::log(Vs * theta / Dc)) / a);
    double x = std::log(V0);
    bool conv = false;
    for (int it = 0; it < maxiter; ++it) {
        const double V = std::exp(x);
        const double u = V / (2 * Vs) * C;
        const double g = eta * V + sig * a * std::asinh(u) - T;
        const double dg = eta * V + sig * a * u / std::sqrt(1 + u * u);
        const double dx = std::clamp(-g / dg, -maxstep, maxstep);
        x += dx;
        if (std::fabs(dx) < tol) {
            conv = true;
            break;
        }
    }
    if (ok) *ok = conv;
    else if (!conv) throw NewtonFail("solve_slip_rate failed to converge");
    return std::exp(x);
}

struct Model {
    eqo_bp8_params p;
    int nf, n2f, injection;
    std::vector<double> K;  // column-major 2nf x 2nf
    std::vector<double> Lx;  // 1D Neumann-SAT Laplacian factor (n2f x n2f, row-major)
    std::vector<double> source, weights, x2;
    std::vector<char> active;
    double tau0[2], eta, lam, q0, WI, S_e;
    int well_cell;
    // cache (src/BP8.jl:104-121)
    std::vector<double> dtau, V2, V3, Vmag, tau2, tau3, Vprev;
    double sig_lowest;
    int64_t floor_hits;
    std::vector<char> floor_nodes;
    int64_t nrhs = 0;

    int state_len() const { return 4 * nf + (injection == 1 ? 1 : 0); }

    void reset_cache() {
        dtau.assign(2 * nf, 0.0);
        V2.assign(nf, 0.0);
        V3.assign(nf, 0.0);
        Vmag.assign(nf, 0.0);
        tau2.assign(nf, 0.0);
        tau3.assign(nf, 0.0);
        Vprev.assign(nf, 1e-12);
        sig_lowest = std::numeric_limits<double>::infinity();
        floor_hits = 0;
        floor_nodes.assign(nf, 0);
    }

    // src/BP8.jl:355-396
    void evaluate(const double* u) {
        const int n2 = 2 * nf;
        for (int i = 0; i < n2; ++i) dtau[i] = 0.0;
        for (int j = 0; j < n2; ++j) {  // column sweep: dtau += K[:, j] * slip[j]
            const double s = u[j];
            if (s == 0.0) continue;
            const double* Kj = &K[(size_t)j * n2];
            for (int i = 0; i < n2; ++i) dtau[i] += Kj[i] * s;
        }
        const double* phi = u + 2 * nf;
        const double* pres = u + 3 * nf;
        for (int i = 0; i < nf; ++i) {
            if (!active[i]) {
                V2[i] = V3[i] = Vmag[i] = 0.0;
                tau2[i] = tau0[0] + dtau[i];
                tau3[i] = tau0[1] + dtau[nf + i];
                continue;
            }
            const double sraw = p.sigma0 - pres[i];
            if (sraw < sig_lowest) sig_lowest = sraw;
            if (sraw < p.sigma_min) {
                floor_hits++;
                floor_nodes[i] = 1;
            }
            const double sig = std::max(sraw, p.sigma_min);
            const double theta = std::exp(phi[i]);
            const double T2 = tau0[0] + dtau[i], T3 = tau0[1] + dtau[nf + i];
            const double T = std::sqrt(T2 * T2 + T3 * T3);
            double v2 = 0.0, v3 = 0.0;
            if (T != 0.0) {
                const double V = slip_rate(T, theta, sig, eta, p.a, p.b, p.D_RS, p.V_star, p.f_star, Vprev[i]);
                v2 = V * T2 / T;
                v3 = V * T3 / T;
            }
            V2[i] = v2;
            V3[i] = v3;
            Vmag[i] = std::sqrt(v2 * v2 + v3 * v3);
            Vprev[i] = std::max(Vmag[i], 1e-30);
            tau2[i] = tau0[0] + dtau[i] - eta * v2;
            tau3[i] = tau0[1] + dtau[nf + i] - eta * v3;
        }
    }

    // src/BP8.jl:403-433
    void rhs(const double* u, double t, double* du) {
        ++nrhs;
        evaluate(u);
        for (int i = 0; i < nf; ++i) {
            du[i] = V2[i];
            du[nf + i] = V3[i];
            du[2 * nf + i] = std::exp(-u[2 * nf + i]) - Vmag[i] / p.D_RS;
        }
        const double* pres = u + 3 * nf;
        double* dp = du + 3 * nf;
        const int n = n2f;
        for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
                double s = 0.0;
                for (int c = 0; c < n; ++c) s += Lx[(size_t)a * n + c] * pres[c + n * b];
                for (int c = 0; c < n; ++c) s += Lx[(size_t)b * n + c] * pres[a + n * c];
                dp[a + n * b] = p.alpha * s;
            }
        if (injection == 0) {
            const double q = t < p.t_off ? q0 : 0.0;
            for (int i = 0; i < nf; ++i) dp[i] += q / (p.beta * p.phi) * source[i];
        } else {
            const double pw = u[4 * nf];
            const double transfer = WI * (pw - pres[well_cell]);
            dp[well_cell] += transfer / S_e;
            const double Qinj = t < p.t_off ? p.Q0 : 0.0;
            du[4 * nf] = (Qinj - transfer) / p.S_well;
        }
    }
};

std::unique_ptr<Model> create(const eqo_bp8_params& p, const double* K, int n2f, int injection) {
    auto mp = std::make_unique<Model>();
    Model& m = *mp;
    m.p = p;
    m.n2f = n2f;
    m.nf = n2f * n2f;
    m.injection = injection;
    if (injection != 0 && injection != 1) throw std::runtime_error("injection must be :gaussian or :peaceman");
    if (n2f < 9) throw std::runtime_error("Omega_f spans fewer than 9 nodes; the pore-pressure operator needs 9");
    const size_t n2 = 2 * (size_t)m.nf;
    m.K.assign(K, K + n2 * n2);
    const double h = 2.0 * p.l_f / (n2f - 1);
    sbp4::Ops1D o = sbp4::build(n2f, h);
    // D2 + Neumann SATs = -H^-1 M along one axis (src/PorePressure.jl:168-175)
    m.Lx = o.D2;
    for (int j = 0; j < n2f; ++j) {
        m.Lx[j] += o.dl[j] / o.H[0];
        m.Lx[(size_t)(n2f - 1) * n2f + j] -= o.dr[j] / o.H[n2f - 1];
    }
    m.x2.resize(n2f);
    for (int i = 0; i < n2f; ++i) m.x2[i] = -p.l_f + i * h;
    m.source.resize(m.nf);
    m.weights.resize(m.nf);
    m.active.assign(m.nf, 1);
    double best = std::numeric_limits<double>::infinity();
    m.well_cell = 0;
    for (int b = 0; b < n2f; ++b)
        for (int a = 0; a < n2f; ++a) {
            const int i = a + n2f * b;
            const double x = m.x2[a], y = m.x2[b];
            m.source[i] = std::exp(-(x * x + y * y) / (2 * p.L_gauss * p.L_gauss)) / (2 * M_PI * p.L_gauss * p.L_gauss);
            m.weights[i] = o.H[a] * o.H[b];
            if (a == 0 || a == n2f - 1 || b == 0 || b == n2f - 1) m.active[i] = 0;
            const double r2 = x * x + y * y;
            if (r2 < best) {
                best = r2;
                m.well_cell = i;
            }
        }
    const double vn = std::sqrt(p.V_init * p.V_init + p.V_zero * p.V_zero);
    m.tau0[0] = p.tau_init * p.V_init / vn;
    m.tau0[1] = p.tau_init * p.V_zero / vn;
    m.eta = p.mu / (2 * p.c_s);
    m.lam = 2 * p.mu * p.nu / (1 - 2 * p.nu);
    m.q0 = p.Q0 / p.L_fwid;
    m.WI = 0.0;
    if (injection == 1) {
        const double re = 0.268 * p.dz;
        if (!(re > p.r_well)) throw std::runtime_error("equivalent radius must exceed r_well");
        m.WI = 2 * M_PI * p.k * p.L_fwid / (p.viscosity * std::log(re / p.r_well));
    }
    m.S_e = p.dz * p.dz * p.L_fwid * p.phi * p.beta;
    m.reset_cache();
    return mp;
}

// Tsitouras (2011) tableau, as published.
constexpr double c2 = 0.161, c3 = 0.327, c4 = 0.9, c5 = 0.9800255409045097;
constexpr double a21 = 0.161;
constexpr double a31 = -0.008480655492356989, a32 = 0.335480655492357;
constexpr double a41 = 2.897153057105493, a42 = -6.359448489975075, a43 = 4.3622954328695815;
constexpr double a51 = 5.325864828439257, a52 = -11.748883564062828, a53 = 7.4955393428898365,
                 a54 = -0.09249506636175525;
constexpr double a61 = 5.86145544294642, a62 = -12.92096931784711, a63 = 8.159367898576159,
                 a64 = -0.071584973281401, a65 = -0.028269050394068383;
constexpr double a71 = 0.09646076681806523, a72 = 0.01, a73 = 0.4798896504144996, a74 = 1.379008574103742,
                 a75 = -3.290069515436081, a76 = 2.324710524099774;
constexpr double bt1 = -0.00178001105222577714, bt2 = -0.0008164344596567469, bt3 = 0.007880878010261995,
                 bt4 = -0.1447110071732629, bt5 = 0.5823571654525552, bt6 = -0.45808210592918697,
                 bt7 = 1.0 / 66.0;

// Tsit5 free interpolant weights b_i(theta).
void interp_weights(double th, double b[7]) {
    const double t2 = th * th;
    b[0] = -1.0530884977290216 * th * (th - 1.3299890189751412) * (t2 - 1.4364028541716351 * th + 0.7139816917074209);
    b[1] = 0.1017 * t2 * (t2 - 2.1966568338249754 * th + 1.2949852507374631);
    b[2] = 2.490627285651252793 * t2 * (t2 - 2.38535645472061657 * th + 1.57803468208092486);
    b[3] = -16.54810288924490272 * (th - 1.21712927295533244) * (th - 0.61620406037800089) * t2;
    b[4] = 47.37952196281928122 * (th - 1.203071208372362603) * (th - 0.658047292653547382) * t2;
    b[5] = -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * t2;
    b[6] = 2.5 * (th - 1) * (th - 0.6) * t2;
}

double rms(const std::vector<double>& v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s / (double)v.size());
}

}  // namespace

extern "C" {

void eqo_bp8_default_params(eqo_bp8_params* p) {
    *p = eqo_bp8_params{32.04e9, 0.25, 2670.0, 3464.0, 0.016, 0.010, 0.5e-3, 1e-6, 0.6, 25.0e6, 14.6e6, 1e-12,
                        1e-20, 0.003, 50.0, 1.0, 1e-8, 0.1, 5e-14, 1e-3, 0.05, 1e-7, 0.05, 400.0, 10.0,
                        100 * 3600.0, 30 * 24 * 3600.0, 1.0e3};
}

int eqo_solve_slip_rate(double T, double theta, double sigma, double eta, double a, double b, double Dc,
                        double V_star, double f_star, double V0, double tol, int maxiter, double maxstep, double* V) {
    bool ok = false;
    *V = slip_rate(T, theta, sigma, eta, a, b, Dc, V_star, f_star, V0, tol, maxiter, maxstep, &ok);
    return ok ? 0 : 1;
}

void* eqo_bp8_create(const eqo_bp8_params* p, const double* K, int n2f, int injection, char* err, size_t errlen) {
    try {
        return create(*p, K, n2f, injection).release();
    } catch (const std::exception& e) {
        if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
        return nullptr;
    }
}

void eqo_bp8_destroy(void* h) { delete static_cast<Model*>(h); }
int eqo_bp8_state_length(void* h) { return static_cast<Model*>(h)->state_len(); }
void eqo_bp8_reset_cache(void* h) { static_cast<Model*>(h)->reset_cache(); }

// src/BP8.jl:337-345 with src/RateStateFriction.jl:127-130
int eqo_bp8_initial_state(void* h, double* u) {
    Model& m = *static_cast<Model*>(h);
    const eqo_bp8_params& p = m.p;
    std::fill(u, u + m.state_len(), 0.0);
    const double Vm = std::sqrt(p.V_init * p.V_init + p.V_zero * p.V_zero);
    const double f = (p.tau_init - m.eta * Vm) / p.sigma0;
    const double theta0 = (p.D_RS / p.V_star) * std::exp((p.a * std::log(2 * p.V_star * std::sinh(f / p.a) / Vm) - p.f_star) / p.b);
    for (int i = 0; i < m.nf; ++i) u[2 * m.nf + i] = std::log(theta0);
    return 0;
}

int eqo_bp8_evaluate(void* h, const double* u, double, double* V2, double* V3, double* Vmag, double* tau2,
                     double* tau3) {
    Model& m = *static_cast<Model*>(h);
    try {
        m.evaluate(u);
    } catch (const NewtonFail&) {
        return 1;
    }
    const size_t n = m.nf * sizeof(double);
    if (V2) std::memcpy(V2, m.V2.data(), n);
    if (V3) std::memcpy(V3, m.V3.data(), n);
    if (Vmag) std::memcpy(Vmag, m.Vmag.data(), n);
    if (tau2) std::memcpy(tau2, m.tau2.data(), n);
    if (tau3) std::memcpy(tau3, m.tau3.data(), n);
    return 0;
}

int eqo_bp8_rhs(void* h, const double* u, double t, double* du) {
    try {
        static_cast<Model*>(h)->rhs(u, t, du);
    } catch (const NewtonFail&) {
        return 1;
    }
    return 0;
}

int eqo_bp8_Ap(void* h, double* Ap) {
    Model& m = *static_cast<Model*>(h);
    const int n = m.n2f, nf = m.nf;
    std::fill(Ap, Ap + (size_t)nf * nf, 0.0);
    for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a) {
            const int row = a + n * b;
            for (int c = 0; c < n; ++c) {
                Ap[(size_t)row * nf + c + n * b] += m.p.alpha * m.Lx[(size_t)a * n + c];
                Ap[(size_t)row * nf + a + n * c] += m.p.alpha * m.Lx[(size_t)b * n + c];
            }
        }
    return 0;
}

int eqo_bp8_source(void* h, double* src, double* weights) {
    Model& m = *static_cast<Model*>(h);
    if (src) std::memcpy(src, m.source.data(), m.nf * sizeof(double));
    if (weights) std::memcpy(weights, m.weights.data(), m.nf * sizeof(double));
    return 0;
}

// run_bp8 (src/BP8.jl:442-470): Tsit5 with the integrator defaults of OrdinaryDiffEq.
int eqo_bp8_run(void* h, double t0, double tf, double saveat, double reltol, int nsave_max, double* ts, double* us,
                double* vmax, int* nsaved, int64_t* naccept, int64_t* nreject, int64_t* nrhs) {
    Model& m = *static_cast<Model*>(h);
    const int n = m.state_len(), nf = m.nf;
    std::vector<double> abstol(n);
    for (int i = 0; i < n; ++i)
        abstol[i] = i < 2 * nf ? 1e-14 : (i < 3 * nf ? 1e-10 : 1e-3);
    std::vector<double> u(n), uprev(n), tmp(n), k[7], utilde(n), atmp(n);
    for (auto& v : k) v.resize(n);
    eqo_bp8_initial_state(h, u.data());
    m.nrhs = 0;
    int saved = 0;
    int64_t acc = 0, rej = 0;
    auto save = [&](double t, const double* uu) {
        if (saved >= nsave_max) return;
        ts[saved] = t;
        std::memcpy(us + (size_t)saved * n, uu, n * sizeof(double));
        ++saved;
    };
    try {
        double t = t0;
        m.rhs(u.data(), t, k[0].data());
        // initial dt (OrdinaryDiffEq ode_determine_initdt, non-stiff path)
        const double dtmax = tf - t0;
        double dt;
        {
            std::vector<double> sk(n), f1(n), u1(n);
            for (int i = 0; i < n; ++i) sk[i] = abstol[i] + std::fabs(u[i]) * reltol;
            for (int i = 0; i < n; ++i) tmp[i] = u[i] / sk[i];
            const double d0 = rms(tmp);
            for (int i = 0; i < n; ++i) tmp[i] = k[0][i] / sk[i];
            const double d1 = rms(tmp);
            const double smalldt = 1e-6;
            double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? smalldt : 0.01 * (d0 / d1);
            dt0 = std::min(dt0, dtmax);
            for (int i = 0; i < n; ++i) u1[i] = u[i] + dt0 * k[0][i];
            m.rhs(u1.data(), t + dt0, f1.data());
            for (int i = 0; i < n; ++i) tmp[i] = (f1[i] - k[0][i]) / sk[i];
            const double d2 = rms(tmp) / dt0;
            const double md = std::max(d1, d2);
            const double dt1 = md <= 1e-15 ? std::max(1e-6, dt0 * 1e-3) : std::pow(10.0, -(2 + std::log10(md)) / 5.0);
            dt = std::min(std::min(100 * dt0, dt1), dtmax);
        }
        double qold = 1e-4, q11 = 0.0;
        const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2, qmax = 10.0, gamma = 0.9;
        save(t, u.data());
        double next_save = t0 + saveat;
        int64_t guard = 0;
        while (t < tf) {
            // tstops = [t_off]; a stop outside (t, tf) is ignored and tf always ends the run
            const double tstop = (m.p.t_off > t && m.p.t_off < tf) ? m.p.t_off : tf;
            const bool clipped = dt >= tstop - t;
            if (clipped) dt = tstop - t;
            uprev = u;
            const double* k1 = k[0].data();
            for (int i = 0; i < n; ++i) tmp[i] = uprev[i] + dt * a21 * k1[i];
            m.rhs(tmp.data(), t + c2 * dt, k[1].data());
            for (int i = 0; i < n; ++i) tmp[i] = uprev[i] + dt * (a31 * k1[i] + a32 * k[1][i]);
            m.rhs(tmp.data(), t + c3 * dt, k[2].data());
            for (int i = 0; i < n; ++i) tmp[i] = uprev[i] + dt * (a41 * k1[i] + a42 * k[1][i] + a43 * k[2][i]);
            m.rhs(tmp.data(), t + c4 * dt, k[3].data());
            for (int i = 0; i < n; ++i)
                tmp[i] = uprev[i] + dt * (a51 * k1[i] + a52 * k[1][i] + a53 * k[2][i] + a54 * k[3][i]);
            m.rhs(tmp.data(), t + c5 * dt, k[4].data());
            for (int i = 0; i < n; ++i)
                tmp[i] = uprev[i] + dt * (a61 * k1[i] + a62 * k[1][i] + a63 * k[2][i] + a64 * k[3][i] + a65 * k[4][i]);
            m.rhs(tmp.data(), t + dt, k[5].data());
            for (int i = 0; i < n; ++i)
                u[i] = uprev[i] + dt * (a71 * k1[i] + a72 * k[1][i] + a73 * k[2][i] + a74 * k[3][i] + a75 * k[4][i] +
                                        a76 * k[5][i]);
            m.rhs(u.data(), t + dt, k[6].data());
            for (int i = 0; i < n; ++i) {
                utilde[i] = dt * (bt1 * k1[i] + bt2 * k[1][i] + bt3 * k[2][i] + bt4 * k[3][i] + bt5 * k[4][i] +
                                  bt6 * k[5][i] + bt7 * k[6][i]);
                atmp[i] = utilde[i] / (abstol[i] + std::max(std::fabs(uprev[i]), std::fabs(u[i])) * reltol);
            }
            const double EEst = rms(atmp);
            double q;
            if (EEst == 0.0) q = 1.0 / qmax;
            else {
                q11 = std::pow(EEst, beta1);
                q = q11 / std::pow(qold, beta2);
                q = std::max(1.0 / qmax, std::min(1.0 / qmin, q / gamma));
            }
            if (EEst <= 1.0) {
                const double tnew = clipped ? tstop : t + dt;
                while (saved < nsave_max && next_save <= tnew + 1e-9 * saveat) {
                    if (std::fabs(next_save - tnew) <= 1e-9 * saveat) {
                        save(next_save, u.data());
                    } else {
                        const double th = (next_save - t) / dt;
                        double b[7];
                        interp_weights(th, b);
                        for (int i = 0; i < n; ++i) {
                            double s = 0.0;
                            for (int j = 0; j < 7; ++j) s += b[j] * k[j][i];
                            tmp[i] = uprev[i] + dt * s;
                        }
                        save(next_save, tmp.data());
                    }
                    next_save = t0 + (saved) * saveat;
                }
                t = tnew;
                // save_end: the final state is saved even when tf is off the saveat grid
                if (t >= tf && saved < nsave_max && std::fabs(ts[saved - 1] - tf) > 1e-9 * saveat) save(tf, u.data());
                std::swap(k[0], k[6]);
                ++acc;
                qold = std::max(EEst, 1e-4);
                dt = dt / q;
            } else {
                u = uprev;
                ++rej;
                dt = dt / std::min(1.0 / qmin, q11 / gamma);
            }
            if (++guard > 100000000) break;
        }
    } catch (const NewtonFail&) {
        if (nsaved) *nsaved = saved;
        return 1;
    }
    if (vmax) {  // V_max at every saved state, evaluated afterwards as write_outputs does (src/BP8.jl:628)
        try {
            for (int j = 0; j < saved; ++j) {
                m.evaluate(us + (size_t)j * n);
                double mx = 0.0;
                for (int i = 0; i < nf; ++i) mx = std::max(mx, m.Vmag[i]);
                vmax[j] = mx;
            }
        } catch (const NewtonFail&) {
            if (nsaved) *nsaved = saved;
            return 1;
        }
    }
    if (nsaved) *nsaved = saved;
    if (naccept) *naccept = acc;
    if (nreject) *nreject = rej;
    if (nrhs) *nrhs = m.nrhs;
    return 0;
}

}  // extern "C"
