// BP8 time loop on the device (C ABI in include/eqd.h).
//
// State u = [s2; s3; ln theta; p (; p_well)] (src/BP8.jl:262,362-364) for the nf fault
// nodes of Omega_f. One right-hand-side evaluation is one kernel: a wavefront owns a
// fault node i and computes
//   dtau2_i, dtau3_i   rows i and nf+i of K times the slip (K kept row-major in HBM, so
//                      both rows stream coalesced, 16 B per lane per load),
//   V                  the rate-state Newton solve in ln V (src/RateStateFriction.jl:77-96),
//   (A_p p)_i          the pore-pressure operator as two 1D SBP rows (src/PorePressure.jl:36-43),
// and writes the four derivative entries of node i plus, for a Tsit5 stage, the same
// entries of the next stage's input (which only need this node's earlier stage values).
// A step is six stage kernels, a one-block controller (error norm, PI step control,
// accept/reject, saveat bookkeeping) and a commit kernel (saveat interpolation, FSAL
// copy, first stage input of the next step). Several steps are captured in one HIP
// graph; the host replays it and polls the controller block between replays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/eqd.h"
#include "eqd_internal.h"
#include "fft_ks.h"
#include "host_common.h"

using namespace eqd;
using namespace eqd::host;

namespace eqd {
namespace bp8 {

// Tsitouras (2011) Tsit5 tableau and the embedded error weights (b - b_hat), as published.
__constant__ double kA[8][7] = {
    {0}, {0}, {0.161},
    {-0.008480655492356989, 0.335480655492357},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774}};
__constant__ double kC[8] = {0, 0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
__constant__ double kBt[7] = {-0.00178001105222577714, -0.0008164344596567469, 0.007880878010261995,
                              -0.1447110071732629,     0.5823571654525552,     -0.45808210592918697,
                              1.0 / 66.0};
constexpr double kA21 = 0.161;

// Tsit5 free interpolant: u(t + theta dt) = uprev + dt sum_j b_j(theta) k_j.
__device__ inline void interp_weights(double th, double b[7]) {
    const double t2 = th * th;
    b[0] = -1.0530884977290216 * th * (th - 1.3299890189751412) * (t2 - 1.4364028541716351 * th + 0.7139816917074209);
    b[1] = 0.1017 * t2 * (t2 - 2.1966568338249754 * th + 1.2949852507374631);
    b[2] = 2.490627285651252793 * t2 * (t2 - 2.38535645472061657 * th + 1.57803468208092486);
    b[3] = -16.54810288924490272 * (th - 1.21712927295533244) * (th - 0.61620406037800089) * t2;
    b[4] = 47.37952196281928122 * (th - 1.203071208372362603) * (th - 0.658047292653547382) * t2;
    b[5] = -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * t2;
    b[6] = 2.5 * (th - 1) * (th - 0.6) * t2;
}

struct Consts {
    int nf, n2f, n2, nstate, injection, well_cell;
    int lw;  // band width of the 1D pore-pressure factor rows (see band_start)
    double tau0_2, tau0_3, eta, a, b, Dc, Vs, fs, sigma0, sigma_min, alpha, D_RS;
    double src_rate;  // q0 / (beta phi): Gaussian source amplitude while injecting
    double t_off, Q0, WI, S_e, S_well;
};

// Integrator state, written only by the one-block controller (and the host at start).
struct Ctl {
    double t, dt, qold, q11, t0, tf, saveat, reltol, tstop;
    double t_step, dt_step, t_new, EEst;
    long long naccept, nreject, nrhs, nsaved, nsave_max, maxiters, iters;
    int done, retcode, accepted, clipped, skip_commit, save_first, save_count, end_save;
};

struct Dev {
    Consts c;
    const double* KR;      // K row-major, 2nf x 2nf
    const double* Tk;      // Toeplitz K: 4 blocks per separation (see k_toeplitz_kernel), or null
    double* ks;            // with Tk: K s partials, ks_split x nf x 2 (k_toeplitz_ks, or one chunk from the FFT)
    int ks_split;          // chunks of ks the stage kernel adds (0: it forms K s from Tk itself)
    const double* Lx;      // n2f x lw: band of each row of the 1D pore-pressure factor (alpha not applied)
    const double* src;     // nf
    const unsigned char* active;
    double* V2;
    double* V3;
    double* Vm;
    double* tau2;
    double* tau3;
    double* Vprev;
    double* sig_low;
    long long* floor_hits;
    unsigned char* floor_mask;
    int* fail;
    double* k;             // 7 x nstate (k1 .. k7)
    double* uprev;
    double* unew;
    double* partial;       // nf: per-node sum of squared scaled errors
    Ctl* ctl;
    double* saves;         // nsave_max rows of ldS (>= nstate, even: rows stay 16-byte aligned)
    double* save_t;        // nsave_max
    int ldS;
};

// First column of row r's band in the n x lw banded form of the 1D pore-pressure factor (the
// SBP-4 D2 rows and their Neumann SATs reach at most 5 columns from the diagonal).
__host__ __device__ inline int band_start(int r, int n, int lw) {
    const int s = r - (lw - 1) / 2;
    return s < 0 ? 0 : (s > n - lw ? n - lw : s);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Newton in x = ln V on g(V) = eta V + sigma a asinh(V C / (2 V*)) - T (src/RateStateFriction.jl:77-96):
// step clamped to +-5, converged when |dx| < 1e-12, at most 50 iterations.
__device__ inline bool slip_rate(double T, double theta, double sig, const Consts& c, double V0, double& V) {
    const double C = exp((c.fs + c.b * log(c.Vs * theta / c.Dc)) / c.a);
    double x = log(V0);
    for (int it = 0; it < 50; ++it) {
        const double v = exp(x);
        const double u = v / (2 * c.Vs) * C;
        const double g = c.eta * v + sig * c.a * asinh(u) - T;
        const double dg = c.eta * v + sig * c.a * u / sqrt(1 + u * u);
        const double dx = fmin(fmax(-g / dg, -5.0), 5.0);
        x += dx;
        if (fabs(dx) < 1e-12) {
            V = exp(x);
            return true;
        }
    }
    V = exp(x);
    return false;
}

__device__ inline double abstol_of(const Consts& c, int e) {
    return e < 2 * c.nf ? 1e-14 : (e < 3 * c.nf ? 1e-10 : 1e-3);
}

// One right-hand-side evaluation at state y (src/BP8.jl:355-433), a wavefront per node.
// stage 0: plain rhs! into kout at time t_fixed.
// stage 2..6: Tsit5 stage s of the current step (time t + c_s dt from the controller):
//             kout = k_s, and ynext = the input of stage s+1 (u_new for s = 6).
// stage 7:    k_7 = f(u_new) and the node's sum of squared scaled error estimates.
__global__ __launch_bounds__(256) void k_bp8_stage(Dev d, int stage, const double* __restrict__ y,
                                                   double* __restrict__ kout, double* __restrict__ ynext,
                                                   double t_fixed) {
    const Consts& c = d.c;
    if (stage != 0 && d.ctl->done) return;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= c.nf) return;
    const double dt = stage != 0 ? d.ctl->dt : 0.0;
    const double t = stage != 0 ? d.ctl->t + kC[stage] * dt : t_fixed;

    // dtau = K s: rows i and nf + i
    double a2 = 0.0, a3 = 0.0;
    if (d.Tk && d.ks_split) {
        // Toeplitz K: k_toeplitz_ks left one partial per source chunk; sum them in chunk order
        for (int q = lane; q < d.ks_split; q += 64) {
            a2 += d.ks[((size_t)q * c.nf + i) * 2];
            a3 += d.ks[((size_t)q * c.nf + i) * 2 + 1];
        }
    } else if (d.Tk) {
        // Toeplitz K on a small patch: row i of K is a window of the separation kernel (key =
        // base - aj - W2 bj), read here directly (the kernel is L2-resident; a separate launch
        // would cost more than the product)
        const int n2 = c.n2f, W2 = 2 * n2 - 1;
        const int base = (i % n2 + n2 - 1) + W2 * (i / n2 + n2 - 1);
        const float inv = 1.0f / (float)n2;  // exact source row for j < nf <= 48^2
        const double4* T = reinterpret_cast<const double4*>(d.Tk);
        for (int j = lane; j < c.nf; j += 64) {
            const int bj = (int)(((float)j + 0.5f) * inv), aj = j - bj * n2;
            const double4 k = T[base - aj - W2 * bj];
            const double s2 = y[j], s3 = y[c.nf + j];
            a2 += k.x * s2 + k.z * s3;
            a3 += k.y * s2 + k.w * s3;
        }
    } else {
        const int nh = c.n2 >> 1;
        const double2* r2 = reinterpret_cast<const double2*>(d.KR + (size_t)i * c.n2);
        const double2* r3 = reinterpret_cast<const double2*>(d.KR + (size_t)(c.nf + i) * c.n2);
        const double2* sv = reinterpret_cast<const double2*>(y);
        for (int j = lane; j < nh; j += 64) {
            const double2 s = sv[j], k2 = r2[j], k3 = r3[j];
            a2 += k2.x * s.x + k2.y * s.y;
            a3 += k3.x * s.x + k3.y * s.y;
        }
    }
    const double dtau2 = wave_sum(a2), dtau3 = wave_sum(a3);

    // A_p p at node i = (a, b): alpha (sum_c Lx[a][c] p[c + n b] + sum_c Lx[b][c] p[a + n c]), the
    // rows' nonzeros from their bands
    const int n = c.n2f, ia = i % n, ib = i / n;
    const double* P = y + 3 * c.nf;
    double ap = 0.0;
    if (lane < c.lw) {
        const int sa = band_start(ia, n, c.lw), sb = band_start(ib, n, c.lw);
        ap = d.Lx[ia * c.lw + lane] * P[sa + lane + n * ib] + d.Lx[ib * c.lw + lane] * P[ia + n * (sb + lane)];
    }
    ap = c.alpha * wave_sum(ap);

    // rate-state slip rate (every lane computes the same scalar path)
    const double phi = y[2 * c.nf + i], p = P[i];
    double v2 = 0.0, v3 = 0.0, vm = 0.0, t2, t3;
    const double T2 = c.tau0_2 + dtau2, T3 = c.tau0_3 + dtau3;
    if (!d.active[i]) {
        t2 = T2;
        t3 = T3;
    } else {
        const double sraw = c.sigma0 - p;
        const double sig = fmax(sraw, c.sigma_min);
        const double theta = exp(phi);
        const double T = sqrt(T2 * T2 + T3 * T3);
        if (T != 0.0) {
            double V;
            const bool ok = slip_rate(T, theta, sig, c, d.Vprev[i], V);
            if (!ok && lane == 0) d.fail[0] = 1;
            v2 = V * T2 / T;
            v3 = V * T3 / T;
        }
        vm = sqrt(v2 * v2 + v3 * v3);
        t2 = T2 - c.eta * v2;
        t3 = T3 - c.eta * v3;
        if (lane == 0) {
            d.Vprev[i] = fmax(vm, 1e-30);
            if (sraw < d.sig_low[i]) d.sig_low[i] = sraw;
            if (sraw < c.sigma_min) {
                d.floor_hits[i] += 1;
                d.floor_mask[i] = 1;
            }
        }
    }
    double f[5];
    f[0] = v2;
    f[1] = v3;
    f[2] = exp(-phi) - vm / c.D_RS;
    f[3] = ap;
    int ne = 4;
    int idx[5] = {i, c.nf + i, 2 * c.nf + i, 3 * c.nf + i, 4 * c.nf};
    if (c.injection == 0) {
        f[3] += (t < c.t_off ? c.src_rate : 0.0) * d.src[i];
    } else if (i == c.well_cell) {
        const double pw = y[4 * c.nf];
        const double transfer = c.WI * (pw - p);
        f[3] += transfer / c.S_e;
        f[4] = ((t < c.t_off ? c.Q0 : 0.0) - transfer) / c.S_well;
        ne = 5;
    }
    if (lane == 0) {
        d.V2[i] = v2;
        d.V3[i] = v3;
        d.Vm[i] = vm;
        d.tau2[i] = t2;
        d.tau3[i] = t3;
    }
    if (lane < ne) {
        const int e = idx[lane];
        double fe = f[0];
#pragma unroll
        for (int q = 1; q < 5; ++q)
            if (lane == q) fe = f[q];
        kout[e] = fe;
        if (stage >= 2 && stage <= 6) {
            double s = 0.0;
            for (int j = 0; j < stage - 1; ++j) s += kA[stage + 1][j] * d.k[(size_t)j * c.nstate + e];
            s += kA[stage + 1][stage - 1] * fe;
            ynext[e] = d.uprev[e] + dt * s;
        } else if (stage == 7) {
            double s = 0.0;
            for (int j = 0; j < 6; ++j) s += kBt[j] * d.k[(size_t)j * c.nstate + e];
            s += kBt[6] * fe;
            const double ut = dt * s;
            const double sc = abstol_of(c, e) + fmax(fabs(d.uprev[e]), fabs(d.unew[e])) * d.ctl->reltol;
            const double r = ut / sc;
            double r2 = r * r;
            // lanes 0..ne-1 hold the node's entries: sum them in entry order
            double tot = 0.0;
            for (int q = 0; q < ne; ++q) tot += __shfl(r2, q, 64);
            if (lane == 0) d.partial[i] = tot;
        }
    }
}

// The stage of a model whose K s comes precomputed per node (ks_split > 0: from k_toeplitz_ks or
// the FFT convolution):
// one node per thread, where the wave-per-node k_bp8_stage would repeat each node's scalar
// Newton on 64 lanes. The per-node arithmetic is k_bp8_stage's except the order of two sums
// (the K s chunks and the A_p row are added here one after another instead of across lanes).
__global__ __launch_bounds__(256) void k_bp8_stage_nodes(Dev d, int stage, const double* __restrict__ y,
                                                         double* __restrict__ kout, double* __restrict__ ynext,
                                                         double t_fixed) {
    const Consts& c = d.c;
    if (stage != 0 && d.ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.nf) return;
    const double dt = stage != 0 ? d.ctl->dt : 0.0;
    const double t = stage != 0 ? d.ctl->t + kC[stage] * dt : t_fixed;

    double dtau2 = 0.0, dtau3 = 0.0;
    for (int q = 0; q < d.ks_split; ++q) {
        dtau2 += d.ks[((size_t)q * c.nf + i) * 2];
        dtau3 += d.ks[((size_t)q * c.nf + i) * 2 + 1];
    }
    const int n = c.n2f, ia = i % n, ib = i / n;
    const double* P = y + 3 * c.nf;
    double ap = 0.0;
    {
        const int sa = band_start(ia, n, c.lw), sb = band_start(ib, n, c.lw);
        for (int l = 0; l < c.lw; ++l)
            ap += d.Lx[ia * c.lw + l] * P[sa + l + n * ib] + d.Lx[ib * c.lw + l] * P[ia + n * (sb + l)];
    }
    ap = c.alpha * ap;

    const double phi = y[2 * c.nf + i], p = P[i];
    double v2 = 0.0, v3 = 0.0, vm = 0.0, t2, t3;
    const double T2 = c.tau0_2 + dtau2, T3 = c.tau0_3 + dtau3;
    if (!d.active[i]) {
        t2 = T2;
        t3 = T3;
    } else {
        const double sraw = c.sigma0 - p;
        const double sig = fmax(sraw, c.sigma_min);
        const double theta = exp(phi);
        const double T = sqrt(T2 * T2 + T3 * T3);
        if (T != 0.0) {
            double V;
            if (!slip_rate(T, theta, sig, c, d.Vprev[i], V)) d.fail[0] = 1;
            v2 = V * T2 / T;
            v3 = V * T3 / T;
        }
        vm = sqrt(v2 * v2 + v3 * v3);
        t2 = T2 - c.eta * v2;
        t3 = T3 - c.eta * v3;
        d.Vprev[i] = fmax(vm, 1e-30);
        if (sraw < d.sig_low[i]) d.sig_low[i] = sraw;
        if (sraw < c.sigma_min) {
            d.floor_hits[i] += 1;
            d.floor_mask[i] = 1;
        }
    }
    double f[5];
    f[0] = v2;
    f[1] = v3;
    f[2] = exp(-phi) - vm / c.D_RS;
    f[3] = ap;
    f[4] = 0.0;
    int ne = 4;
    const int idx[5] = {i, c.nf + i, 2 * c.nf + i, 3 * c.nf + i, 4 * c.nf};
    if (c.injection == 0) {
        f[3] += (t < c.t_off ? c.src_rate : 0.0) * d.src[i];
    } else if (i == c.well_cell) {
        const double pw = y[4 * c.nf];
        const double transfer = c.WI * (pw - p);
        f[3] += transfer / c.S_e;
        f[4] = ((t < c.t_off ? c.Q0 : 0.0) - transfer) / c.S_well;
        ne = 5;
    }
    d.V2[i] = v2;
    d.V3[i] = v3;
    d.Vm[i] = vm;
    d.tau2[i] = t2;
    d.tau3[i] = t3;
    double tot = 0.0;  // stage 7: the node's squared scaled error estimates, in entry order
    for (int q = 0; q < ne; ++q) {
        const int e = idx[q];
        const double fe = f[q];
        kout[e] = fe;
        if (stage >= 2 && stage <= 6) {
            double s = 0.0;
            for (int j = 0; j < stage - 1; ++j) s += kA[stage + 1][j] * d.k[(size_t)j * c.nstate + e];
            s += kA[stage + 1][stage - 1] * fe;
            ynext[e] = d.uprev[e] + dt * s;
        } else if (stage == 7) {
            double s = 0.0;
            for (int j = 0; j < 6; ++j) s += kBt[j] * d.k[(size_t)j * c.nstate + e];
            s += kBt[6] * fe;
            const double ut = dt * s;
            const double sc = abstol_of(c, e) + fmax(fabs(d.uprev[e]), fabs(d.unew[e])) * d.ctl->reltol;
            const double r = ut / sc;
            tot += r * r;
        }
    }
    if (stage == 7) d.partial[i] = tot;
}

// Controller (one block): RMS error norm, PI step control with the OrdinaryDiffEq defaults,
// accept/reject, saveat bookkeeping and the next step size clipped to the next stop.
__global__ __launch_bounds__(256) void k_bp8_control(Dev d) {
    Ctl& s = *d.ctl;
    __shared__ double red[256];
    __shared__ int skip;
    if (threadIdx.x == 0) {
        skip = s.done;
        if (s.done) s.skip_commit = 1;
    }
    __syncthreads();
    if (skip) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < d.c.nf; i += 256) acc += d.partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2, qmax = 10.0, gamma = 0.9;
    const double EEst = sqrt(red[0] / d.c.nstate);
    s.EEst = EEst;
    s.nrhs += 6;
    s.iters += 1;
    s.skip_commit = 0;
    s.end_save = 0;
    s.save_count = 0;
    double q;
    if (EEst == 0.0) {
        q = 1.0 / qmax;
    } else {
        s.q11 = pow(EEst, beta1);
        q = s.q11 / pow(s.qold, beta2);
        q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
    }
    s.t_step = s.t;
    s.dt_step = s.dt;
    if (d.fail[0]) {
        s.done = 1;
        s.retcode = 2;
        s.accepted = 0;
        s.skip_commit = 1;
        return;
    }
    if (EEst <= 1.0) {
        const double tnew = s.clipped ? s.tstop : s.t + s.dt;
        s.t_new = tnew;
        s.accepted = 1;
        s.naccept += 1;
        long long cnt = 0;
        while (s.nsaved + cnt < s.nsave_max && s.t0 + (double)(s.nsaved + cnt) * s.saveat <= tnew + 1e-9 * s.saveat)
            ++cnt;
        s.save_first = (int)s.nsaved;
        s.save_count = (int)cnt;
        s.nsaved += cnt;
        s.t = tnew;
        s.qold = fmax(EEst, 1e-4);
        s.dt = s.dt / q;
        if (s.t >= s.tf) {
            // the end point is always saved (it is on the saveat grid in the usual case)
            const double last = s.t0 + (double)(s.nsaved - 1) * s.saveat;
            if (s.nsaved < s.nsave_max && fabs(last - s.tf) > 1e-9 * s.saveat) {
                s.end_save = 1;
                s.nsaved += 1;
            }
        }
    } else {
        s.accepted = 0;
        s.nreject += 1;
        s.dt = s.dt / fmin(1.0 / qmin, s.q11 / gamma);
    }
    if (s.t >= s.tf) {
        s.done = 1;
        s.retcode = 0;
        return;
    }
    if (s.iters >= s.maxiters) {
        s.done = 1;
        s.retcode = 1;
        return;
    }
    const double tstop = (d.c.t_off > s.t && d.c.t_off < s.tf) ? d.c.t_off : s.tf;
    s.tstop = tstop;
    s.clipped = s.dt >= tstop - s.t;
    if (s.clipped) s.dt = tstop - s.t;
}

// Commit: saveat outputs of an accepted step (free interpolant, or u_new on a grid hit),
// FSAL (uprev <- u_new, k1 <- k7), then the first stage input y2 = uprev + dt a21 k1.
// init = 1 only builds y2 (start of integration).
__global__ __launch_bounds__(256) void k_bp8_commit(Dev d, double* __restrict__ y2, int init) {
    const Ctl& s = *d.ctl;
    if (!init && s.skip_commit) return;
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int ns = d.c.nstate;
    if (e >= ns) return;
    double* k1 = d.k;
    double up = d.uprev[e];
    if (!init && s.accepted) {
        const double un = d.unew[e];
        for (int q = 0; q < s.save_count; ++q) {
            const long long id = s.save_first + q;
            const double ts = s.t0 + (double)id * s.saveat;
            double v;
            if (fabs(ts - s.t_new) <= 1e-9 * s.saveat) {
                v = un;
            } else {
                double b[7];
                interp_weights((ts - s.t_step) / s.dt_step, b);
                double acc = 0.0;
                for (int j = 0; j < 7; ++j) acc += b[j] * d.k[(size_t)j * ns + e];
                v = up + s.dt_step * acc;
            }
            d.saves[(size_t)id * d.ldS + e] = v;
            if (e == 0) d.save_t[id] = ts;
        }
        if (s.end_save) {
            d.saves[(size_t)(s.nsaved - 1) * d.ldS + e] = un;
            if (e == 0) d.save_t[s.nsaved - 1] = s.tf;
        }
        up = un;
        d.uprev[e] = un;
        k1[e] = d.k[(size_t)6 * ns + e];
    }
    y2[e] = up + s.dt * kA21 * k1[e];
}

// Row-major copy of a column-major device K.
__global__ void k_transpose(const double* __restrict__ K, long long ldK, int n, double* __restrict__ KR) {
    __shared__ double tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int i = bx + threadIdx.x, j = by + r;  // element K(i, j) = K[i + j ldK]
        if (i < n && j < n) tile[r][threadIdx.x] = K[(size_t)i + (size_t)j * ldK];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int i = bx + r, j = by + threadIdx.x;  // KR[i n + j] = K(i, j)
        if (i < n && j < n) KR[(size_t)i * n + j] = tile[threadIdx.x][r];
    }
}

// K s of a Toeplitz K (src/FaultResponse.jl:378-390) from its separation kernel: K's four nf x nf
// blocks depend on the node separation only, so the contribution of source x3 row bj to target
// x3 row ib is a 1D correlation of the kernel row (ib - bj) with the slip row bj. Block
// (ib, chunk) stages those kernel rows (2 n2f - 1 entries of 4 blocks) and slip rows in LDS, R
// rows of its chunk at a time with one thread per (row slot, target node), and leaves its partial
// sums of target row ib in ks[chunk][node][2]; the stage kernel adds the chunks in order. No row
// of K is read (at nf = 6561: 41 MB of L2 reads per evaluation instead of 2.8 GB).
//
// Threads per block of k_toeplitz_ks, and the source rows a block works on at once: as many
// as its threads cover with one thread per target node of the row.
constexpr int kKsThreads = 256;
__host__ __device__ inline int ks_slots(int n2) { return n2 <= kKsThreads ? kKsThreads / n2 : 1; }

__global__ __launch_bounds__(kKsThreads) void k_toeplitz_ks(Dev d, int stage, const double* __restrict__ y) {
    const Consts& c = d.c;
    if (stage != 0 && d.ctl->done) return;
    extern __shared__ __attribute__((aligned(32))) double ks_lds[];
    const int n2 = c.n2f, W2 = 2 * n2 - 1, nf = c.nf;
    // R source rows per round (row slot r = thread / T), T targets per pass
    const int R = ks_slots(n2), T = n2 <= kKsThreads ? n2 : kKsThreads;
    double4* Trow = reinterpret_cast<double4*>(ks_lds);                 // R x W2
    double2* srow = reinterpret_cast<double2*>(Trow + (size_t)R * W2);   // R x n2
    double2* red = reinterpret_cast<double2*>(ks_lds);                   // R x T, after the rounds
    const int ib = blockIdx.x, chunk = blockIdx.y;
    const int per = (n2 + d.ks_split - 1) / d.ks_split;
    const int b0 = chunk * per, b1 = min(n2, b0 + per);
    const double4* Tk = reinterpret_cast<const double4*>(d.Tk);
    const int t = threadIdx.x, r = t / T, it = t - r * T;
    for (int ia0 = 0; ia0 < n2; ia0 += T) {
        const int ia = ia0 + it;
        double a2 = 0.0, a3 = 0.0;
        for (int rb = b0; rb < b1; rb += R) {
            __syncthreads();  // the previous round's readers are done
            for (int q = t; q < R * W2; q += blockDim.x) {
                const int sr = q / W2, bj = rb + sr;
                if (bj < b1) Trow[q] = Tk[(size_t)W2 * (ib - bj + n2 - 1) + (q - sr * W2)];
            }
            for (int q = t; q < R * n2; q += blockDim.x) {
                const int sr = q / n2, bj = rb + sr, aj = q - sr * n2;
                if (bj < b1) srow[q] = make_double2(y[aj + (size_t)n2 * bj], y[nf + aj + (size_t)n2 * bj]);
            }
            __syncthreads();
            if (r < R && rb + r < b1 && ia < n2) {
                const double4* Tr = Trow + (size_t)r * W2 + ia + n2 - 1;
                const double2* sr = srow + (size_t)r * n2;
                for (int aj = 0; aj < n2; ++aj) {
                    const double4 k = Tr[-aj];
                    const double2 sv = sr[aj];
                    a2 += k.x * sv.x + k.z * sv.y;
                    a3 += k.y * sv.x + k.w * sv.y;
                }
            }
        }
        // the row slots' sums, added in slot order
        __syncthreads();
        if (r < R) red[t] = make_double2(a2, a3);
        __syncthreads();
        if (r == 0 && ia < n2) {
            double s2 = 0.0, s3 = 0.0;
            for (int q = 0; q < R; ++q) {
                const double2 v = red[q * T + it];
                s2 += v.x;
                s3 += v.y;
            }
            const size_t o = ((size_t)chunk * nf + ia + (size_t)n2 * ib) * 2;
            d.ks[o] = s2;
            d.ks[o + 1] = s3;
        }
    }
}

// Separation kernel of a Toeplitz K (column-major K, leading dimension ldK): for node
// separation (da, db), key = (da + n - 1) + (2n - 1)(db + n - 1), Tk[4 key + c] holds the
// entries K[i, j] (c = 0), K[nf + i, j] (1), K[i, nf + j] (2), K[nf + i, nf + j] (3) of any
// pair with i - j = (da, db); the pair taken is i = (max(da, 0), max(db, 0)).
__global__ void k_toeplitz_kernel(const double* __restrict__ K, long long ldK, int n, double* __restrict__ Tk) {
    const int W = 2 * n - 1;
    const int key = blockIdx.x * blockDim.x + threadIdx.x;
    if (key >= W * W) return;
    const int da = key % W - (n - 1), db = key / W - (n - 1);
    const long long nf = (long long)n * n;
    const long long i = (da > 0 ? da : 0) + (long long)n * (db > 0 ? db : 0);
    const long long j = (da < 0 ? -da : 0) + (long long)n * (db < 0 ? -db : 0);
    Tk[4 * (size_t)key + 0] = K[i + j * ldK];
    Tk[4 * (size_t)key + 1] = K[nf + i + j * ldK];
    Tk[4 * (size_t)key + 2] = K[i + (nf + j) * ldK];
    Tk[4 * (size_t)key + 3] = K[nf + i + (nf + j) * ldK];
}

__global__ void k_max(const double* __restrict__ v, int n, double* __restrict__ out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// write_outputs's derived fields of one evaluated state (src/BP8.jl:617-630), one block:
// go[0] = max |V|, go[1] = mu sum_i w_i |V_i| (moment rate; fixed-order block reduction), and
// for every listed node k = idx[i]: fo[6 i ..] = V2, V3, tau2, tau3, q2, q3 with the Darcy
// velocity q_j = kfac (D1_j p)(k), kfac = -k / viscosity (src/PorePressure.jl:52-56).
__global__ __launch_bounds__(256) void k_outputs(Dev d, const double* __restrict__ u, const double* __restrict__ D1,
                                                 const double* __restrict__ w, double mu, double kfac,
                                                 const int* __restrict__ idx, int nidx, double* __restrict__ fo,
                                                 double* __restrict__ go) {
    __shared__ double smax[256], ssum[256];
    const int nf = d.c.nf, n2f = d.c.n2f;
    double mx = 0.0, sm = 0.0;
    for (int i = threadIdx.x; i < nf; i += 256) {
        mx = fmax(mx, d.Vm[i]);
        sm += w[i] * d.Vm[i];
    }
    smax[threadIdx.x] = mx;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && go) {
        go[0] = smax[0];
        go[1] = mu * ssum[0];
    }
    if (!fo) return;
    const double* p = u + 3 * (size_t)nf;
    for (int i = threadIdx.x; i < nidx; i += 256) {
        const int k = idx[i], a = k % n2f, b = k / n2f;
        double q2 = 0.0, q3 = 0.0;
        for (int j = 0; j < n2f; ++j) {
            q2 += D1[(size_t)a * n2f + j] * p[j + (size_t)n2f * b];
            q3 += D1[(size_t)b * n2f + j] * p[a + (size_t)n2f * j];
        }
        double* o = fo + (size_t)i * 6;
        o[0] = d.V2[k];
        o[1] = d.V3[k];
        o[2] = d.tau2[k];
        o[3] = d.tau3[k];
        o[4] = kfac * q2;
        o[5] = kfac * q3;
    }
}

__global__ void k_reset_cache(Dev d) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= d.c.nf) return;
    d.Vprev[i] = 1e-12;
    d.sig_low[i] = __builtin_huge_val();
    d.floor_hits[i] = 0;
    d.floor_mask[i] = 0;
    d.V2[i] = d.V3[i] = d.Vm[i] = d.tau2[i] = d.tau3[i] = 0.0;
    if (i == 0) d.fail[0] = 0;
}

// ---------------------------------------------------------------------------------
// Ensemble: M members (per-member Q0, a, sigma0) integrated in lockstep Tsit5 stages with
// per-member step control. The stiffness product of a stage is one GEMM D = K S over all
// members (S = the members' slip vectors), done on the FP64 matrix cores.

typedef double dvec4 __attribute__((ext_vector_type(4)));

// Split-K partial products P_z = A[:, kz] B[kz, :] (A = K column-major with leading
// dimension lda, B = the members' slip vectors column-major with leading dimension ldB;
// z = blockIdx.z covers k in [z kchunk, (z + 1) kchunk)), written to P + z ldP nm (column-major,
// leading dimension ldP). FP64 v_mfma_f64_16x16x4f64: lane l feeds A[l % 16][k] and B[k][l % 16]
// and holds C[4 j + l / 16][l % 16] in element j. Block tile kGemmBM x kGemmBN over 4 waves
// (2 x 2, each 64 x 32), k chunks of kGemmBK double-buffered in LDS (one barrier per chunk, the
// next chunk's global loads in flight during the MFMAs). Within a chunk, lane group g = l / 16
// takes k = g kGemmBK / 4 + s in MFMA step s, so a lane's B operands for the chunk are
// contiguous (ds_read_b128). A tile rows are contiguous in HBM (column-major K): BK runs of
// BM doubles per chunk.
constexpr int kGemmBM = 128, kGemmBN = 64, kGemmBK = 16;
__global__ __launch_bounds__(256) void k_ks_gemm(int nr, int nm, int nk, const double* __restrict__ A, int64_t lda,
                                                 const double* __restrict__ B, int64_t ldB, double* __restrict__ P,
                                                 int64_t ldP, int kchunk) {
    constexpr int BM = kGemmBM, BN = kGemmBN, BK = kGemmBK;
    constexpr int TM = BM / 2 / 16, TN = BN / 2 / 16;
    constexpr int LDA_S = BM + 4, LDK = BK + 2;
    constexpr int AP = BM / 2, ACS = 256 / AP, LA = BK / ACS;  // A: row pairs x columns
    constexpr int KP = BK / 2, RS = 256 / KP, LB = BN / RS;     // B: k pairs x members
    constexpr int KS = BK / 4;
    __shared__ __attribute__((aligned(16))) double As[2][BK * LDA_S];
    __shared__ __attribute__((aligned(16))) double Bs[2][BN * LDK];
    // Block -> (row block, member block, k chunk). Workgroups go round-robin over the 8 XCDs in
    // linear order; the member blocks that share one A tile (row block, k chunk) are dealt to the
    // same XCD one after another, so that XCD's L2 fetches the K tile once for all of them (with
    // the grid's natural order they landed on different XCDs and K crossed the fabric once per
    // member block). Bijective over the grid: whole groups of 8 A tiles per XCD, the remaining
    // tiles' blocks spread evenly.
    int rblk, mblk, kz;
    {
        const int nbx = (int)gridDim.x, nby = (int)gridDim.y;
        const int b = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
        const int tiles = nbx * (int)gridDim.z;
        const int xcd = b & 7, q = b >> 3;
        const int whole = (tiles >> 3) * nby;
        int tile;
        if (q < whole) {
            tile = ((q / nby) << 3) + xcd;
            mblk = q % nby;
        } else {
            // the per = (tiles % 8) nby remainder blocks: XCD x holds a + (x < r0) of them
            // (per = 8 a + r0), numbered contiguously from x a + min(x, r0)
            const int per = (tiles & 7) * nby, a = per >> 3, r0 = per & 7;
            const int r = xcd * a + min(xcd, r0) + (q - whole);
            tile = (tiles & ~7) + r / nby;
            mblk = r % nby;
        }
        rblk = tile % nbx;
        kz = tile / nbx;
    }
    const int i0 = rblk * BM, m0 = mblk * BN;
    const int kbeg = kz * kchunk, kend = min(nk, kbeg + kchunk);
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63, g = lane >> 4, r16 = lane & 15;
    const int wr = (wv >> 1) * (BM / 2), wc = (wv & 1) * (BN / 2);
    dvec4 acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
    const int ap_t = t % AP, ac_t = t / AP;
    const int kp_t = t % KP, r_t = t / KP;
    double2 ra[LA], rb[LB];
    auto gload = [&](int k0) {
        const int gi = i0 + 2 * ap_t;
#pragma unroll
        for (int j = 0; j < LA; ++j) {
            const int gk = k0 + ac_t + ACS * j;
            double2 v{0.0, 0.0};
            if (gk < kend) {
                const double* p = A + (size_t)gk * lda + gi;
                if (gi + 1 < nr) v = *reinterpret_cast<const double2*>(p);
                else if (gi < nr) v.x = p[0];
            }
            ra[j] = v;
        }
        const int gk = k0 + 2 * kp_t;
#pragma unroll
        for (int j = 0; j < LB; ++j) {
            const int gm = m0 + r_t + RS * j;
            double2 v{0.0, 0.0};
            if (gm < nm) {
                const double* p = B + (size_t)gk + (size_t)ldB * gm;
                if (gk + 1 < kend) v = *reinterpret_cast<const double2*>(p);
                else if (gk < kend) v.x = p[0];
            }
            rb[j] = v;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int j = 0; j < LA; ++j) *reinterpret_cast<double2*>(&As[buf][(ac_t + ACS * j) * LDA_S + 2 * ap_t]) = ra[j];
#pragma unroll
        for (int j = 0; j < LB; ++j) *reinterpret_cast<double2*>(&Bs[buf][(r_t + RS * j) * LDK + 2 * kp_t]) = rb[j];
    };
    gload(kbeg);
    lstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const bool more = k0 + BK < kend;
        if (more) gload(k0 + BK);
        double av[TM][KS], bv[TN][KS];
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int s = 0; s < KS; ++s) av[a][s] = As[buf][(g * KS + s) * LDA_S + wr + a * 16 + r16];
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int s = 0; s < KS; s += 2) {
                const double2 v = *reinterpret_cast<const double2*>(&Bs[buf][(wc + b * 16 + r16) * LDK + g * KS + s]);
                bv[b][s] = v.x;
                bv[b][s + 1] = v.y;
            }
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a][s], bv[b][s], acc[a][b], 0, 0, 0);
        if (more) lstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    double* Pz = P + (size_t)kz * ldP * nm;
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = i0 + wr + a * 16 + 4 * j + g, col = m0 + wc + b * 16 + r16;
                if (row < nr && col < nm) Pz[(size_t)row + (size_t)ldP * col] = acc[a][b][j];
            }
}

struct EnsDev {
    Consts c;              // shared constants (a, sigma0, src_rate, Q0 overridden per member)
    const double* Lx;
    const double* src;
    const unsigned char* active;
    const double* pa;      // per-member a
    const double* psig;    // per-member sigma0
    const double* pQ0;     // per-member Q0 (well injection rate)
    const double* psrc;    // per-member Gaussian source amplitude (Q0 / L_fwid) / (beta phi)
    int M, ldU;
    double* Vprev;         // nf x M
    double* Vm;            // nf x M (|V| of the last evaluation)
    int* fail;             // M: Newton failure flags
    double* k;             // 7 x ldU x M
    double* uprev;         // ldU x M
    double* unew;
    double* partial;       // blocks x M
    Ctl* ctl;              // M
    double* saves;         // (M x nsave_max) x ldU
    int64_t nsave_max;
    const double* D;       // splits x 2nf x M: partial K y of the current stage
    int splits;
    double* tv;            // null, or 4 x nf x M: T2, T3, V2, V3 of the last evaluation (implicit-local)
    const int* rmask;      // null, or M: the members an evaluation covers (implicit-local)
};

// A member mask in the form the FFT K S kernels skip members by (done = leave the member alone).
struct SkipFlag {
    int done;
};

// One rhs! per (node, member): thread i over nodes, member m = blockIdx.y. d.D holds the
// split-K partials of K y (splits x 2nf x M).
__global__ __launch_bounds__(256) void k_ens_stage(EnsDev d, int stage, const double* __restrict__ y, int64_t ldY,
                                                   double* __restrict__ kout, double* __restrict__ ynext,
                                                   const double* __restrict__ t_fixed) {
    const Consts& c = d.c;
    const int m = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const Ctl& cm = d.ctl[m];
    const bool run = (stage == 0 || !cm.done) && (!d.rmask || d.rmask[m]);
    double ssum = 0.0;
    if (run && i < c.nf) {
        const double dt = stage != 0 ? cm.dt : 0.0;
        const double t = stage != 0 ? cm.t + kC[stage] * dt : t_fixed[m];
        const size_t um = (size_t)m * d.ldU;
        const double* ym = y + (size_t)m * ldY;
        const double a = d.pa[m], sigma0 = d.psig[m];
        // K y: the split-K partials summed in a fixed order
        const double* Dm = d.D + (size_t)c.n2 * m;
        const size_t zs = (size_t)c.n2 * d.M;
        double dtau2 = Dm[i], dtau3 = Dm[c.nf + i];
        for (int z = 1; z < d.splits; ++z) {
            dtau2 += Dm[z * zs + i];
            dtau3 += Dm[z * zs + c.nf + i];
        }
        const int n = c.n2f, ia = i % n, ib = i / n;
        const double* P = ym + 3 * c.nf;
        const int sa = band_start(ia, n, c.lw), sb = band_start(ib, n, c.lw);
        double ap = 0.0;
        for (int q = 0; q < c.lw; ++q)
            ap += d.Lx[ia * c.lw + q] * P[sa + q + n * ib] + d.Lx[ib * c.lw + q] * P[ia + n * (sb + q)];
        ap *= c.alpha;
        const double phi = ym[2 * c.nf + i], p = P[i];
        double v2 = 0.0, v3 = 0.0, vm = 0.0;
        const double T2 = c.tau0_2 + dtau2, T3 = c.tau0_3 + dtau3;
        const size_t nm = (size_t)m * c.nf + i;
        if (d.active[i]) {
            const double sig = fmax(sigma0 - p, c.sigma_min);
            const double T = sqrt(T2 * T2 + T3 * T3);
            if (T != 0.0) {
                Consts cc = c;
                cc.a = a;
                double V;
                if (!slip_rate(T, exp(phi), sig, cc, d.Vprev[nm], V)) d.fail[m] = 1;
                v2 = V * T2 / T;
                v3 = V * T3 / T;
            }
            vm = sqrt(v2 * v2 + v3 * v3);
            d.Vprev[nm] = fmax(vm, 1e-30);
        }
        d.Vm[nm] = vm;
        if (d.tv) {
            const size_t zs = (size_t)c.nf * d.M;
            d.tv[nm] = T2;
            d.tv[zs + nm] = T3;
            d.tv[2 * zs + nm] = v2;
            d.tv[3 * zs + nm] = v3;
        }
        double f[5];
        f[0] = v2;
        f[1] = v3;
        f[2] = exp(-phi) - vm / c.D_RS;
        f[3] = ap;
        int ne = 4;
        const int idx[5] = {i, c.nf + i, 2 * c.nf + i, 3 * c.nf + i, 4 * c.nf};
        if (c.injection == 0) {
            f[3] += (t < c.t_off ? d.psrc[m] : 0.0) * d.src[i];
        } else if (i == c.well_cell) {
            const double pw = ym[4 * c.nf];
            const double transfer = c.WI * (pw - p);
            f[3] += transfer / c.S_e;
            f[4] = ((t < c.t_off ? d.pQ0[m] : 0.0) - transfer) / c.S_well;
            ne = 5;
        }
        for (int q = 0; q < ne; ++q) {
            const int e = idx[q];
            const double fe = f[q];
            kout[um + e] = fe;
            if (stage >= 2 && stage <= 6) {
                double s = 0.0;
                for (int j = 0; j < stage - 1; ++j) s += kA[stage + 1][j] * d.k[(size_t)j * d.ldU * d.M + um + e];
                s += kA[stage + 1][stage - 1] * fe;
                ynext[um + e] = d.uprev[um + e] + dt * s;
            } else if (stage == 7) {
                double s = 0.0;
                for (int j = 0; j < 6; ++j) s += kBt[j] * d.k[(size_t)j * d.ldU * d.M + um + e];
                s += kBt[6] * fe;
                const double sc = abstol_of(c, e) + fmax(fabs(d.uprev[um + e]), fabs(d.unew[um + e])) * cm.reltol;
                const double r = dt * s / sc;
                ssum += r * r;
            }
        }
    }
    if (stage == 7) {  // per-(block, member) partial of the squared scaled errors, fixed order
        __shared__ double red[256];
        red[threadIdx.x] = ssum;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) d.partial[(size_t)m * gridDim.x + blockIdx.x] = red[0];
    }
}

// Per-member controller: one thread per member, the same logic as k_bp8_control.
__global__ void k_ens_control(EnsDev d, int nblocks) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= d.M) return;
    Ctl& s = d.ctl[m];
    if (s.done) {
        s.skip_commit = 1;
        return;
    }
    double acc = 0.0;
    for (int b = 0; b < nblocks; ++b) acc += d.partial[(size_t)m * nblocks + b];
    const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2, qmax = 10.0, gamma = 0.9;
    const double EEst = sqrt(acc / d.c.nstate);
    s.EEst = EEst;
    s.nrhs += 6;
    s.iters += 1;
    s.skip_commit = 0;
    s.end_save = 0;
    s.save_count = 0;
    double q;
    if (EEst == 0.0) {
        q = 1.0 / qmax;
    } else {
        s.q11 = pow(EEst, beta1);
        q = s.q11 / pow(s.qold, beta2);
        q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
    }
    s.t_step = s.t;
    s.dt_step = s.dt;
    if (d.fail[m]) {
        s.done = 1;
        s.retcode = 2;
        s.accepted = 0;
        s.skip_commit = 1;
        return;
    }
    if (EEst <= 1.0) {
        const double tnew = s.clipped ? s.tstop : s.t + s.dt;
        s.t_new = tnew;
        s.accepted = 1;
        s.naccept += 1;
        long long cnt = 0;
        while (s.nsaved + cnt < s.nsave_max && s.t0 + (double)(s.nsaved + cnt) * s.saveat <= tnew + 1e-9 * s.saveat)
            ++cnt;
        s.save_first = (int)s.nsaved;
        s.save_count = (int)cnt;
        s.nsaved += cnt;
        s.t = tnew;
        s.qold = fmax(EEst, 1e-4);
        s.dt = s.dt / q;
        if (s.t >= s.tf) {
            const double last = s.t0 + (double)(s.nsaved - 1) * s.saveat;
            if (s.nsaved < s.nsave_max && fabs(last - s.tf) > 1e-9 * s.saveat) {
                s.end_save = 1;
                s.nsaved += 1;
            }
        }
    } else {
        s.accepted = 0;
        s.nreject += 1;
        s.dt = s.dt / fmin(1.0 / qmin, s.q11 / gamma);
    }
    if (s.t >= s.tf) {
        s.done = 1;
        s.retcode = 0;
        return;
    }
    if (s.iters >= s.maxiters) {
        s.done = 1;
        s.retcode = 1;
        return;
    }
    const double tstop = (d.c.t_off > s.t && d.c.t_off < s.tf) ? d.c.t_off : s.tf;
    s.tstop = tstop;
    s.clipped = s.dt >= tstop - s.t;
    if (s.clipped) s.dt = tstop - s.t;
}

// Per-(entry, member) commit: saves (interpolated or on-grid), FSAL, next y2.
__global__ __launch_bounds__(256) void k_ens_commit(EnsDev d, double* __restrict__ y2, int init) {
    const int m = blockIdx.y;
    const int e = blockIdx.x * 256 + threadIdx.x;
    const Ctl& s = d.ctl[m];
    if (!init && s.skip_commit) return;
    if (e >= d.c.nstate) return;
    const size_t um = (size_t)m * d.ldU, stride = (size_t)d.ldU * d.M;
    double* k1 = d.k + um;
    double up = d.uprev[um + e];
    if (!init && s.accepted) {
        const double un = d.unew[um + e];
        double* sv = d.saves + (size_t)m * d.nsave_max * d.ldU;
        for (int q = 0; q < s.save_count; ++q) {
            const long long id = s.save_first + q;
            const double ts = s.t0 + (double)id * s.saveat;
            double v;
            if (fabs(ts - s.t_new) <= 1e-9 * s.saveat) {
                v = un;
            } else {
                double b[7];
                interp_weights((ts - s.t_step) / s.dt_step, b);
                double acc = 0.0;
                for (int j = 0; j < 7; ++j) acc += b[j] * d.k[(size_t)j * stride + um + e];
                v = up + s.dt_step * acc;
            }
            sv[(size_t)id * d.ldU + e] = v;
        }
        if (s.end_save) sv[(size_t)(s.nsaved - 1) * d.ldU + e] = un;
        up = un;
        d.uprev[um + e] = un;
        k1[e] = d.k[(size_t)6 * stride + um + e];
    }
    y2[um + e] = up + s.dt * kA21 * k1[e];
}

// max_i |V| per member of the last k_ens_stage evaluation (one block per member).
__global__ void k_ens_vmax(const double* __restrict__ Vm, int nf, double* __restrict__ out, int64_t ldo) {
    const int m = blockIdx.x;
    __shared__ double red[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < nf; i += 256) v = fmax(v, Vm[(size_t)m * nf + i]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)m * ldo] = red[0];
}


// ---------------------------------------------------------------------------------
// Implicit-local integrator for BP8-PW (no counterpart in the reference, which scopes it in
// PROGRESS.md:872-894 and TODO.md:126-153): the stiff mode of the Peaceman variant is the well
// node's slip against its own self-stiffness (rate K_ii / (dg/dV)), local to one node, so the
// stage equations of an L-stable SDIRK method are solved by Newton iterations whose Jacobian is
// the per-node 3 x 3 block of (s2, s3, ln theta): nf independent 3 x 3 solves per iteration, the
// dense K coupling between nodes and the pore pressure converging as a fixed point.
//
// Method: the 5-stage, order-4 SDIRK of Hairer & Wanner (Solving ODEs II, IV.6), gamma = 1/4,
// stiffly accurate, with its embedded order-3 solution for the error estimate.
__constant__ double kSdA[5][5] = {{0.25, 0, 0, 0, 0},
                                  {0.5, 0.25, 0, 0, 0},
                                  {17.0 / 50.0, -1.0 / 25.0, 0.25, 0, 0},
                                  {371.0 / 1360.0, -137.0 / 2720.0, 15.0 / 544.0, 0.25, 0},
                                  {25.0 / 24.0, -49.0 / 48.0, 125.0 / 16.0, -85.0 / 12.0, 0.25}};
__constant__ double kSdBh[5] = {59.0 / 48.0, -17.0 / 96.0, 225.0 / 32.0, -85.0 / 12.0, 0.0};

// The node's 2 x 2 block of K (dtau_k / ds_m at the node itself), from the row-major K.
__global__ void k_imex_kdiag(Dev d, double* __restrict__ Kd) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= d.c.nf) return;
    const size_t n2 = d.c.n2, nf = d.c.nf;
    Kd[4 * i + 0] = d.KR[i * n2 + i];
    Kd[4 * i + 1] = d.KR[i * n2 + nf + i];
    Kd[4 * i + 2] = d.KR[(nf + i) * n2 + i];
    Kd[4 * i + 3] = d.KR[(nf + i) * n2 + nf + i];
}

// Stage start: base = u_n + h sum_{j < stage} A[stage][j] F_j, and the Newton start
// U = base + h gamma Fg (Fg: the last known stage derivative).
__global__ void k_imex_base(int ns, int stage, double h, const double* __restrict__ un, const double* __restrict__ F,
                            const double* __restrict__ Fg, double* __restrict__ base, double* __restrict__ U) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ns) return;
    double acc = 0.0;
    for (int j = 0; j < stage; ++j) acc += kSdA[stage][j] * F[(size_t)j * ns + e];
    const double b = un[e] + h * acc;
    base[e] = b;
    U[e] = b + h * 0.25 * Fg[e];
}

// One Newton iteration on U = base + hg f(U), f(U) in F (just evaluated, with the slip-rate cache
// of that evaluation): per node, dU = -(I - hg J)^{-1} G on (s2, s3, ln theta) with J the node's
// 3 x 3 block, dU = -G on the pressure entries; U += dU and partial[i] = the node's sum of
// (dU / scale)^2, scale = abstol + reltol |u_n|.
__global__ void k_imex_newton(Dev d, const double* __restrict__ Kd, const double* __restrict__ F,
                              const double* __restrict__ base, double hg, const double* __restrict__ un,
                              double reltol, double* __restrict__ U, double* __restrict__ partial) {
    const Consts& c = d.c;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c.nf) return;
    const int nf = c.nf;
    const int e0 = i, e1 = nf + i, e2 = 2 * nf + i, e3 = 3 * nf + i;
    const double G0 = U[e0] - base[e0] - hg * F[e0];
    const double G1 = U[e1] - base[e1] - hg * F[e1];
    const double G2 = U[e2] - base[e2] - hg * F[e2];
    const double G3 = U[e3] - base[e3] - hg * F[e3];
    const double phi = U[e2];
    // Jacobian block J[r][q], rows (dV2/dt, dV3/dt, dphi/dt), columns (s2, s3, phi)
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, -exp(-phi)}};
    const double V = d.Vm[i];
    const double T2 = d.tau2[i] + c.eta * d.V2[i], T3 = d.tau3[i] + c.eta * d.V3[i];
    const double T = sqrt(T2 * T2 + T3 * T3);
    if (d.active[i] && T > 0.0) {
        const double sig = fmax(c.sigma0 - U[e3], c.sigma_min);
        const double C = exp((c.fs + c.b * (log(c.Vs / c.Dc) + phi)) / c.a);
        const double u = V * C / (2 * c.Vs);
        const double r = sqrt(1 + u * u);
        const double D = c.eta + sig * c.a * (C / (2 * c.Vs)) / r;  // dg/dV
        const double dVdT = 1.0 / D, dVdphi = -(sig * c.b * u / r) / D;
        const double n2 = T2 / T, n3 = T3 / T;
        const double K22 = Kd[4 * i], K23 = Kd[4 * i + 1], K32 = Kd[4 * i + 2], K33 = Kd[4 * i + 3];
        const double dT2 = n2 * K22 + n3 * K32, dT3 = n2 * K23 + n3 * K33;  // dT / ds2, dT / ds3
        J[0][0] = dVdT * dT2 * n2 + V * (K22 - n2 * dT2) / T;
        J[0][1] = dVdT * dT3 * n2 + V * (K23 - n2 * dT3) / T;
        J[1][0] = dVdT * dT2 * n3 + V * (K32 - n3 * dT2) / T;
        J[1][1] = dVdT * dT3 * n3 + V * (K33 - n3 * dT3) / T;
        J[0][2] = dVdphi * n2;
        J[1][2] = dVdphi * n3;
        J[2][0] = -dVdT * dT2 / c.D_RS;
        J[2][1] = -dVdT * dT3 / c.D_RS;
        J[2][2] -= dVdphi / c.D_RS;
    }
    // (I - hg J) x = -G by Gaussian elimination with partial pivoting (3 x 3)
    double M[3][4];
    const double Gv[3] = {G0, G1, G2};
    for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) M[r][q] = (r == q ? 1.0 : 0.0) - hg * J[r][q];
        M[r][3] = -Gv[r];
    }
    for (int col = 0; col < 3; ++col) {
        int piv = col;
        for (int r = col + 1; r < 3; ++r)
            if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
        if (piv != col)
            for (int q = 0; q < 4; ++q) {
                const double t = M[col][q];
                M[col][q] = M[piv][q];
                M[piv][q] = t;
            }
        for (int r = col + 1; r < 3; ++r) {
            const double f = M[r][col] / M[col][col];
            for (int q = col; q < 4; ++q) M[r][q] -= f * M[col][q];
        }
    }
    double x[3];
    for (int r = 2; r >= 0; --r) {
        double acc = M[r][3];
        for (int q = r + 1; q < 3; ++q) acc -= M[r][q] * x[q];
        x[r] = acc / M[r][r];
    }
    const double dU[4] = {x[0], x[1], x[2], -G3};
    const int es[4] = {e0, e1, e2, e3};
    double tot = 0.0;
    for (int q = 0; q < 4; ++q) {
        const int e = es[q];
        U[e] += dU[q];
        const double sc = abstol_of(c, e) + reltol * fabs(un[e]);
        const double r = dU[q] / sc;
        tot += r * r;
    }
    if (c.injection == 1 && i == c.well_cell) {
        const int e = 4 * nf;
        const double G = U[e] - base[e] - hg * F[e];
        U[e] -= G;
        const double sc = abstol_of(c, e) + reltol * fabs(un[e]);
        tot += (G / sc) * (G / sc);
    }
    partial[i] = tot;
}

// One block: out[0] = sum of v[0..n) (fixed order), out[1] = the slip-rate failure flag, which is
// cleared for the next evaluation.
__global__ __launch_bounds__(256) void k_imex_sum(const double* __restrict__ v, int n, int* __restrict__ fail,
                                                   double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += v[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0];
        out[1] = (double)fail[0];
        fail[0] = 0;
    }
}

// The converged stage derivative F = (U - base) / hg (the stage equation solved for f).
__global__ void k_imex_stage_f(int ns, const double* __restrict__ U, const double* __restrict__ base, double hg,
                               double* __restrict__ Fs) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ns) return;
    Fs[e] = (U[e] - base[e]) / hg;
}

// Squared scaled error of the step, per entry: u_new = U (stiffly accurate) against the embedded
// u_n + h sum bh_j F_j, scale = abstol + reltol max(|u_n|, |u_new|).
__global__ void k_imex_err(Dev d, const double* __restrict__ un, const double* __restrict__ U,
                           const double* __restrict__ F, double h, double reltol, double* __restrict__ out) {
    const int ns = d.c.nstate;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ns) return;
    double acc = 0.0;
    for (int j = 0; j < 5; ++j) acc += kSdBh[j] * F[(size_t)j * ns + e];
    const double est = U[e] - (un[e] + h * acc);
    const double sc = abstol_of(d.c, e) + reltol * fmax(fabs(un[e]), fabs(U[e]));
    const double r = est / sc;
    out[e] = r * r;
}


// ---------------------------------------------------------------------------------------------
// The implicit-local integrator over an ensemble: the members advance in lockstep steps, each with
// its own step size, stages and Newton iterations of k_imex_newton's form (per (node, member)),
// every rhs! and K S over all members at once. Member m's entries sit at m ldU + e; a member
// takes part in a launch when act[m] is set.

// Stage start per member: base = u_n + h sum_{j < stage} A[stage][j] F_j, U = base + h gamma Fg.
__global__ void k_eimex_base(int ns, int ld, int stage, const double* __restrict__ h, const int* __restrict__ act,
                             const double* __restrict__ un, const double* __restrict__ F, int64_t fst,
                             const double* __restrict__ Fg, double* __restrict__ base, double* __restrict__ U) {
    const int m = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ns || !act[m]) return;
    const size_t o = (size_t)m * ld + e;
    double acc = 0.0;
    for (int j = 0; j < stage; ++j) acc += kSdA[stage][j] * F[(size_t)j * fst + o];
    const double b = un[o] + h[m] * acc;
    base[o] = b;
    U[o] = b + h[m] * 0.25 * Fg[o];
}

// One Newton iteration per (node, member) on U = base + hg f(U), f(U) in F (just evaluated, its
// tractions and slip rates in d.tv); partial[m nf + i] = the node's sum of (dU / scale)^2.
__global__ void k_eimex_newton(EnsDev d, const double* __restrict__ Kd, const double* __restrict__ F,
                               const double* __restrict__ base, const double* __restrict__ hgv,
                               const double* __restrict__ rtol, const int* __restrict__ act,
                               const double* __restrict__ un, double* __restrict__ U, double* __restrict__ partial) {
    const Consts& c = d.c;
    const int m = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
    const int nf = c.nf;
    if (i >= nf) return;
    const size_t nm = (size_t)m * nf + i;
    if (!act[m]) return;
    const double hg = hgv[m], reltol = rtol[m], a = d.pa[m];
    const size_t um = (size_t)m * d.ldU;
    const size_t e0 = um + i, e1 = um + nf + i, e2 = um + 2 * nf + i, e3 = um + 3 * nf + i;
    const double G0 = U[e0] - base[e0] - hg * F[e0];
    const double G1 = U[e1] - base[e1] - hg * F[e1];
    const double G2 = U[e2] - base[e2] - hg * F[e2];
    const double G3 = U[e3] - base[e3] - hg * F[e3];
    const double phi = U[e2];
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, -exp(-phi)}};
    const size_t zs = (size_t)nf * d.M;
    const double V = d.Vm[nm];
    const double T2 = d.tv[nm], T3 = d.tv[zs + nm];
    const double T = sqrt(T2 * T2 + T3 * T3);
    if (d.active[i] && T > 0.0) {
        const double sig = fmax(d.psig[m] - U[e3], c.sigma_min);
        const double C = exp((c.fs + c.b * (log(c.Vs / c.Dc) + phi)) / a);
        const double u = V * C / (2 * c.Vs);
        const double r = sqrt(1 + u * u);
        const double D = c.eta + sig * a * (C / (2 * c.Vs)) / r;
        const double dVdT = 1.0 / D, dVdphi = -(sig * c.b * u / r) / D;
        const double n2 = T2 / T, n3 = T3 / T;
        const double K22 = Kd[4 * i], K23 = Kd[4 * i + 1], K32 = Kd[4 * i + 2], K33 = Kd[4 * i + 3];
        const double dT2 = n2 * K22 + n3 * K32, dT3 = n2 * K23 + n3 * K33;
        J[0][0] = dVdT * dT2 * n2 + V * (K22 - n2 * dT2) / T;
        J[0][1] = dVdT * dT3 * n2 + V * (K23 - n2 * dT3) / T;
        J[1][0] = dVdT * dT2 * n3 + V * (K32 - n3 * dT2) / T;
        J[1][1] = dVdT * dT3 * n3 + V * (K33 - n3 * dT3) / T;
        J[0][2] = dVdphi * n2;
        J[1][2] = dVdphi * n3;
        J[2][0] = -dVdT * dT2 / c.D_RS;
        J[2][1] = -dVdT * dT3 / c.D_RS;
        J[2][2] -= dVdphi / c.D_RS;
    }
    double Mx[3][4];
    const double Gv[3] = {G0, G1, G2};
    for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) Mx[r][q] = (r == q ? 1.0 : 0.0) - hg * J[r][q];
        Mx[r][3] = -Gv[r];
    }
    for (int col = 0; col < 3; ++col) {
        int piv = col;
        for (int r = col + 1; r < 3; ++r)
            if (fabs(Mx[r][col]) > fabs(Mx[piv][col])) piv = r;
        if (piv != col)
            for (int q = 0; q < 4; ++q) {
                const double t = Mx[col][q];
                Mx[col][q] = Mx[piv][q];
                Mx[piv][q] = t;
            }
        for (int r = col + 1; r < 3; ++r) {
            const double f = Mx[r][col] / Mx[col][col];
            for (int q = col; q < 4; ++q) Mx[r][q] -= f * Mx[col][q];
        }
    }
    double x[3];
    for (int r = 2; r >= 0; --r) {
        double acc = Mx[r][3];
        for (int q = r + 1; q < 3; ++q) acc -= Mx[r][q] * x[q];
        x[r] = acc / Mx[r][r];
    }
    const double dU[4] = {x[0], x[1], x[2], -G3};
    const size_t es[4] = {e0, e1, e2, e3};
    const int el[4] = {i, nf + i, 2 * nf + i, 3 * nf + i};
    double tot = 0.0;
    for (int q = 0; q < 4; ++q) {
        const size_t e = es[q];
        U[e] += dU[q];
        const double sc = abstol_of(c, el[q]) + reltol * fabs(un[e]);
        const double r = dU[q] / sc;
        tot += r * r;
    }
    if (c.injection == 1 && i == c.well_cell) {
        const size_t e = um + 4 * nf;
        const double G = U[e] - base[e] - hg * F[e];
        U[e] -= G;
        const double sc = abstol_of(c, 4 * nf) + reltol * fabs(un[e]);
        tot += (G / sc) * (G / sc);
    }
    partial[nm] = tot;
}

// One block per member: out[2m] = sum of v[m n .. m n + n) (the order of k_imex_sum), out[2m+1] =
// the member's slip-rate failure flag, which is cleared.
__global__ __launch_bounds__(256) void k_eimex_sum(const double* __restrict__ v, int n, int* __restrict__ fail,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    const int m = blockIdx.x;
    const double* vm = v + (size_t)m * n;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += vm[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * m] = red[0];
        out[2 * m + 1] = (double)fail[m];
        fail[m] = 0;
    }
}

// The Newton test of one iteration for the members still iterating (nset), on the device: the
// member's sum of squared scaled updates (v, n entries per member, summed as k_eimex_sum does)
// against tol, or a failed slip-rate solve. A member that stops leaves nset (conv = 1 when it
// converged), skip follows nset for the FFT K S, cnt counts the member's evaluations.
__global__ __launch_bounds__(256) void k_eimex_decide(const double* __restrict__ v, int n, int ns, double tol,
                                                       int* __restrict__ fail, int* __restrict__ nset,
                                                       int* __restrict__ skip, int* __restrict__ conv,
                                                       long long* __restrict__ cnt) {
    __shared__ double red[256];
    const int m = blockIdx.x;
    if (!nset[m]) return;
    const double* vm = v + (size_t)m * n;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += vm[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ++cnt[m];
        const double s = red[0];
        const int fl = fail[m];
        fail[m] = 0;
        if (fl != 0 || !isfinite(s)) {
            nset[m] = 0;
        } else if (sqrt(s / ns) < tol) {
            nset[m] = 0;
            conv[m] = 1;
        }
        skip[m] = !nset[m];
    }
}

// Member masks around a stage's Newton iterations. start: every member of the step still live
// iterates (nset = live, skip = !live, conv = 0); end: a member whose stage did not converge is
// no longer live.
__global__ void k_eimex_masks(int M, int start, int* __restrict__ live, int* __restrict__ nset,
                              int* __restrict__ skip, int* __restrict__ conv) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= M) return;
    if (start) {
        nset[m] = live[m];
        skip[m] = !live[m];
        conv[m] = 0;
    } else {
        live[m] = live[m] && conv[m];
    }
}

// Converged stage derivative F = (U - base) / hg.
__global__ void k_eimex_stage_f(int ns, int ld, const double* __restrict__ U, const double* __restrict__ base,
                                const double* __restrict__ hgv, const int* __restrict__ act, double* __restrict__ Fs) {
    const int m = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ns || !act[m]) return;
    const size_t o = (size_t)m * ld + e;
    Fs[o] = (U[o] - base[o]) / hgv[m];
}

// Squared scaled error per entry (packed out[m ns + e]), as k_imex_err.
__global__ void k_eimex_err(Consts c, int ld, const double* __restrict__ un, const double* __restrict__ U,
                            const double* __restrict__ F, int64_t fst, const double* __restrict__ h,
                            const double* __restrict__ rtol, const int* __restrict__ act, double* __restrict__ out) {
    const int ns = c.nstate;
    const int m = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ns) return;
    if (!act[m]) {
        out[(size_t)m * ns + e] = 0.0;
        return;
    }
    const size_t o = (size_t)m * ld + e;
    double acc = 0.0;
    for (int j = 0; j < 5; ++j) acc += kSdBh[j] * F[(size_t)j * fst + o];
    const double est = U[o] - (un[o] + h[m] * acc);
    const double sc = abstol_of(c, e) + rtol[m] * fmax(fabs(un[o]), fabs(U[o]));
    const double r = est / sc;
    out[(size_t)m * ns + e] = r * r;
}

// Accepted members: u_n = U, f(u_n) = F_4 (stiffly accurate), and the save slot slot[m] >= 0
// (saves: member m's rows at m save_cap ld).
__global__ void k_eimex_commit(int ns, int ld, const int* __restrict__ acc, const long long* __restrict__ slot,
                               int64_t save_cap, const double* __restrict__ U, const double* __restrict__ F4,
                               double* __restrict__ un, double* __restrict__ Fn, double* __restrict__ saves) {
    const int m = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ns || !acc[m]) return;
    const size_t o = (size_t)m * ld + e;
    const double v = U[o];
    un[o] = v;
    Fn[o] = F4[o];
    if (slot[m] >= 0) saves[((size_t)m * save_cap + (size_t)slot[m]) * ld + e] = v;
}

}  // namespace bp8
}  // namespace eqd

using namespace eqd::bp8;

struct eqd_bp8 {
    eqd_bp8_params p{};
    Consts c{};
    std::vector<double> Lx_h, src_h, w_h;
    std::vector<unsigned char> active_h;
    DevMem KR, Tk, ks, Lx, src, active, V2, V3, Vm, tau2, tau3, Vprev, sig_low, floor_hits, floor_mask, fail;
    int ks_split = 0;  // source-row chunks of the direct Toeplitz K s (0: dense K or in-kernel)
    int ks_method = EQD_KS_DIRECT;  // Toeplitz K s: direct correlation or FFT convolution
    FftPlan fft;                    // the FFT convolution's plan (ks_method == EQD_KS_FFT)
    DevMem k, uprev, unew, ya, yb, partial, ctl, saves, save_t, scratch;
    DevMem Kd, imex_out;  // the implicit-local integrator: per-node 2 x 2 K blocks, reduction results
    int64_t save_cap = 0, nsaved = 0;
    hipStream_t st = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 8;
    Ctl* h_ctl = nullptr;

    ~eqd_bp8() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (st) (void)hipStreamDestroy(st);
    }
    Dev dev() const {
        Dev d;
        d.c = c;
        d.KR = KR.as<double>();
        d.Tk = Tk.p ? Tk.as<double>() : nullptr;
        d.ks = ks.p ? ks.as<double>() : nullptr;
        d.ks_split = ks_chunks();
        d.Lx = Lx.as<double>();
        d.src = src.as<double>();
        d.active = active.as<unsigned char>();
        d.V2 = V2.as<double>();
        d.V3 = V3.as<double>();
        d.Vm = Vm.as<double>();
        d.tau2 = tau2.as<double>();
        d.tau3 = tau3.as<double>();
        d.Vprev = Vprev.as<double>();
        d.sig_low = sig_low.as<double>();
        d.floor_hits = floor_hits.as<long long>();
        d.floor_mask = floor_mask.as<unsigned char>();
        d.fail = fail.as<int>();
        d.k = k.as<double>();
        d.uprev = uprev.as<double>();
        d.unew = unew.as<double>();
        d.partial = partial.as<double>();
        d.ctl = ctl.as<Ctl>();
        d.saves = saves.as<double>();
        d.save_t = save_t.as<double>();
        d.ldS = ldS();
        return d;
    }
    int node_blocks() const { return (c.nf + 3) / 4; }
    // K s chunks the stage kernel adds (0: the stage kernel forms K s itself)
    int ks_chunks() const { return ks_method == EQD_KS_FFT ? 1 : ks_split; }
    int ldS() const { return (c.nstate + 1) & ~1; }
    int state_blocks() const { return (c.nstate + 255) / 256; }
};

struct eqd_bp8_ensemble {
    eqd_bp8* base = nullptr;
    int M = 0, ldU = 0;
    std::vector<double> ha, hsig, hQ0;
    DevMem Kc, pa, psig, pQ0, psrc, Vprev, Vm, fail, k, uprev, unew, ya, yb, D, partial, ctl, saves, tfix;
    // the implicit-local integrator: tractions and slip rates of the last evaluation, per-entry
    // partials, per-member step sizes, tolerances, masks, save slots and reduction results
    DevMem itv, ipart, ih, ihg, irt, iact, iacc, islot, iout, iskip, ilive, iconv, icnt;
    double* h_iout = nullptr;
    int* h_imask = nullptr;  // 2 M: the iterating set of the last iteration, the live set of a step
    int64_t save_cap = 0, nsaved = 0;
    hipStream_t st = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 4;
    Ctl* h_ctl = nullptr;
    double gemm_flops_per_call = 0;
    int splits = 1, kchunk = 0;  // split-K of the K S GEMM
    bool use_fft = false;        // Toeplitz model: K S as the batched FFT convolution (no Kc)
    FftPlan fft;
    ~eqd_bp8_ensemble() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (h_iout) (void)hipHostFree(h_iout);
        if (h_imask) (void)hipHostFree(h_imask);
        if (st) (void)hipStreamDestroy(st);
    }
    int node_blocks() const { return (base->c.nf + 255) / 256; }
    EnsDev dev() const {
        EnsDev d;
        d.c = base->c;
        d.Lx = base->Lx.as<double>();
        d.src = base->src.as<double>();
        d.active = base->active.as<unsigned char>();
        d.pa = pa.as<double>();
        d.psig = psig.as<double>();
        d.pQ0 = pQ0.as<double>();
        d.psrc = psrc.as<double>();
        d.M = M;
        d.ldU = ldU;
        d.Vprev = Vprev.as<double>();
        d.Vm = Vm.as<double>();
        d.fail = fail.as<int>();
        d.k = k.as<double>();
        d.uprev = uprev.as<double>();
        d.unew = unew.as<double>();
        d.partial = partial.as<double>();
        d.ctl = ctl.as<Ctl>();
        d.saves = saves.as<double>();
        d.nsave_max = save_cap;
        d.D = D.as<double>();
        d.splits = splits;
        d.tv = nullptr;
        d.rmask = nullptr;
        return d;
    }
};

namespace {

void init_model(eqd_bp8& m, const eqd_bp8_params& p, int n2f, int injection) {
    if (injection != EQD_INJECTION_GAUSSIAN && injection != EQD_INJECTION_PEACEMAN)
        throw Error(EQD_ERR_ARG, "injection must be :gaussian or :peaceman, got " + std::to_string(injection));
    if (n2f < 9)
        throw Error(EQD_ERR_GRID, "Omega_f spans only " + std::to_string(n2f) +
                                      " nodes; the pore-pressure operator needs at least 9");
    if (!(p.l_f > 0) || !(p.a > 0) || !(p.D_RS > 0) || !(p.c_s > 0))
        throw Error(EQD_ERR_ARG, "l_f, a, D_RS and c_s must be positive");
    m.p = p;
    Consts& c = m.c;
    c.n2f = n2f;
    c.nf = n2f * n2f;
    c.n2 = 2 * c.nf;
    c.injection = injection;
    c.nstate = 4 * c.nf + (injection == EQD_INJECTION_PEACEMAN ? 1 : 0);
    const double h = 2.0 * p.l_f / (n2f - 1);
    std::vector<double> H;
    sbp4_neumann_1d(n2f, h, m.Lx_h, H);
    m.src_h.resize(c.nf);
    m.w_h.resize(c.nf);
    m.active_h.assign(c.nf, 1);
    double best = std::numeric_limits<double>::infinity();
    c.well_cell = 0;
    for (int b = 0; b < n2f; ++b)
        for (int a = 0; a < n2f; ++a) {
            const int i = a + n2f * b;
            const double x = -p.l_f + a * h, yv = -p.l_f + b * h;
            const double r2 = x * x + yv * yv;
            m.src_h[i] = std::exp(-r2 / (2 * p.L_gauss * p.L_gauss)) / (2 * M_PI * p.L_gauss * p.L_gauss);
            m.w_h[i] = H[a] * H[b];
            if (a == 0 || a == n2f - 1 || b == 0 || b == n2f - 1) m.active_h[i] = 0;
            if (r2 < best) {
                best = r2;
                c.well_cell = i;
            }
        }
    const double vn = std::sqrt(p.V_init * p.V_init + p.V_zero * p.V_zero);
    c.tau0_2 = p.tau_init * p.V_init / vn;
    c.tau0_3 = p.tau_init * p.V_zero / vn;
    c.eta = p.mu / (2 * p.c_s);
    c.a = p.a;
    c.b = p.b;
    c.Dc = p.D_RS;
    c.D_RS = p.D_RS;
    c.Vs = p.V_star;
    c.fs = p.f_star;
    c.sigma0 = p.sigma0;
    c.sigma_min = p.sigma_min;
    c.alpha = p.alpha;
    c.src_rate = (p.Q0 / p.L_fwid) / (p.beta * p.phi);
    c.t_off = p.t_off;
    c.Q0 = p.Q0;
    c.S_well = p.S_well;
    c.WI = 0.0;
    if (injection == EQD_INJECTION_PEACEMAN) {
        const double re = 0.268 * p.dz;
        if (!(re > p.r_well)) throw Error(EQD_ERR_ARG, "Peaceman equivalent radius 0.268*dz must exceed r_well");
        c.WI = 2 * M_PI * p.k * p.L_fwid / (p.viscosity * std::log(re / p.r_well));
    }
    c.S_e = p.dz * p.dz * p.L_fwid * p.phi * p.beta;

    check_device();
    HIPCHK(hipStreamCreateWithFlags(&m.st, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc((void**)&m.h_ctl, sizeof(Ctl), hipHostMallocDefault));
    const size_t nf = c.nf, ns = c.nstate;
    m.KR.alloc((size_t)c.n2 * c.n2 * 8);
    // banded device copy of the 1D factor: every nonzero of row r lies in [band_start(r), +lw)
    c.lw = std::min(11, n2f);
    std::vector<double> Lb((size_t)n2f * c.lw);
    for (int r = 0; r < n2f; ++r) {
        const int s0 = band_start(r, n2f, c.lw);
        for (int q = 0; q < n2f; ++q) {
            const double v = m.Lx_h[(size_t)r * n2f + q];
            if (q >= s0 && q < s0 + c.lw) Lb[(size_t)r * c.lw + (q - s0)] = v;
            else if (v != 0.0) throw Error(EQD_ERR_DEVICE, "pore-pressure factor row exceeds its band");
        }
    }
    m.Lx.alloc(Lb.size() * 8);
    m.src.alloc(nf * 8);
    m.active.alloc(nf);
    for (DevMem* v : {&m.V2, &m.V3, &m.Vm, &m.tau2, &m.tau3, &m.Vprev, &m.sig_low, &m.partial}) v->alloc(nf * 8);
    m.floor_hits.alloc(nf * 8);
    m.floor_mask.alloc(nf);
    m.fail.alloc(sizeof(int));
    m.k.alloc(7 * ns * 8);
    for (DevMem* v : {&m.uprev, &m.unew, &m.ya, &m.yb, &m.scratch}) v->alloc(ns * 8);
    m.ctl.alloc(sizeof(Ctl));
    HIPCHK(hipMemcpy(m.Lx.p, Lb.data(), Lb.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.src.p, m.src_h.data(), nf * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.active.p, m.active_h.data(), nf, hipMemcpyHostToDevice));
    std::memset(m.h_ctl, 0, sizeof(Ctl));
    HIPCHK(hipMemcpy(m.ctl.p, m.h_ctl, sizeof(Ctl), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_reset_cache, dim3((c.nf + 255) / 256), dim3(256), 0, m.st, m.dev());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m.st));
}

void set_K_device(eqd_bp8& m, const double* K_dev, int64_t ldK) {
    const int n = m.c.n2;
    hipLaunchKernelGGL(k_transpose, dim3((n + 31) / 32, (n + 31) / 32), dim3(32, 8), 0, m.st, K_dev, (long long)ldK,
                       n, m.KR.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m.st));
}

void initial_state(const eqd_bp8& m, double* u) {
    const eqd_bp8_params& p = m.p;
    std::fill(u, u + m.c.nstate, 0.0);
    // initial_state_from_strength (src/RateStateFriction.jl:127-130) at tau_init, |V_init|
    const double Vm = std::sqrt(p.V_init * p.V_init + p.V_zero * p.V_zero);
    const double f = (p.tau_init - m.c.eta * Vm) / p.sigma0;
    const double theta0 =
        (p.D_RS / p.V_star) * std::exp((p.a * std::log(2 * p.V_star * std::sinh(f / p.a) / Vm) - p.f_star) / p.b);
    for (int i = 0; i < m.c.nf; ++i) u[2 * m.c.nf + i] = std::log(theta0);
}

// The host copy of a model's separation kernel (4 (2 n2f - 1)^2 values).
std::vector<double> separation_kernel(const eqd_bp8& m) {
    std::vector<double> Tk(m.Tk.bytes / 8);
    HIPCHK(hipMemcpy(Tk.data(), m.Tk.p, m.Tk.bytes, hipMemcpyDeviceToHost));
    return Tk;
}

// Selects how a Toeplitz-kernel model forms K s: EQD_KS_DIRECT (the separation-kernel sum, in the
// stage kernel up to 48 x 48 nodes, k_toeplitz_ks beyond) or EQD_KS_FFT (the FFT convolution).
void set_ks_method(eqd_bp8& m, int method) {
    if (method != EQD_KS_DIRECT && method != EQD_KS_FFT)
        throw Error(EQD_ERR_ARG, "K s method must be EQD_KS_DIRECT or EQD_KS_FFT, got " + std::to_string(method));
    if (!m.Tk.p) throw Error(EQD_ERR_ARG, "the model has a dense K (no separation kernel)");
    if (method == EQD_KS_FFT && !m.fft.M) {
        fft_plan_build(m.fft, separation_kernel(m), m.c.n2f, 1);
        if (m.ks.bytes < (size_t)m.c.nf * 2 * 8) m.ks.alloc((size_t)m.c.nf * 2 * 8);
    }
    if (method != m.ks_method && m.exec) {
        (void)hipGraphExecDestroy(m.exec);
        m.exec = nullptr;
    }
    m.ks_method = method;
}

// K s partials of a Toeplitz-K model for the stage about to run on y (no-op for a dense K).
void toeplitz_ks(eqd_bp8& m, const Dev& d, int stage, const double* y) {
    if (m.ks_method == EQD_KS_FFT) {
        fft_ks_launch(m.fft, m.st, y, 0, m.c.nf, d.ks, 0, 2, 1, 1, d.ctl, stage != 0);
        return;
    }
    if (!m.ks_split) return;
    const int n2 = m.c.n2f, R = ks_slots(n2);
    const size_t lds = (size_t)R * (2 * n2 - 1) * 32 + (size_t)R * n2 * 16;
    hipLaunchKernelGGL(k_toeplitz_ks, dim3(n2, m.ks_split), dim3(kKsThreads), lds, m.st, d, stage, y);
}

// One stage kernel: a thread per node when K s comes precomputed, else a wave per node.
void stage_launch(eqd_bp8& m, const Dev& d, int stage, const double* y, double* kout, double* ynext, double t) {
    if (m.ks_chunks())
        hipLaunchKernelGGL(k_bp8_stage_nodes, dim3((m.c.nf + 255) / 256), dim3(256), 0, m.st, d, stage, y, kout,
                           ynext, t);
    else
        hipLaunchKernelGGL(k_bp8_stage, dim3(m.node_blocks()), dim3(256), 0, m.st, d, stage, y, kout, ynext, t);
}

// One rhs! at a device state into a device output, then the Newton status.
void rhs_device(eqd_bp8& m, const double* y_dev, double t, double* out_dev) {
    Dev d = m.dev();
    toeplitz_ks(m, d, 0, y_dev);
    stage_launch(m, d, 0, y_dev, out_dev, (double*)nullptr, t);
    HIPCHK(hipGetLastError());
}

void check_newton(eqd_bp8& m) {
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, m.fail.p, sizeof(int), hipMemcpyDeviceToHost, m.st));
    HIPCHK(hipStreamSynchronize(m.st));
    if (f) {
        const int zero = 0;
        HIPCHK(hipMemcpy(m.fail.p, &zero, sizeof(int), hipMemcpyHostToDevice));
        throw Error(EQD_ERR_NEWTON, "solve_slip_rate failed to converge");
    }
}

double rms(const std::vector<double>& v) {
    double s = 0.0;
    for (double x : v) s += x * x;
    return std::sqrt(s / (double)v.size());
}

void build_graph(eqd_bp8& m) {
    if (m.exec) return;
    Dev d = m.dev();
    double* k = m.k.as<double>();
    const size_t ns = m.c.nstate;
    double* ya = m.ya.as<double>();
    double* yb = m.yb.as<double>();
    double* un = m.unew.as<double>();
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(m.st, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < m.graph_steps; ++s) {
        // stage s reads y_s and writes y_{s+1}: y2 = ya, y3 = yb, y4 = ya, y5 = yb, y6 = ya, u_new
        const double* in[8] = {nullptr, nullptr, ya, yb, ya, yb, ya, un};
        double* out[8] = {nullptr, nullptr, yb, ya, yb, ya, un, nullptr};
        for (int stage = 2; stage <= 7; ++stage) {
            toeplitz_ks(m, d, stage, in[stage]);
            stage_launch(m, d, stage, in[stage], k + (size_t)(stage - 1) * ns, out[stage], 0.0);
        }
        hipLaunchKernelGGL(k_bp8_control, dim3(1), dim3(256), 0, m.st, d);
        hipLaunchKernelGGL(k_bp8_commit, dim3(m.state_blocks()), dim3(256), 0, m.st, d, ya, 0);
    }
    HIPCHK(hipStreamEndCapture(m.st, &g));
    HIPCHK(hipGraphInstantiate(&m.exec, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
}

// u_n = u0 on the device (uprev), f(u0) in k[0], and Hairer's initial step (OrdinaryDiffEq
// ode_determine_initdt, order + 1 = 5).
double start_state(eqd_bp8& m, double t0, double tf, double reltol) {
    const int ns = m.c.nstate;
    // u0, k1 = f(u0), Hairer's initial step (OrdinaryDiffEq ode_determine_initdt)
    std::vector<double> u0(ns), f0(ns), f1(ns), u1(ns), tmp(ns), abst(ns);
    initial_state(m, u0.data());
    for (int e = 0; e < ns; ++e) abst[e] = e < 2 * m.c.nf ? 1e-14 : (e < 3 * m.c.nf ? 1e-10 : 1e-3);
    HIPCHK(hipMemcpyAsync(m.uprev.p, u0.data(), ns * 8, hipMemcpyHostToDevice, m.st));
    rhs_device(m, m.uprev.as<double>(), t0, m.k.as<double>());
    HIPCHK(hipMemcpyAsync(f0.data(), m.k.p, ns * 8, hipMemcpyDeviceToHost, m.st));
    check_newton(m);
    std::vector<double> sk(ns);
    for (int e = 0; e < ns; ++e) sk[e] = abst[e] + std::fabs(u0[e]) * reltol;
    for (int e = 0; e < ns; ++e) tmp[e] = u0[e] / sk[e];
    const double d0 = rms(tmp);
    for (int e = 0; e < ns; ++e) tmp[e] = f0[e] / sk[e];
    const double d1 = rms(tmp);
    const double dtmax = tf - t0;
    double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
    dt0 = std::min(dt0, dtmax);
    for (int e = 0; e < ns; ++e) u1[e] = u0[e] + dt0 * f0[e];
    HIPCHK(hipMemcpyAsync(m.ya.p, u1.data(), ns * 8, hipMemcpyHostToDevice, m.st));
    rhs_device(m, m.ya.as<double>(), t0 + dt0, m.scratch.as<double>());
    HIPCHK(hipMemcpyAsync(f1.data(), m.scratch.p, ns * 8, hipMemcpyDeviceToHost, m.st));
    check_newton(m);
    for (int e = 0; e < ns; ++e) tmp[e] = (f1[e] - f0[e]) / sk[e];
    const double d2 = rms(tmp) / dt0;
    const double md = std::max(d1, d2);
    const double dt1 = md <= 1e-15 ? std::max(1e-6, dt0 * 1e-3) : std::pow(10.0, -(2 + std::log10(md)) / 5.0);
    return std::min(std::min(100 * dt0, dt1), dtmax);
}


void integrate(eqd_bp8& m, double t0, double tf, double saveat, double reltol, int64_t maxiters, int64_t nsave_max,
               double* ts, double* us, eqd_bp8_stats* stats) {
    if (!(tf > t0)) throw Error(EQD_ERR_ARG, "tspan must have tf > t0");
    if (!(saveat > 0)) throw Error(EQD_ERR_ARG, "saveat must be positive");
    if (!(reltol > 0)) throw Error(EQD_ERR_ARG, "reltol must be positive");
    if (nsave_max < 1) throw Error(EQD_ERR_ARG, "nsave_max must be at least 1");
    if (maxiters <= 0) maxiters = 100000;
    const int ns = m.c.nstate;
    if (nsave_max > m.save_cap) {
        m.saves.alloc((size_t)nsave_max * m.ldS() * 8);
        m.save_t.alloc((size_t)nsave_max * 8);
        m.save_cap = nsave_max;
        if (m.exec) {
            HIPCHK(hipGraphExecDestroy(m.exec));
            m.exec = nullptr;
        }
    }
    m.nsaved = 0;
    double dt = start_state(m, t0, tf, reltol);

    Ctl& c = *m.h_ctl;
    std::memset(&c, 0, sizeof(Ctl));
    c.t = t0;
    c.t0 = t0;
    c.tf = tf;
    c.saveat = saveat;
    c.reltol = reltol;
    c.qold = 1e-4;
    c.nsave_max = nsave_max;
    c.maxiters = maxiters;
    c.nrhs = 2;
    c.nsaved = 1;
    c.tstop = (m.c.t_off > t0 && m.c.t_off < tf) ? m.c.t_off : tf;
    c.clipped = dt >= c.tstop - t0;
    c.dt = c.clipped ? c.tstop - t0 : dt;
    HIPCHK(hipMemcpyAsync(m.ctl.p, &c, sizeof(Ctl), hipMemcpyHostToDevice, m.st));
    HIPCHK(hipMemcpyAsync(m.saves.p, m.uprev.p, ns * 8, hipMemcpyDeviceToDevice, m.st));
    HIPCHK(hipMemcpyAsync(m.save_t.p, &c.t0, 8, hipMemcpyHostToDevice, m.st));
    hipLaunchKernelGGL(k_bp8_commit, dim3(m.state_blocks()), dim3(256), 0, m.st, m.dev(), m.ya.as<double>(), 1);
    HIPCHK(hipGetLastError());
    build_graph(m);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, m.st));
    int64_t replays = 0;
    const int64_t max_replays = maxiters / m.graph_steps + 2;
    for (;;) {
        HIPCHK(hipGraphLaunch(m.exec, m.st));
        HIPCHK(hipMemcpyAsync(m.h_ctl, m.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, m.st));
        HIPCHK(hipStreamSynchronize(m.st));
        if (m.h_ctl->done) break;
        if (++replays > max_replays) throw Error(EQD_ERR_DEVICE, "BP8 step loop did not terminate");
    }
    HIPCHK(hipEventRecord(e1, m.st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const Ctl& r = *m.h_ctl;
    m.nsaved = r.nsaved;
    if (stats) {
        stats->naccept = r.naccept;
        stats->nreject = r.nreject;
        stats->nrhs = r.nrhs;
        stats->nsaved = r.nsaved;
        stats->retcode = r.retcode == 1 ? 1 : 0;
        stats->ms = ms;
    }
    if (r.retcode == 2) {
        const int zero = 0;
        HIPCHK(hipMemcpy(m.fail.p, &zero, sizeof(int), hipMemcpyHostToDevice));
        throw Error(EQD_ERR_NEWTON, "solve_slip_rate failed to converge");
    }
    if (ts) HIPCHK(hipMemcpy(ts, m.save_t.p, (size_t)r.nsaved * 8, hipMemcpyDeviceToHost));
    if (us)
        HIPCHK(hipMemcpy2D(us, (size_t)ns * 8, m.saves.p, (size_t)m.ldS() * 8, (size_t)ns * 8, (size_t)r.nsaved,
                           hipMemcpyDeviceToHost));
}

// The implicit-local integrator (see k_imex_newton): SDIRK stages solved by per-node block Newton,
// the step controlled by the embedded order-3 estimate, steps ending on every save time and on
// t_off (so saves are exact step ends, no interpolation). Newton converges when the RMS of the
// scaled update is below kNewtonTol; a stage that has not converged after kNewtonMax iterations
// (or whose slip-rate solve failed) rejects the step with dt / 4.
constexpr double kNewtonTol = 0.03;
constexpr int kNewtonMax = 10;

void integrate_imex(eqd_bp8& m, double t0, double tf, double saveat, double reltol, int64_t maxiters,
                    int64_t nsave_max, double* ts, double* us, eqd_bp8_stats* stats) {
    if (!(tf > t0)) throw Error(EQD_ERR_ARG, "tspan must have tf > t0");
    if (!(saveat > 0)) throw Error(EQD_ERR_ARG, "saveat must be positive");
    if (!(reltol > 0)) throw Error(EQD_ERR_ARG, "reltol must be positive");
    if (nsave_max < 1) throw Error(EQD_ERR_ARG, "nsave_max must be at least 1");
    if (maxiters <= 0) maxiters = 100000;
    const int ns = m.c.nstate, nf = m.c.nf;
    if (nsave_max > m.save_cap) {
        m.saves.alloc((size_t)nsave_max * m.ldS() * 8);
        m.save_t.alloc((size_t)nsave_max * 8);
        m.save_cap = nsave_max;
        if (m.exec) {
            HIPCHK(hipGraphExecDestroy(m.exec));
            m.exec = nullptr;
        }
    }
    if (!m.Kd.p) {
        m.Kd.alloc((size_t)4 * nf * 8);
        m.imex_out.alloc(2 * 8);
        hipLaunchKernelGGL(k_imex_kdiag, dim3((nf + 255) / 256), dim3(256), 0, m.st, m.dev(), m.Kd.as<double>());
        HIPCHK(hipGetLastError());
    }
    m.nsaved = 0;
    double dt = start_state(m, t0, tf, reltol);  // u_n in uprev, f(u_n) in k[0]
    Dev d = m.dev();
    double* F = m.k.as<double>();                   // stage derivatives F_0 .. F_4 in k[0..4]
    double* Fn = F + (size_t)5 * ns;                // f(u_n) (the first stage's Newton start)
    double* Ft = F + (size_t)6 * ns;                // f(U) of the current Newton iterate
    double* U = m.ya.as<double>();
    double* base = m.yb.as<double>();
    double* un = m.uprev.as<double>();
    double* ev = m.scratch.as<double>();
    HIPCHK(hipMemcpyAsync(Fn, F, (size_t)ns * 8, hipMemcpyDeviceToDevice, m.st));
    HIPCHK(hipMemcpyAsync(m.saves.p, un, (size_t)ns * 8, hipMemcpyDeviceToDevice, m.st));
    HIPCHK(hipMemcpyAsync(m.save_t.p, &t0, 8, hipMemcpyHostToDevice, m.st));
    int64_t nsaved = 1, naccept = 0, nreject = 0, nrhs = 2, iters = 0;
    int retcode = 0;
    double t = t0;
    const double sc_t = 1e-9 * saveat;
    double next_save = t0 + saveat;
    const int sb = m.state_blocks(), nb = (nf + 255) / 256;
    double red[2];
    auto reduce = [&](const double* v, int n) {
        hipLaunchKernelGGL(k_imex_sum, dim3(1), dim3(256), 0, m.st, v, n, d.fail, m.imex_out.as<double>());
        HIPCHK(hipMemcpyAsync(red, m.imex_out.p, 16, hipMemcpyDeviceToHost, m.st));
        HIPCHK(hipStreamSynchronize(m.st));
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, m.st));
    while (t < tf) {
        if (iters >= maxiters) {
            retcode = 1;
            break;
        }
        ++iters;
        // the step ends on the next save time, on t_off and on tf when it would pass them
        double stop = std::min(tf, next_save);
        if (m.c.t_off > t && m.c.t_off < stop) stop = m.c.t_off;
        const bool clipped = dt >= stop - t;
        const double h = clipped ? stop - t : dt;
        const double hg = 0.25 * h;
        static const double cs[5] = {0.25, 0.75, 11.0 / 20.0, 0.5, 1.0};
        bool ok = true;
        for (int st = 0; st < 5 && ok; ++st) {
            const double* Fg = st == 0 ? Fn : F + (size_t)(st - 1) * ns;
            hipLaunchKernelGGL(k_imex_base, dim3(sb), dim3(256), 0, m.st, ns, st, h, un, F, Fg, base, U);
            bool conv = false;
            for (int it = 0; it < kNewtonMax; ++it) {
                toeplitz_ks(m, d, 0, U);
                stage_launch(m, d, 0, U, Ft, (double*)nullptr, t + cs[st] * h);
                hipLaunchKernelGGL(k_imex_newton, dim3(nb), dim3(256), 0, m.st, d, m.Kd.as<double>(), Ft, base, hg,
                                   un, reltol, U, m.partial.as<double>());
                HIPCHK(hipGetLastError());
                reduce(m.partial.as<double>(), nf);
                ++nrhs;
                if (red[1] != 0.0 || !std::isfinite(red[0])) break;  // slip-rate solve failed
                if (std::sqrt(red[0] / ns) < kNewtonTol) {
                    conv = true;
                    break;
                }
            }
            if (!conv) {
                ok = false;
                break;
            }
            hipLaunchKernelGGL(k_imex_stage_f, dim3(sb), dim3(256), 0, m.st, ns, U, base, hg, F + (size_t)st * ns);
        }
        double err = 0.0;
        if (ok) {
            hipLaunchKernelGGL(k_imex_err, dim3(sb), dim3(256), 0, m.st, d, un, U, F, h, reltol, ev);
            reduce(ev, ns);
            err = std::sqrt(red[0] / ns);
            ok = std::isfinite(err);
        }
        if (ok && err <= 1.0) {
            t = clipped ? stop : t + h;
            ++naccept;
            HIPCHK(hipMemcpyAsync(un, U, (size_t)ns * 8, hipMemcpyDeviceToDevice, m.st));
            HIPCHK(hipMemcpyAsync(Fn, F + (size_t)4 * ns, (size_t)ns * 8, hipMemcpyDeviceToDevice, m.st));
            const bool on_save = std::fabs(t - next_save) <= sc_t;
            const bool at_end = t >= tf;
            if ((on_save || at_end) && nsaved < nsave_max) {
                HIPCHK(hipMemcpyAsync(m.saves.as<double>() + (size_t)nsaved * m.ldS(), un, (size_t)ns * 8,
                                      hipMemcpyDeviceToDevice, m.st));
                const double tsv = on_save ? next_save : tf;
                HIPCHK(hipMemcpy(m.save_t.as<double>() + nsaved, &tsv, 8, hipMemcpyHostToDevice));
                ++nsaved;
            }
            if (on_save) next_save = t0 + (double)nsaved * saveat;
            const double fac = std::min(4.0, std::max(0.2, 0.9 * std::pow(std::max(err, 1e-10), -0.25)));
            dt = clipped ? std::max(dt, h * fac) : h * fac;
        } else {
            ++nreject;
            dt = ok ? h * std::max(0.2, 0.9 * std::pow(err, -0.25)) : h * 0.25;
            if (dt < 1e-12 * std::max(1.0, std::fabs(t))) throw Error(EQD_ERR_NEWTON, "implicit stage solve failed at the minimum step");
        }
    }
    HIPCHK(hipEventRecord(e1, m.st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    m.nsaved = nsaved;
    if (stats) {
        stats->naccept = naccept;
        stats->nreject = nreject;
        stats->nrhs = nrhs;
        stats->nsaved = nsaved;
        stats->retcode = retcode;
        stats->ms = ms;
    }
    if (ts) HIPCHK(hipMemcpy(ts, m.save_t.p, (size_t)nsaved * 8, hipMemcpyDeviceToHost));
    if (us)
        HIPCHK(hipMemcpy2D(us, (size_t)ns * 8, m.saves.p, (size_t)m.ldS() * 8, (size_t)ns * 8, (size_t)nsaved,
                           hipMemcpyDeviceToHost));
}

// K S over the members' slip vectors (rows 0 .. 2nf-1 of y, member stride ldY) into E.D: the
// batched FFT convolution for a Toeplitz model (members already done skipped when skip_done),
// else the MFMA GEMM with the dense K.
void ens_gemm(eqd_bp8_ensemble& E, const double* y, int64_t ldY, bool skip_done = false) {
    const int n2 = E.base->c.n2;
    if (E.use_fft) {
        const int nf = E.base->c.nf;
        fft_ks_launch(E.fft, E.st, y, ldY, nf, E.D.as<double>(), (int64_t)n2, 1, (int64_t)nf, E.M, E.ctl.as<Ctl>(),
                      skip_done ? 1 : 0);
        return;
    }
    hipLaunchKernelGGL(k_ks_gemm, dim3((n2 + kGemmBM - 1) / kGemmBM, (E.M + kGemmBN - 1) / kGemmBN, E.splits),
                       dim3(256), 0, E.st, n2, E.M, n2, E.Kc.as<double>(), (int64_t)n2, y, ldY, E.D.as<double>(),
                       (int64_t)n2, E.kchunk);
}

void ens_stage(eqd_bp8_ensemble& E, int stage, const double* y, int64_t ldY, double* kout, double* ynext) {
    hipLaunchKernelGGL(k_ens_stage, dim3(E.node_blocks(), E.M), dim3(256), 0, E.st, E.dev(), stage, y, ldY, kout,
                       ynext, E.tfix.as<double>());
}

// rhs! of every member at y (member stride ldY) into out (member stride ldU), at times tfix.
void ens_rhs(eqd_bp8_ensemble& E, const double* y, int64_t ldY, double* out) {
    ens_gemm(E, y, ldY);
    ens_stage(E, 0, y, ldY, out, nullptr);
    HIPCHK(hipGetLastError());
}

void ens_check_newton(eqd_bp8_ensemble& E) {
    std::vector<int> f(E.M);
    HIPCHK(hipMemcpyAsync(f.data(), E.fail.p, E.M * sizeof(int), hipMemcpyDeviceToHost, E.st));
    HIPCHK(hipStreamSynchronize(E.st));
    for (int m = 0; m < E.M; ++m)
        if (f[m]) {
            HIPCHK(hipMemset(E.fail.p, 0, E.M * sizeof(int)));
            throw Error(EQD_ERR_NEWTON, "solve_slip_rate failed to converge (member " + std::to_string(m) + ")");
        }
}

void ens_initial_state(const eqd_bp8_ensemble& E, int m, double* u) {
    const eqd_bp8& B = *E.base;
    const eqd_bp8_params& p = B.p;
    std::fill(u, u + B.c.nstate, 0.0);
    const double a = E.ha[m], sig = E.hsig[m];
    const double Vm = std::sqrt(p.V_init * p.V_init + p.V_zero * p.V_zero);
    const double f = (p.tau_init - B.c.eta * Vm) / sig;
    const double theta0 = (p.D_RS / p.V_star) * std::exp((a * std::log(2 * p.V_star * std::sinh(f / a) / Vm) - p.f_star) / p.b);
    for (int i = 0; i < B.c.nf; ++i) u[2 * B.c.nf + i] = std::log(theta0);
}

// Every member's u0 in E.uprev and f(u0) in E.k (stage 0), and its first step by Hairer's rule
// (as start_state does for one trajectory).
std::vector<double> ens_start(eqd_bp8_ensemble& E, double t0, double tf, const std::vector<double>& rtol) {
    const eqd_bp8& B = *E.base;
    const int M = E.M, ns = B.c.nstate, ld = E.ldU, nf = B.c.nf;
    std::vector<double> u0((size_t)ld * M, 0.0), f0((size_t)ld * M), f1((size_t)ld * M), u1((size_t)ld * M, 0.0);
    for (int m = 0; m < M; ++m) ens_initial_state(E, m, u0.data() + (size_t)m * ld);
    std::vector<double> tf0(M, t0);
    HIPCHK(hipMemcpyAsync(E.tfix.p, tf0.data(), M * 8, hipMemcpyHostToDevice, E.st));
    HIPCHK(hipMemcpyAsync(E.uprev.p, u0.data(), u0.size() * 8, hipMemcpyHostToDevice, E.st));
    ens_rhs(E, E.uprev.as<double>(), ld, E.k.as<double>());
    HIPCHK(hipMemcpyAsync(f0.data(), E.k.p, f0.size() * 8, hipMemcpyDeviceToHost, E.st));
    ens_check_newton(E);
    // Hairer's initial step, per member
    std::vector<double> dt0(M), sk(ns), dtv(M);
    for (int m = 0; m < M; ++m) {
        const double reltol = rtol[m];
        const double* u = u0.data() + (size_t)m * ld;
        const double* f = f0.data() + (size_t)m * ld;
        double s0 = 0, s1 = 0;
        for (int e = 0; e < ns; ++e) {
            sk[e] = (e < 2 * nf ? 1e-14 : (e < 3 * nf ? 1e-10 : 1e-3)) + std::fabs(u[e]) * reltol;
            s0 += (u[e] / sk[e]) * (u[e] / sk[e]);
            s1 += (f[e] / sk[e]) * (f[e] / sk[e]);
        }
        const double d0 = std::sqrt(s0 / ns), d1 = std::sqrt(s1 / ns);
        double d = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
        dt0[m] = std::min(d, tf - t0);
        for (int e = 0; e < ns; ++e) u1[(size_t)m * ld + e] = u[e] + dt0[m] * f[e];
        tf0[m] = t0 + dt0[m];
    }
    HIPCHK(hipMemcpyAsync(E.tfix.p, tf0.data(), M * 8, hipMemcpyHostToDevice, E.st));
    HIPCHK(hipMemcpyAsync(E.ya.p, u1.data(), u1.size() * 8, hipMemcpyHostToDevice, E.st));
    ens_rhs(E, E.ya.as<double>(), ld, E.yb.as<double>());
    HIPCHK(hipMemcpyAsync(f1.data(), E.yb.p, f1.size() * 8, hipMemcpyDeviceToHost, E.st));
    ens_check_newton(E);
    std::vector<double> dts(M);
    for (int m = 0; m < M; ++m) {
        const double reltol = rtol[m];
        const double* u = u0.data() + (size_t)m * ld;
        const double* f = f0.data() + (size_t)m * ld;
        const double* g = f1.data() + (size_t)m * ld;
        double s2 = 0, s1 = 0;
        for (int e = 0; e < ns; ++e) {
            const double skv = (e < 2 * nf ? 1e-14 : (e < 3 * nf ? 1e-10 : 1e-3)) + std::fabs(u[e]) * reltol;
            s2 += ((g[e] - f[e]) / skv) * ((g[e] - f[e]) / skv);
            s1 += (f[e] / skv) * (f[e] / skv);
        }
        const double d1 = std::sqrt(s1 / ns), d2 = std::sqrt(s2 / ns) / dt0[m];
        const double md = std::max(d1, d2);
        const double dt1 = md <= 1e-15 ? std::max(1e-6, dt0[m] * 1e-3) : std::pow(10.0, -(2 + std::log10(md)) / 5.0);
        dts[m] = std::min(std::min(100 * dt0[m], dt1), tf - t0);
    }
    return dts;
}

void ens_integrate(eqd_bp8_ensemble& E, double t0, double tf, double saveat, double reltol_all,
                   const double* reltols, int64_t maxiters, int64_t nsave_max, eqd_bp8_stats* stats,
                   double* loop_ms) {
    if (!(tf > t0) || !(saveat > 0) || nsave_max < 1)
        throw Error(EQD_ERR_ARG, "need tf > t0, saveat > 0 and nsave_max >= 1");
    std::vector<double> rtol(E.M, reltol_all);
    if (reltols) std::copy(reltols, reltols + E.M, rtol.begin());
    for (double r : rtol)
        if (!(r > 0)) throw Error(EQD_ERR_ARG, "reltol must be positive");
    if (maxiters <= 0) maxiters = 100000;
    const eqd_bp8& B = *E.base;
    const int M = E.M, ns = B.c.nstate, ld = E.ldU, nf = B.c.nf;
    if (nsave_max > E.save_cap) {
        E.saves.alloc((size_t)M * nsave_max * ld * 8);
        E.save_cap = nsave_max;
        if (E.exec) {
            HIPCHK(hipGraphExecDestroy(E.exec));
            E.exec = nullptr;
        }
    }
    E.nsaved = 0;
    const std::vector<double> dts = ens_start(E, t0, tf, rtol);
    std::vector<Ctl> hc(M);
    for (int m = 0; m < M; ++m) {
        const double reltol = rtol[m];
        const double dt = dts[m];
        Ctl& c = hc[m];
        std::memset(&c, 0, sizeof(Ctl));
        c.t = t0;
        c.t0 = t0;
        c.tf = tf;
        c.saveat = saveat;
        c.reltol = reltol;
        c.qold = 1e-4;
        c.nsave_max = nsave_max;
        c.maxiters = maxiters;
        c.nrhs = 2;
        c.nsaved = 1;
        c.tstop = (B.c.t_off > t0 && B.c.t_off < tf) ? B.c.t_off : tf;
        c.clipped = dt >= c.tstop - t0;
        c.dt = c.clipped ? c.tstop - t0 : dt;
    }
    HIPCHK(hipMemcpyAsync(E.ctl.p, hc.data(), M * sizeof(Ctl), hipMemcpyHostToDevice, E.st));
    HIPCHK(hipMemcpy2DAsync(E.saves.p, (size_t)E.save_cap * ld * 8, E.uprev.p, (size_t)ld * 8, (size_t)ld * 8, M,
                            hipMemcpyDeviceToDevice, E.st));
    const dim3 cgrid((ns + 255) / 256, M);
    hipLaunchKernelGGL(k_ens_commit, cgrid, dim3(256), 0, E.st, E.dev(), E.ya.as<double>(), 1);
    HIPCHK(hipGetLastError());
    if (!E.exec) {
        double* kk = E.k.as<double>();
        const size_t kst = (size_t)ld * M;
        double* ya = E.ya.as<double>();
        double* yb = E.yb.as<double>();
        double* un = E.unew.as<double>();
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(E.st, hipStreamCaptureModeThreadLocal));
        for (int stp = 0; stp < E.graph_steps; ++stp) {
            const double* in[8] = {nullptr, nullptr, ya, yb, ya, yb, ya, un};
            double* out[8] = {nullptr, nullptr, yb, ya, yb, ya, un, nullptr};
            for (int stage = 2; stage <= 7; ++stage) {
                ens_gemm(E, in[stage], ld, true);
                ens_stage(E, stage, in[stage], ld, kk + (size_t)(stage - 1) * kst, out[stage]);
            }
            hipLaunchKernelGGL(k_ens_control, dim3((M + 63) / 64), dim3(64), 0, E.st, E.dev(), E.node_blocks());
            hipLaunchKernelGGL(k_ens_commit, cgrid, dim3(256), 0, E.st, E.dev(), ya, 0);
        }
        HIPCHK(hipStreamEndCapture(E.st, &g));
        HIPCHK(hipGraphInstantiate(&E.exec, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, E.st));
    const int64_t max_replays = maxiters / E.graph_steps + 2;
    for (int64_t rep = 0;; ++rep) {
        HIPCHK(hipGraphLaunch(E.exec, E.st));
        HIPCHK(hipMemcpyAsync(E.h_ctl, E.ctl.p, M * sizeof(Ctl), hipMemcpyDeviceToHost, E.st));
        HIPCHK(hipStreamSynchronize(E.st));
        bool all = true;
        for (int m = 0; m < M; ++m) all = all && E.h_ctl[m].done;
        if (all) break;
        if (rep > max_replays) throw Error(EQD_ERR_DEVICE, "ensemble step loop did not terminate");
    }
    HIPCHK(hipEventRecord(e1, E.st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (loop_ms) *loop_ms = ms;
    int64_t nsv = E.h_ctl[0].nsaved;
    int failed = -1;
    for (int m = 0; m < M; ++m) {
        const Ctl& r = E.h_ctl[m];
        nsv = std::min<int64_t>(nsv, r.nsaved);
        if (r.retcode == 2 && failed < 0) failed = m;
        if (stats) {
            stats[m].naccept = r.naccept;
            stats[m].nreject = r.nreject;
            stats[m].nrhs = r.nrhs;
            stats[m].nsaved = r.nsaved;
            stats[m].retcode = r.retcode;
            stats[m].ms = ms;
        }
    }
    E.nsaved = nsv;
    if (failed >= 0) {
        HIPCHK(hipMemset(E.fail.p, 0, M * sizeof(int)));
        throw Error(EQD_ERR_NEWTON, "solve_slip_rate failed to converge (member " + std::to_string(failed) + ")");
    }
}


// The implicit-local integrator over the members (the arithmetic of integrate_imex per member):
// every member keeps its own step size, Newton iterations, error control, saves and counters, and
// the host drives the lockstep loop: per Newton iteration one K S and one rhs! over all members,
// the per-(node, member) block Newton of the members still iterating, and one read of their
// update norms. A member whose stage Newton has converged waits for the others (its U no longer
// moves); a member that has reached tf, or failed, leaves the launches.
void ens_integrate_imex(eqd_bp8_ensemble& E, double t0, double tf, double saveat, double reltol_all,
                        const double* reltols, int64_t maxiters, int64_t nsave_max, eqd_bp8_stats* stats,
                        double* loop_ms) {
    if (!(tf > t0) || !(saveat > 0) || nsave_max < 1)
        throw Error(EQD_ERR_ARG, "need tf > t0, saveat > 0 and nsave_max >= 1");
    std::vector<double> rtol(E.M, reltol_all);
    if (reltols) std::copy(reltols, reltols + E.M, rtol.begin());
    for (double r : rtol)
        if (!(r > 0)) throw Error(EQD_ERR_ARG, "reltol must be positive");
    if (maxiters <= 0) maxiters = 100000;
    eqd_bp8& B = *E.base;
    const int M = E.M, ns = B.c.nstate, ld = E.ldU, nf = B.c.nf;
    if (nsave_max > E.save_cap) {
        E.saves.alloc((size_t)M * nsave_max * ld * 8);
        E.save_cap = nsave_max;
        if (E.exec) {
            HIPCHK(hipGraphExecDestroy(E.exec));
            E.exec = nullptr;
        }
    }
    if (!E.itv.p) {
        E.itv.alloc((size_t)4 * nf * M * 8);
        E.ipart.alloc((size_t)std::max(ns, nf) * M * 8);
        E.ih.alloc((size_t)M * 8);
        E.ihg.alloc((size_t)M * 8);
        E.irt.alloc((size_t)M * 8);
        E.iact.alloc((size_t)M * sizeof(int));
        E.iacc.alloc((size_t)M * sizeof(int));
        E.islot.alloc((size_t)M * sizeof(long long));
        E.iout.alloc((size_t)2 * M * 8);
        E.iskip.alloc((size_t)M * sizeof(int));
        E.ilive.alloc((size_t)M * sizeof(int));
        E.iconv.alloc((size_t)M * sizeof(int));
        E.icnt.alloc((size_t)M * sizeof(long long));
        HIPCHK(hipHostMalloc((void**)&E.h_iout, (size_t)2 * M * 8, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&E.h_imask, (size_t)2 * M * sizeof(int), hipHostMallocDefault));
    }
    if (!B.Kd.p) {
        B.Kd.alloc((size_t)4 * nf * 8);
        B.imex_out.alloc(2 * 8);
        hipLaunchKernelGGL(k_imex_kdiag, dim3((nf + 255) / 256), dim3(256), 0, E.st, B.dev(), B.Kd.as<double>());
        HIPCHK(hipGetLastError());
    }
    E.nsaved = 0;
    const std::vector<double> dts = ens_start(E, t0, tf, rtol);  // u_n in uprev, f(u_n) in k[0]
    HIPCHK(hipMemcpyAsync(E.irt.p, rtol.data(), M * 8, hipMemcpyHostToDevice, E.st));
    EnsDev d = E.dev();
    d.tv = E.itv.as<double>();
    d.rmask = E.iact.as<int>();  // an evaluation covers the members still iterating
    const size_t fst = (size_t)ld * M;
    double* F = E.k.as<double>();  // stage derivatives F_0 .. F_4 in k[0..4]
    double* Fn = F + 5 * fst;      // f(u_n)
    double* Ft = F + 6 * fst;      // f(U) of the current Newton iterate
    double* U = E.ya.as<double>();
    double* base = E.yb.as<double>();
    double* un = E.uprev.as<double>();
    HIPCHK(hipMemcpyAsync(Fn, F, fst * 8, hipMemcpyDeviceToDevice, E.st));
    HIPCHK(hipMemcpy2DAsync(E.saves.p, (size_t)E.save_cap * ld * 8, un, (size_t)ld * 8, (size_t)ns * 8, M,
                            hipMemcpyDeviceToDevice, E.st));
    std::vector<double> t(M, t0), dt(dts), next_save(M, t0 + saveat), h(M, 0.0), hg(M, 0.0), stop(M, tf), tfix(M);
    std::vector<int64_t> nsaved(M, 1), naccept(M, 0), nreject(M, 0), nrhs(M, 2), iters(M, 0);
    std::vector<int> retcode(M, 0), done(M, 0), act(M, 0), okv(M, 0), clipped(M, 0), accv(M, 0), end_save(M, 0);
    std::vector<long long> slot(M, -1);
    const double sc_t = 1e-9 * saveat;
    static const double cs[5] = {0.25, 0.75, 11.0 / 20.0, 0.5, 1.0};
    const dim3 sgrid((ns + 255) / 256, M), ngrid((nf + 255) / 256, M), mgrid((M + 255) / 256);
    double* out = E.h_iout;
    int stage_its[5] = {1, 1, 1, 1, 1};  // Newton iterations each stage took in the last step
    auto put_int = [&](DevMem& dm, const std::vector<int>& v) {
        HIPCHK(hipMemcpyAsync(dm.p, v.data(), M * sizeof(int), hipMemcpyHostToDevice, E.st));
    };
    auto reduce = [&](int n) {  // per-member sums of E.ipart (n entries each) into out[2 M]
        hipLaunchKernelGGL(k_eimex_sum, dim3(M), dim3(256), 0, E.st, E.ipart.as<double>(), n, E.fail.as<int>(),
                           E.iout.as<double>());
        HIPCHK(hipMemcpyAsync(out, E.iout.p, (size_t)2 * M * 8, hipMemcpyDeviceToHost, E.st));
        HIPCHK(hipStreamSynchronize(E.st));
    };
    HIPCHK(hipMemsetAsync(E.icnt.p, 0, M * sizeof(long long), E.st));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, E.st));
    for (;;) {
        // the members that take a step now, and its size (ending on the next save, t_off or tf)
        bool any = false;
        for (int m = 0; m < M; ++m) {
            act[m] = 0;
            if (done[m]) continue;
            if (iters[m] >= maxiters) {
                retcode[m] = 1;
                done[m] = 1;
                continue;
            }
            ++iters[m];
            double s = std::min(tf, next_save[m]);
            if (B.c.t_off > t[m] && B.c.t_off < s) s = B.c.t_off;
            clipped[m] = dt[m] >= s - t[m];
            h[m] = clipped[m] ? s - t[m] : dt[m];
            hg[m] = 0.25 * h[m];
            stop[m] = s;
            act[m] = 1;
            any = true;
        }
        if (!any) break;
        HIPCHK(hipMemcpyAsync(E.ih.p, h.data(), M * 8, hipMemcpyHostToDevice, E.st));
        HIPCHK(hipMemcpyAsync(E.ihg.p, hg.data(), M * 8, hipMemcpyHostToDevice, E.st));
        // E.ilive: the members in the step whose stages have all converged so far; E.iact: those
        // still iterating in the current stage. The Newton test runs on the device
        // (k_eimex_decide), and the host reads the iterating set back after a run of iterations,
        // as many as the stage took in the last step minus one, then one at a time: iterations
        // past a member's exit leave it untouched, so the results are those of testing after
        // every iteration.
        put_int(E.ilive, act);
        for (int st = 0; st < 5; ++st) {
            const double* Fg = st == 0 ? Fn : F + (size_t)(st - 1) * fst;
            hipLaunchKernelGGL(k_eimex_base, sgrid, dim3(256), 0, E.st, ns, ld, st, E.ih.as<double>(),
                               E.ilive.as<int>(), un, F, (int64_t)fst, Fg, base, U);
            for (int m = 0; m < M; ++m) tfix[m] = t[m] + cs[st] * h[m];
            HIPCHK(hipMemcpyAsync(E.tfix.p, tfix.data(), M * 8, hipMemcpyHostToDevice, E.st));
            hipLaunchKernelGGL(k_eimex_masks, mgrid, dim3(256), 0, E.st, M, 1, E.ilive.as<int>(), E.iact.as<int>(),
                               E.iskip.as<int>(), E.iconv.as<int>());
            int done_it = 0, run = std::max(1, stage_its[st] - 1);
            while (done_it < kNewtonMax) {
                const int nrun = std::min(run, kNewtonMax - done_it);
                for (int it = 0; it < nrun; ++it) {
                    if (E.use_fft) {  // K U of the members still iterating
                        fft_ks_launch(E.fft, E.st, U, (int64_t)ld, nf, E.D.as<double>(), (int64_t)B.c.n2, 1,
                                      (int64_t)nf, M, reinterpret_cast<const SkipFlag*>(E.iskip.as<int>()), 1);
                    } else {
                        ens_gemm(E, U, ld);
                    }
                    hipLaunchKernelGGL(k_ens_stage, dim3(E.node_blocks(), M), dim3(256), 0, E.st, d, 0, U,
                                       (int64_t)ld, Ft, (double*)nullptr, E.tfix.as<double>());
                    hipLaunchKernelGGL(k_eimex_newton, ngrid, dim3(256), 0, E.st, d, B.Kd.as<double>(), Ft, base,
                                       E.ihg.as<double>(), E.irt.as<double>(), E.iact.as<int>(), un, U,
                                       E.ipart.as<double>());
                    hipLaunchKernelGGL(k_eimex_decide, dim3(M), dim3(256), 0, E.st, E.ipart.as<double>(), nf, ns,
                                       kNewtonTol, E.fail.as<int>(), E.iact.as<int>(), E.iskip.as<int>(),
                                       E.iconv.as<int>(), E.icnt.as<long long>());
                    HIPCHK(hipGetLastError());
                }
                done_it += nrun;
                run = 1;
                HIPCHK(hipMemcpyAsync(E.h_imask, E.iact.p, M * sizeof(int), hipMemcpyDeviceToHost, E.st));
                HIPCHK(hipStreamSynchronize(E.st));
                bool left = false;
                for (int m = 0; m < M && !left; ++m) left = E.h_imask[m] != 0;
                if (!left) break;
            }
            stage_its[st] = done_it;
            // a member whose stage failed or did not converge after kNewtonMax iterations drops out
            hipLaunchKernelGGL(k_eimex_masks, mgrid, dim3(256), 0, E.st, M, 0, E.ilive.as<int>(), E.iact.as<int>(),
                               E.iskip.as<int>(), E.iconv.as<int>());
            hipLaunchKernelGGL(k_eimex_stage_f, sgrid, dim3(256), 0, E.st, ns, ld, U, base, E.ihg.as<double>(),
                               E.ilive.as<int>(), F + (size_t)st * fst);
        }
        // error estimates of the members whose stages all converged (E.ilive)
        hipLaunchKernelGGL(k_eimex_err, sgrid, dim3(256), 0, E.st, B.c, ld, un, U, F, (int64_t)fst,
                           E.ih.as<double>(), E.irt.as<double>(), E.ilive.as<int>(), E.ipart.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(E.h_imask + M, E.ilive.p, M * sizeof(int), hipMemcpyDeviceToHost, E.st));
        reduce(ns);
        for (int m = 0; m < M; ++m) okv[m] = act[m] && E.h_imask[M + m];
        for (int m = 0; m < M; ++m) {
            accv[m] = 0;
            slot[m] = -1;
            if (!act[m]) continue;
            double err = 0.0;
            bool ok = okv[m] != 0;
            if (ok) {
                err = std::sqrt(out[2 * m] / ns);
                ok = std::isfinite(err);
            }
            if (ok && err <= 1.0) {
                t[m] = clipped[m] ? stop[m] : t[m] + h[m];
                ++naccept[m];
                accv[m] = 1;
                const bool on_save = std::fabs(t[m] - next_save[m]) <= sc_t;
                const bool at_end = t[m] >= tf;
                if ((on_save || at_end) && nsaved[m] < nsave_max) {
                    slot[m] = nsaved[m];
                    if (!on_save) end_save[m] = 1;
                    ++nsaved[m];
                }
                if (on_save) next_save[m] = t0 + (double)nsaved[m] * saveat;
                const double fac = std::min(4.0, std::max(0.2, 0.9 * std::pow(std::max(err, 1e-10), -0.25)));
                dt[m] = clipped[m] ? std::max(dt[m], h[m] * fac) : h[m] * fac;
                if (at_end) done[m] = 1;
            } else {
                ++nreject[m];
                dt[m] = ok ? h[m] * std::max(0.2, 0.9 * std::pow(err, -0.25)) : h[m] * 0.25;
                if (dt[m] < 1e-12 * std::max(1.0, std::fabs(t[m]))) {
                    retcode[m] = 2;
                    done[m] = 1;
                }
            }
        }
        put_int(E.iacc, accv);
        HIPCHK(hipMemcpyAsync(E.islot.p, slot.data(), M * sizeof(long long), hipMemcpyHostToDevice, E.st));
        hipLaunchKernelGGL(k_eimex_commit, sgrid, dim3(256), 0, E.st, ns, ld, E.iacc.as<int>(),
                           E.islot.as<long long>(), (int64_t)E.save_cap, U, F + (size_t)4 * fst, un, Fn,
                           E.saves.as<double>());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(e1, E.st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (loop_ms) *loop_ms = ms;
    {
        std::vector<long long> cnt(M);  // evaluations inside the Newton iterations
        HIPCHK(hipMemcpy(cnt.data(), E.icnt.p, M * sizeof(long long), hipMemcpyDeviceToHost));
        for (int m = 0; m < M; ++m) nrhs[m] += cnt[m];
    }
    // the saves' bookkeeping where eqd_bp8_ensemble_saved reads it
    int failed = -1;
    int64_t nsv = nsaved[0];
    for (int m = 0; m < M; ++m) {
        Ctl& c = E.h_ctl[m];
        std::memset(&c, 0, sizeof(Ctl));
        c.t0 = t0;
        c.tf = tf;
        c.saveat = saveat;
        c.nsaved = nsaved[m];
        c.end_save = end_save[m];
        c.naccept = naccept[m];
        c.nreject = nreject[m];
        c.nrhs = nrhs[m];
        c.retcode = retcode[m];
        c.done = 1;
        nsv = std::min<int64_t>(nsv, nsaved[m]);
        if (retcode[m] == 2 && failed < 0) failed = m;
        if (stats) {
            stats[m].naccept = naccept[m];
            stats[m].nreject = nreject[m];
            stats[m].nrhs = nrhs[m];
            stats[m].nsaved = nsaved[m];
            stats[m].retcode = retcode[m];
            stats[m].ms = ms;
        }
    }
    E.nsaved = nsv;
    if (failed >= 0)
        throw Error(EQD_ERR_NEWTON,
                    "implicit stage solve failed at the minimum step (member " + std::to_string(failed) + ")");
}

}  // namespace

extern "C" {

void eqd_bp8_default_params(eqd_bp8_params* p) {
    if (!p) return;
    // Table 1 (src/BP8.jl:50-87)
    *p = eqd_bp8_params{32.04e9, 0.25, 2670.0, 3464.0, 0.016, 0.010, 0.5e-3, 1e-6, 0.6, 25.0e6, 14.6e6, 1e-12,
                        1e-20, 0.003, 50.0, 1.0, 1e-8, 0.1, 5e-14, 1e-3, 0.05, 1e-7, 0.05, 400.0, 10.0,
                        100 * 3600.0, 30 * 24 * 3600.0, 1.0e3};
}

eqd_status eqd_bp8_create(const eqd_bp8_params* p, const double* K, int64_t ldK, int n2f, int injection,
                          eqd_bp8** out) {
    return guard([&] {
        if (!p || !K || !out) throw Error(EQD_ERR_ARG, "null argument");
        auto m = std::make_unique<eqd_bp8>();
        init_model(*m, *p, n2f, injection);
        const int n = m->c.n2;
        if (ldK < n) throw Error(EQD_ERR_ARG, "ldK must be at least 2nf");
        std::vector<double> KR((size_t)n * n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) KR[(size_t)i * n + j] = K[(size_t)i + (size_t)j * ldK];
        HIPCHK(hipMemcpy(m->KR.p, KR.data(), KR.size() * 8, hipMemcpyHostToDevice));
        *out = m.release();
    });
}

eqd_status eqd_bp8_create_device(const eqd_bp8_params* p, const double* K_dev, int64_t ldK, int n2f, int injection,
                                 eqd_bp8** out) {
    return guard([&] {
        if (!p || !K_dev || !out) throw Error(EQD_ERR_ARG, "null argument");
        auto m = std::make_unique<eqd_bp8>();
        init_model(*m, *p, n2f, injection);
        if (ldK < m->c.n2) throw Error(EQD_ERR_ARG, "ldK must be at least 2nf");
        set_K_device(*m, K_dev, ldK);
        *out = m.release();
    });
}

eqd_status eqd_bp8_create_from_solver(eqd_solver* s, int stiffness, const eqd_bp8_params* p, int injection,
                                      eqd_bp8** out) {
    return guard([&] {
        if (!s || !p || !out) throw Error(EQD_ERR_ARG, "null argument");
        if (stiffness != EQD_STIFFNESS_EXACT && stiffness != EQD_STIFFNESS_TOEPLITZ)
            throw Error(EQD_ERR_ARG, "stiffness must be :exact or :toeplitz, got " + std::to_string(stiffness));
        int nf = 0, n2f = 0, n3f = 0;
        solver_fault_dims(s, nf, n2f, n3f);
        if (n2f != n3f) throw Error(EQD_ERR_GRID, "Omega_f must be square (n2f == n3f)");
        auto m = std::make_unique<eqd_bp8>();
        init_model(*m, *p, n2f, injection);
        DevMem K((size_t)2 * nf * 2 * nf * 8);
        const eqd_status rc = eqd_fault_stiffness_device(s, stiffness, nullptr, 0, K.as<double>(), 2 * nf);
        if (rc != EQD_OK) throw Error(rc, g_err);
        set_K_device(*m, K.as<double>(), 2 * nf);
        if (stiffness == EQD_STIFFNESS_TOEPLITZ && n2f <= 683) {
            // the time loop reads K s from the separation kernel: one block per (target x3 row,
            // source chunk), about 512 blocks in all; the kernel row and the slip row take
            // (2 n2f - 1) 32 + n2f 16 bytes of LDS (< 64 KiB up to n2f = 683)
            // (patches of up to 48 x 48 nodes read the kernel inside the stage kernel instead)
            const int W = 2 * n2f - 1;
            m->ks_split = n2f <= 48 ? 0 : std::max(1, std::min(n2f, (512 + n2f - 1) / n2f));
            if (m->ks_split) m->ks.alloc((size_t)m->ks_split * nf * 2 * 8);
            m->Tk.alloc((size_t)W * W * 4 * 8);
            hipLaunchKernelGGL(k_toeplitz_kernel, dim3((W * W + 255) / 256), dim3(256), 0, m->st, K.as<double>(),
                               (long long)(2 * nf), n2f, m->Tk.as<double>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(m->st));
            // beyond 48 x 48 nodes K s is the FFT convolution by default (up to 256 x 256)
            if (m->ks_split && FftPlan::size_for(n2f) <= kFftMaxM) set_ks_method(*m, EQD_KS_FFT);
        }
        *out = m.release();
    });
}

eqd_status eqd_bp8_set_ks_method(eqd_bp8* m, int method) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        set_ks_method(*m, method);
    });
}

eqd_status eqd_bp8_ks(eqd_bp8* m, const double* s, double* out) {
    return guard([&] {
        if (!m || !s || !out) throw Error(EQD_ERR_ARG, "null argument");
        if (!m->Tk.p) throw Error(EQD_ERR_ARG, "the model has a dense K (no separation kernel)");
        if (!m->ks_chunks()) throw Error(EQD_ERR_ARG, "the model forms K s inside its stage kernel (EQD_KS_DIRECT at "
                                                     "48 x 48 nodes or fewer)");
        const int nf = m->c.nf, chunks = m->ks_chunks();
        HIPCHK(hipMemcpyAsync(m->ya.p, s, (size_t)2 * nf * 8, hipMemcpyHostToDevice, m->st));
        toeplitz_ks(*m, m->dev(), 0, m->ya.as<double>());
        HIPCHK(hipGetLastError());
        std::vector<double> part((size_t)chunks * nf * 2);
        HIPCHK(hipMemcpyAsync(part.data(), m->ks.p, part.size() * 8, hipMemcpyDeviceToHost, m->st));
        HIPCHK(hipStreamSynchronize(m->st));
        // the chunks added in the order the stage kernel adds them
        for (int i = 0; i < nf; ++i) {
            double a2 = 0.0, a3 = 0.0;
            for (int q = 0; q < chunks; ++q) {
                a2 += part[((size_t)q * nf + i) * 2];
                a3 += part[((size_t)q * nf + i) * 2 + 1];
            }
            out[i] = a2;
            out[nf + i] = a3;
        }
    });
}

eqd_status eqd_bp8_ks_timing(eqd_bp8* m, int reps, double* ms_per_product) {
    return guard([&] {
        if (!m || reps < 1) throw Error(EQD_ERR_ARG, "null model or reps < 1");
        if (!m->Tk.p) throw Error(EQD_ERR_ARG, "the model has a dense K (no separation kernel)");
        if (!m->ks_chunks()) throw Error(EQD_ERR_ARG, "the model forms K s inside its stage kernel (EQD_KS_DIRECT "
                                                     "at 48 x 48 nodes or fewer): nothing to time");
        std::vector<double> u0(m->c.nstate);
        initial_state(*m, u0.data());
        HIPCHK(hipMemcpyAsync(m->ya.p, u0.data(), u0.size() * 8, hipMemcpyHostToDevice, m->st));
        Dev d = m->dev();
        toeplitz_ks(*m, d, 0, m->ya.as<double>());
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, m->st));
        for (int r = 0; r < reps; ++r) toeplitz_ks(*m, d, 0, m->ya.as<double>());
        HIPCHK(hipEventRecord(e1, m->st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCHK(hipGetLastError());
        if (ms_per_product) *ms_per_product = ms / reps;
    });
}

eqd_status eqd_bp8_destroy(eqd_bp8* m) {
    return guard([&] { delete m; });
}

eqd_status eqd_bp8_sizes(const eqd_bp8* m, int* nf, int* state_len) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        if (nf) *nf = m->c.nf;
        if (state_len) *state_len = m->c.nstate;
    });
}

eqd_status eqd_bp8_stiffness(const eqd_bp8* m, double* K, int64_t ldK) {
    return guard([&] {
        if (!m || !K) throw Error(EQD_ERR_ARG, "null argument");
        const int n = m->c.n2;
        if (ldK < n) throw Error(EQD_ERR_ARG, "ldK must be at least 2nf");
        std::vector<double> KR((size_t)n * n);
        HIPCHK(hipMemcpy(KR.data(), m->KR.p, KR.size() * 8, hipMemcpyDeviceToHost));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) K[(size_t)i + (size_t)j * ldK] = KR[(size_t)i * n + j];
    });
}

eqd_status eqd_bp8_operators(const eqd_bp8* m, double* Ap, double* source, double* weights) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        const int n = m->c.n2f, nf = m->c.nf;
        if (Ap) {
            std::fill(Ap, Ap + (size_t)nf * nf, 0.0);
            for (int b = 0; b < n; ++b)
                for (int a = 0; a < n; ++a) {
                    const size_t row = (size_t)(a + n * b) * nf;
                    for (int q = 0; q < n; ++q) {
                        Ap[row + q + n * b] += m->p.alpha * m->Lx_h[(size_t)a * n + q];
                        Ap[row + a + n * q] += m->p.alpha * m->Lx_h[(size_t)b * n + q];
                    }
                }
        }
        if (source) std::copy(m->src_h.begin(), m->src_h.end(), source);
        if (weights) std::copy(m->w_h.begin(), m->w_h.end(), weights);
    });
}

eqd_status eqd_bp8_initial_state(const eqd_bp8* m, double* u) {
    return guard([&] {
        if (!m || !u) throw Error(EQD_ERR_ARG, "null argument");
        initial_state(*m, u);
    });
}

eqd_status eqd_bp8_evaluate(eqd_bp8* m, const double* u, double t, double* V2, double* V3, double* Vmag, double* tau2,
                            double* tau3) {
    return guard([&] {
        if (!m || !u) throw Error(EQD_ERR_ARG, "null argument");
        const int ns = m->c.nstate, nf = m->c.nf;
        HIPCHK(hipMemcpyAsync(m->ya.p, u, ns * 8, hipMemcpyHostToDevice, m->st));
        rhs_device(*m, m->ya.as<double>(), t, m->scratch.as<double>());
        check_newton(*m);
        if (V2) HIPCHK(hipMemcpy(V2, m->V2.p, nf * 8, hipMemcpyDeviceToHost));
        if (V3) HIPCHK(hipMemcpy(V3, m->V3.p, nf * 8, hipMemcpyDeviceToHost));
        if (Vmag) HIPCHK(hipMemcpy(Vmag, m->Vm.p, nf * 8, hipMemcpyDeviceToHost));
        if (tau2) HIPCHK(hipMemcpy(tau2, m->tau2.p, nf * 8, hipMemcpyDeviceToHost));
        if (tau3) HIPCHK(hipMemcpy(tau3, m->tau3.p, nf * 8, hipMemcpyDeviceToHost));
    });
}

eqd_status eqd_bp8_rhs(eqd_bp8* m, const double* u, double t, double* du) {
    return guard([&] {
        if (!m || !u || !du) throw Error(EQD_ERR_ARG, "null argument");
        const int ns = m->c.nstate;
        HIPCHK(hipMemcpyAsync(m->ya.p, u, ns * 8, hipMemcpyHostToDevice, m->st));
        rhs_device(*m, m->ya.as<double>(), t, m->scratch.as<double>());
        check_newton(*m);
        HIPCHK(hipMemcpy(du, m->scratch.p, ns * 8, hipMemcpyDeviceToHost));
    });
}

eqd_status eqd_bp8_stress_report(const eqd_bp8* m, double* sigma_lowest, int64_t* floor_hits,
                                 unsigned char* floor_mask) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        const int nf = m->c.nf;
        std::vector<double> sl(nf);
        std::vector<long long> fh(nf);
        HIPCHK(hipMemcpy(sl.data(), m->sig_low.p, nf * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(fh.data(), m->floor_hits.p, nf * 8, hipMemcpyDeviceToHost));
        if (sigma_lowest) *sigma_lowest = *std::min_element(sl.begin(), sl.end());
        if (floor_hits) {
            long long s = 0;
            for (long long v : fh) s += v;
            *floor_hits = s;
        }
        if (floor_mask) HIPCHK(hipMemcpy(floor_mask, m->floor_mask.p, nf, hipMemcpyDeviceToHost));
    });
}

eqd_status eqd_bp8_reset_cache(eqd_bp8* m) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        hipLaunchKernelGGL(k_reset_cache, dim3((m->c.nf + 255) / 256), dim3(256), 0, m->st, m->dev());
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(m->st));
    });
}

eqd_status eqd_bp8_integrate(eqd_bp8* m, double t0, double tf, double saveat, double reltol, int64_t maxiters,
                             int64_t nsave_max, double* ts, double* us, eqd_bp8_stats* stats) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        integrate(*m, t0, tf, saveat, reltol, maxiters, nsave_max, ts, us, stats);
    });
}

eqd_status eqd_bp8_integrate_imex(eqd_bp8* m, double t0, double tf, double saveat, double reltol, int64_t maxiters,
                                  int64_t nsave_max, double* ts, double* us, eqd_bp8_stats* stats) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        integrate_imex(*m, t0, tf, saveat, reltol, maxiters, nsave_max, ts, us, stats);
    });
}

eqd_status eqd_bp8_saved_device(const eqd_bp8* m, const double** us_dev, int64_t* ld, int64_t* nsaved) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        if (us_dev) *us_dev = m->saves.as<double>();
        if (ld) *ld = m->ldS();
        if (nsaved) *nsaved = m->nsaved;
    });
}

eqd_status eqd_bp8_saved_vmax(eqd_bp8* m, double* vmax, int64_t n) {
    return guard([&] {
        if (!m || !vmax) throw Error(EQD_ERR_ARG, "null argument");
        if (n > m->nsaved) throw Error(EQD_ERR_ARG, "more values requested than states saved");
        if (n <= 0) return;
        DevMem out((size_t)n * 8);
        const size_t ld = m->ldS();
        for (int64_t j = 0; j < n; ++j) {
            rhs_device(*m, m->saves.as<double>() + (size_t)j * ld, 0.0, m->scratch.as<double>());
            hipLaunchKernelGGL(k_max, dim3(1), dim3(256), 0, m->st, m->Vm.as<double>(), m->c.nf,
                               out.as<double>() + j);
        }
        HIPCHK(hipGetLastError());
        check_newton(*m);
        HIPCHK(hipMemcpy(vmax, out.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    });
}

eqd_status eqd_bp8_saved_outputs(eqd_bp8* m, int64_t n, const int* idx, int nidx, double* out_fields,
                                 double* out_global) {
    return guard([&] {
        if (!m) throw Error(EQD_ERR_ARG, "null model");
        if (n > m->nsaved) throw Error(EQD_ERR_ARG, "more states requested than saved");
        if (nidx < 0 || (nidx > 0 && !idx)) throw Error(EQD_ERR_ARG, "bad node index list");
        for (int i = 0; i < nidx; ++i)
            if (idx[i] < 0 || idx[i] >= m->c.nf) throw Error(EQD_ERR_ARG, "node index outside Omega_f");
        if (n <= 0) return;
        const int n2f = m->c.n2f;
        std::vector<double> D1;
        sbp4_d1_1d(n2f, m->p.dz, D1);
        DevMem dD1(D1.size() * 8), dw((size_t)m->c.nf * 8), didx((size_t)std::max(nidx, 1) * sizeof(int));
        DevMem fo((size_t)std::max<int64_t>(1, n * nidx) * 6 * 8), go((size_t)n * 2 * 8);
        HIPCHK(hipMemcpyAsync(dD1.p, D1.data(), D1.size() * 8, hipMemcpyHostToDevice, m->st));
        HIPCHK(hipMemcpyAsync(dw.p, m->w_h.data(), (size_t)m->c.nf * 8, hipMemcpyHostToDevice, m->st));
        if (nidx) HIPCHK(hipMemcpyAsync(didx.p, idx, (size_t)nidx * sizeof(int), hipMemcpyHostToDevice, m->st));
        const double kfac = -m->p.k / m->p.viscosity;
        const size_t ld = m->ldS();
        for (int64_t j = 0; j < n; ++j) {  // evaluate! in save order, as write_outputs does
            const double* u = m->saves.as<double>() + (size_t)j * ld;
            rhs_device(*m, u, 0.0, m->scratch.as<double>());
            hipLaunchKernelGGL(k_outputs, dim3(1), dim3(256), 0, m->st, m->dev(), u, dD1.as<double>(), dw.as<double>(),
                               m->p.mu, kfac, didx.as<int>(), nidx,
                               nidx ? fo.as<double>() + (size_t)j * nidx * 6 : nullptr, go.as<double>() + 2 * j);
        }
        HIPCHK(hipGetLastError());
        check_newton(*m);
        if (out_fields && nidx)
            HIPCHK(hipMemcpy(out_fields, fo.p, (size_t)n * nidx * 6 * 8, hipMemcpyDeviceToHost));
        if (out_global) HIPCHK(hipMemcpy(out_global, go.p, (size_t)n * 2 * 8, hipMemcpyDeviceToHost));
        else HIPCHK(hipStreamSynchronize(m->st));
    });
}

eqd_status eqd_bp8_ensemble_create(eqd_bp8* base, int members, const double* Q0, const double* a,
                                   const double* sigma0, eqd_bp8_ensemble** out) {
    return guard([&] {
        if (!base || !out || members < 1) throw Error(EQD_ERR_ARG, "need a model and at least one member");
        auto E = std::make_unique<eqd_bp8_ensemble>();
        E->base = base;
        E->M = members;
        E->ldU = (base->c.nstate + 1) & ~1;
        E->ha.assign(members, base->p.a);
        E->hsig.assign(members, base->p.sigma0);
        E->hQ0.assign(members, base->p.Q0);
        for (int m = 0; m < members; ++m) {
            if (Q0) E->hQ0[m] = Q0[m];
            if (a) E->ha[m] = a[m];
            if (sigma0) E->hsig[m] = sigma0[m];
            if (!(E->ha[m] > 0) || !(E->hsig[m] > 0)) throw Error(EQD_ERR_ARG, "a and sigma0 must be positive");
        }
        const size_t nf = base->c.nf, M = members, ld = E->ldU;
        HIPCHK(hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking));
        HIPCHK(hipHostMalloc((void**)&E->h_ctl, M * sizeof(Ctl), hipHostMallocDefault));
        std::memset(E->h_ctl, 0, M * sizeof(Ctl));
        E->pa.alloc(M * 8);
        E->psig.alloc(M * 8);
        E->pQ0.alloc(M * 8);
        E->psrc.alloc(M * 8);
        E->tfix.alloc(M * 8);
        E->Vprev.alloc(nf * M * 8);
        E->Vm.alloc(nf * M * 8);
        E->fail.alloc(M * sizeof(int));
        E->k.alloc(7 * ld * M * 8);
        for (DevMem* v : {&E->uprev, &E->unew, &E->ya, &E->yb}) v->alloc(ld * M * 8);
        const int n2 = base->c.n2;
        if (base->Tk.p && FftPlan::size_for(base->c.n2f) <= kFftMaxM) {
            // Toeplitz model: K S as the batched FFT convolution from the separation kernel, no
            // copy of K
            E->use_fft = true;
            fft_plan_build(E->fft, separation_kernel(*base), base->c.n2f, members);
            E->splits = 1;
        } else {
            // K column-major (the model keeps it row-major for its GEMV); split-K so that the GEMM
            // has about 12 blocks per CU (at most 8 splits), with at least 64 k per split
            E->Kc.alloc((size_t)n2 * n2 * 8);
            hipLaunchKernelGGL(k_transpose, dim3((n2 + 31) / 32, (n2 + 31) / 32), dim3(32, 8), 0, E->st,
                               base->KR.as<double>(), (long long)n2, n2, E->Kc.as<double>());
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(E->st));
            const int tiles = ((n2 + kGemmBM - 1) / kGemmBM) * ((members + kGemmBN - 1) / kGemmBN);
            int want = std::max(1, std::min(8, (3072 + tiles - 1) / tiles));
            want = std::max(1, std::min(want, n2 / 64));
            E->kchunk = ((n2 + want - 1) / want + kGemmBK - 1) / kGemmBK * kGemmBK;
            E->splits = (n2 + E->kchunk - 1) / E->kchunk;
        }
        E->D.alloc((size_t)E->splits * n2 * M * 8);
        E->partial.alloc((size_t)E->node_blocks() * M * 8);
        E->ctl.alloc(M * sizeof(Ctl));
        HIPCHK(hipMemcpy(E->pa.p, E->ha.data(), M * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(E->psig.p, E->hsig.data(), M * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(E->pQ0.p, E->hQ0.data(), M * 8, hipMemcpyHostToDevice));
        std::vector<double> srate(M);
        const eqd_bp8_params& bp = base->p;
        for (size_t m = 0; m < M; ++m) srate[m] = (E->hQ0[m] / bp.L_fwid) / (bp.beta * bp.phi);
        HIPCHK(hipMemcpy(E->psrc.p, srate.data(), M * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(E->uprev.p, 0, E->uprev.bytes));
        HIPCHK(hipMemset(E->unew.p, 0, E->unew.bytes));
        HIPCHK(hipMemset(E->fail.p, 0, M * sizeof(int)));
        std::vector<double> vp(nf * M, 1e-12);
        HIPCHK(hipMemcpy(E->Vprev.p, vp.data(), vp.size() * 8, hipMemcpyHostToDevice));
        E->gemm_flops_per_call = E->use_fft ? 0.0 : 2.0 * (double)base->c.n2 * base->c.n2 * members;
        *out = E.release();
    });
}

eqd_status eqd_bp8_ensemble_destroy(eqd_bp8_ensemble* e) {
    return guard([&] { delete e; });
}

eqd_status eqd_bp8_ensemble_integrate(eqd_bp8_ensemble* e, double t0, double tf, double saveat, double reltol,
                                      const double* member_reltol, int64_t maxiters, int64_t nsave_max,
                                      eqd_bp8_stats* stats, double* loop_ms) {
    return guard([&] {
        if (!e) throw Error(EQD_ERR_ARG, "null ensemble");
        ens_integrate(*e, t0, tf, saveat, reltol, member_reltol, maxiters, nsave_max, stats, loop_ms);
    });
}

eqd_status eqd_bp8_ensemble_integrate_imex(eqd_bp8_ensemble* e, double t0, double tf, double saveat, double reltol,
                                           const double* member_reltol, int64_t maxiters, int64_t nsave_max,
                                           eqd_bp8_stats* stats, double* loop_ms) {
    return guard([&] {
        if (!e) throw Error(EQD_ERR_ARG, "null ensemble");
        ens_integrate_imex(*e, t0, tf, saveat, reltol, member_reltol, maxiters, nsave_max, stats, loop_ms);
    });
}

eqd_status eqd_bp8_ensemble_saved(const eqd_bp8_ensemble* e, int member, double* ts, double* us, int64_t* nsaved) {
    return guard([&] {
        if (!e) throw Error(EQD_ERR_ARG, "null ensemble");
        if (member < 0 || member >= e->M) throw Error(EQD_ERR_ARG, "member out of range");
        const int ns = e->base->c.nstate;
        const Ctl& c = e->h_ctl[member];
        const int64_t n = c.nsaved;
        if (nsaved) *nsaved = n;
        if (ts) {
            for (int64_t j = 0; j < n; ++j) ts[j] = c.t0 + (double)j * c.saveat;
            if (c.end_save && n > 0) ts[n - 1] = c.tf;
        }
        if (us && n > 0)
            HIPCHK(hipMemcpy2D(us, (size_t)ns * 8, e->saves.as<double>() + (size_t)member * e->save_cap * e->ldU,
                               (size_t)e->ldU * 8, (size_t)ns * 8, (size_t)n, hipMemcpyDeviceToHost));
    });
}

eqd_status eqd_bp8_ensemble_saved_vmax(eqd_bp8_ensemble* e, double* vmax, int64_t n, double* gemm_ms) {
    return guard([&] {
        if (!e || !vmax) throw Error(EQD_ERR_ARG, "null argument");
        if (n > e->nsaved) throw Error(EQD_ERR_ARG, "more values requested than states saved");
        if (n <= 0) return;
        eqd_bp8_ensemble& E = *e;
        DevMem out((size_t)E.M * n * 8);
        std::vector<double> tz(E.M, 0.0);
        HIPCHK(hipMemcpyAsync(E.tfix.p, tz.data(), E.M * 8, hipMemcpyHostToDevice, E.st));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, E.st));
        const int64_t ldY = E.save_cap * E.ldU;
        for (int64_t j = 0; j < n; ++j) {
            const double* y = E.saves.as<double>() + (size_t)j * E.ldU;
            ens_rhs(E, y, ldY, E.yb.as<double>());
            hipLaunchKernelGGL(k_ens_vmax, dim3(E.M), dim3(256), 0, E.st, E.Vm.as<double>(), E.base->c.nf,
                               out.as<double>() + j, (int64_t)n);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e1, E.st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (gemm_ms) *gemm_ms = ms;
        ens_check_newton(E);
        HIPCHK(hipMemcpy(vmax, out.p, (size_t)E.M * n * 8, hipMemcpyDeviceToHost));
    });
}

/* Time of one K S product over all members (measurement hook): average ms over `reps` GEMMs
 * on the members' current states, and the flops of one. */
eqd_status eqd_bp8_ensemble_gemm_timing(eqd_bp8_ensemble* e, int reps, double* ms_per_gemm, double* flops) {
    return guard([&] {
        if (!e || reps < 1) throw Error(EQD_ERR_ARG, "null ensemble or reps < 1");
        eqd_bp8_ensemble& E = *e;
        ens_gemm(E, E.uprev.as<double>(), E.ldU);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, E.st));
        for (int r = 0; r < reps; ++r) ens_gemm(E, E.uprev.as<double>(), E.ldU);
        HIPCHK(hipEventRecord(e1, E.st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (ms_per_gemm) *ms_per_gemm = ms / reps;
        if (flops) *flops = E.gemm_flops_per_call;
    });
}

}  // extern "C"
