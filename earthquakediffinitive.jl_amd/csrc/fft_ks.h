// K s for a Toeplitz K as a 2D circular convolution (src/FaultResponse.jl:378-390 fills K from
// the response to one slip per separation, so each of its four nf x nf blocks is a function of
// the node separation (da, db) only).
//
// With the fault nodes on an n x n grid (node i = x + n y), tau_c(x, y) = sum over (x', y') of
// k_cd(x - x', y - y') s_d(x', y'), c, d in {2, 3}. Zero-padded to M x M (M the power of two
// >= 2n - 1, so that no separation wraps), this is a circular convolution, computed as
//   z = s2 + i s3 (one complex transform for both real slip fields),
//   Z = FFT2(z); S2(k) = (Z(k) + conj Z(-k)) / 2, S3(k) = (Z(k) - conj Z(-k)) / (2i),
//   W(k) = K22(k) S2 + K23(k) S3 + i (K32(k) S2 + K33(k) S3) = G(k) Z(k) + H(k) conj Z(-k),
//   with G = (P - iQ) / 2, H = (P + iQ) / 2, P = K22 + i K32, Q = K23 + i K33,
//   tau2 + i tau3 = IFFT2(W) on the n x n corner.
// The kernel spectra G, H (scaled by 1 / M^2, an exact power of two) are computed once on the
// host from the separation kernel; the device does three passes per product:
//   k_fft_rows_fwd  one block per (fault row y, batch member): row FFT of z(., y) -> Z rows
//   k_fft_cols      one block per column pair (kx, M - kx): column FFTs of both, the spectral
//                   product (which pairs k with -k), inverse column FFTs, rows y < n kept
//   k_fft_rows_inv  one block per (fault row y, member): inverse row FFT, tau for x < n
// Transforms are Stockham autosort passes in LDS (radix 4, and one radix-2 pass when log2 M is
// odd) with a host-made twiddle table exp(-2 pi i j / M).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "host_common.h"

namespace eqd {
namespace bp8 {

// Device view of a planned K s convolution.
struct FftKs {
    int n, M, logM;
    const double2* tw;  // M: exp(-2 pi i j / M)
    const double4* GH;  // M x M, kx-major: {G.x, G.y, H.x, H.y} at (kx M + ky), scaled by 1/M^2
    double2* Z;         // batch x n x M: row transforms (rows y < n only)
    double2* W;         // batch x n x M: spectral products after the inverse column transforms
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// nfft transforms of length M = 2^logM held back to back in a (nfft x M); b is the other half of
// the Stockham ping-pong. Returns the buffer that holds the result (natural order). inv selects
// exp(+2 pi i j / M) (no 1/M scaling). A pass of radix R over sub-transforms of length Ns takes
// a[j + r M/R] (r < R) times exp(-+2 pi i r k / (R Ns)), k = j mod Ns, through an R-point DFT into
// b[(j - k) R + k + r Ns] (Govindaraju et al., SC 2008). All threads of the block take part; LDS
// is synchronised on entry and on return.
__device__ double2* stockham(double2* a, double2* b, const double2* tw, int M, int logM, int nfft, bool inv) {
    __syncthreads();
    const int Q = M >> 2;
    int Ns = 1, lNs = 0;
    for (; 4 * Ns <= M; Ns <<= 2, lNs += 2) {
        for (int q = threadIdx.x; q < nfft * Q; q += blockDim.x) {
            const int f = q >> (logM - 2), j = q & (Q - 1);
            const int k = j & (Ns - 1), st = k << (logM - 2 - lNs);
            double2 w1 = tw[st], w2 = tw[2 * st], w3 = tw[3 * st];
            if (inv) {
                w1.y = -w1.y;
                w2.y = -w2.y;
                w3.y = -w3.y;
            }
            const double2* af = a + (size_t)f * M;
            const double2 v0 = af[j], v1 = cmul(af[j + Q], w1), v2 = cmul(af[j + 2 * Q], w2),
                          v3 = cmul(af[j + 3 * Q], w3);
            const double2 t0 = make_double2(v0.x + v2.x, v0.y + v2.y), t1 = make_double2(v0.x - v2.x, v0.y - v2.y);
            const double2 t2 = make_double2(v1.x + v3.x, v1.y + v3.y), t3 = make_double2(v1.x - v3.x, v1.y - v3.y);
            // i t3 = (-t3.y, t3.x); forward y1 = t1 - i t3, y3 = t1 + i t3 (swapped for inv)
            const double2 p = make_double2(t1.x + t3.y, t1.y - t3.x), m = make_double2(t1.x - t3.y, t1.y + t3.x);
            double2* bf = b + (size_t)f * M + ((j - k) << 2) + k;
            bf[0] = make_double2(t0.x + t2.x, t0.y + t2.y);
            bf[Ns] = inv ? m : p;
            bf[2 * Ns] = make_double2(t0.x - t2.x, t0.y - t2.y);
            bf[3 * Ns] = inv ? p : m;
        }
        __syncthreads();
        double2* t = a;
        a = b;
        b = t;
    }
    if (Ns < M) {  // the radix-2 pass (Ns = M / 2)
        const int H = M >> 1;
        for (int q = threadIdx.x; q < nfft * H; q += blockDim.x) {
            const int f = q >> (logM - 1), j = q & (H - 1);
            double2 w = tw[j];
            if (inv) w.y = -w.y;
            const double2* af = a + (size_t)f * M;
            const double2 v0 = af[j], v1 = cmul(af[j + H], w);
            double2* bf = b + (size_t)f * M + j;
            bf[0] = make_double2(v0.x + v1.x, v0.y + v1.y);
            bf[Ns] = make_double2(v0.x - v1.x, v0.y - v1.y);
        }
        __syncthreads();
        double2* t = a;
        a = b;
        b = t;
    }
    return a;
}

__device__ inline void load_twiddles(double2* tw_lds, const double2* __restrict__ tw, int M) {
    for (int j = threadIdx.x; j < M; j += blockDim.x) tw_lds[j] = tw[j];
}

// Forward row transforms: block (y, b) transforms z(x, y) = s2 + i s3 of batch member b (slip s2
// at y_in + b ldY + node, s3 nf further), zero for x >= n, into Z[b][y][0 .. M).
// skip: a member whose controller (ctl[b]) is done is left alone (ctl may be null).
template <class CtlT>
__global__ __launch_bounds__(256) void k_fft_rows_fwd(FftKs f, const double* __restrict__ y_in, int64_t ldY, int nf,
                                                      const CtlT* ctl, int skip_done) {
    const int row = blockIdx.x, b = blockIdx.y;
    if (skip_done && ctl[b].done) return;
    extern __shared__ __attribute__((aligned(16))) double2 fft_lds[];
    const int M = f.M, n = f.n;
    double2* tw = fft_lds;
    double2* A = tw + M;
    double2* B = A + M;
    load_twiddles(tw, f.tw, M);
    const double* yb = y_in + (size_t)b * ldY + (size_t)n * row;
    for (int x = threadIdx.x; x < M; x += blockDim.x)
        A[x] = x < n ? make_double2(yb[x], yb[nf + x]) : make_double2(0.0, 0.0);
    const double2* R = stockham(A, B, tw, M, f.logM, 1, false);
    double2* z = f.Z + ((size_t)b * n + row) * M;
    for (int x = threadIdx.x; x < M; x += blockDim.x) z[x] = R[x];
}

// Column pass: one block per (kx, b) for kx in [0, M/2] takes columns c0 = kx and c1 = M - kx (one
// column when they coincide, kx = 0 and M/2), whose spectral products need each other's values at
// -k. A column is read and written 16 B per row, so neighbouring kx share cache lines: the 1D grid
// (padded to a multiple of 8) is mapped so that consecutive kx run on the same XCD (blocks are
// dealt to the 8 XCDs round-robin by block index) and share its L2.
template <class CtlT>
__global__ __launch_bounds__(256) void k_fft_cols(FftKs f, int batch, const CtlT* ctl, int skip_done) {
    const int P = (f.M >> 1) + 1;
    const int per = (int)(gridDim.x >> 3);
    const int L = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (L >= P * batch) return;
    const int kx = L % P, b = L / P;
    if (skip_done && ctl[b].done) return;
    extern __shared__ __attribute__((aligned(16))) double2 fft_lds[];
    const int M = f.M, n = f.n, mask = M - 1;
    const int c0 = kx, c1 = (M - kx) & mask;
    const bool two = c1 != c0;
    const int nfft = two ? 2 : 1;
    double2* tw = fft_lds;
    double2* A = tw + M;
    double2* B = A + 2 * M;
    load_twiddles(tw, f.tw, M);
    const double2* Zb = f.Z + (size_t)b * n * M;
    for (int q = threadIdx.x; q < nfft * M; q += blockDim.x) {
        const int col = q < M ? c0 : c1, ky = q & mask;
        A[q] = ky < n ? Zb[(size_t)ky * M + col] : make_double2(0.0, 0.0);
    }
    double2* R = stockham(A, B, tw, M, f.logM, nfft, false);
    double2* O = R == A ? B : A;
    // W(ky, c) = G Z(ky, c) + H conj Z(-ky, -c): column c1's values at -ky pair with column c0's
    const double2* Rc1 = two ? R + M : R;
    for (int q = threadIdx.x; q < nfft * M; q += blockDim.x) {
        const int ky = q & mask, mky = (M - ky) & mask;
        const bool first = q < M;
        const int col = first ? c0 : c1;
        const double2 zk = first ? R[ky] : Rc1[ky];
        const double2 zm = first ? Rc1[mky] : R[mky];
        const double4 gh = f.GH[(size_t)col * M + ky];
        const double2 g = make_double2(gh.x, gh.y), h = make_double2(gh.z, gh.w);
        const double2 u = cmul(g, zk), v = cmulc(h, zm);
        O[q] = make_double2(u.x + v.x, u.y + v.y);
    }
    double2* R2 = stockham(O, R, tw, M, f.logM, nfft, true);
    double2* Wb = f.W + (size_t)b * n * M;
    for (int q = threadIdx.x; q < nfft * n; q += blockDim.x) {
        const int first = q < n, y = first ? q : q - n;
        Wb[(size_t)y * M + (first ? c0 : c1)] = R2[first ? y : M + y];
    }
}

// Inverse row transforms: block (y, b) writes tau2 (real part) and tau3 (imaginary part) of nodes
// x < n of row y to out + b ldOut + node s (tau2) and + off3 (tau3).
template <class CtlT>
__global__ __launch_bounds__(256) void k_fft_rows_inv(FftKs f, double* __restrict__ out, int64_t ldOut, int s,
                                                      int64_t off3, const CtlT* ctl, int skip_done) {
    const int row = blockIdx.x, b = blockIdx.y;
    if (skip_done && ctl[b].done) return;
    extern __shared__ __attribute__((aligned(16))) double2 fft_lds[];
    const int M = f.M, n = f.n;
    double2* tw = fft_lds;
    double2* A = tw + M;
    double2* B = A + M;
    load_twiddles(tw, f.tw, M);
    const double2* w = f.W + ((size_t)b * n + row) * M;
    for (int x = threadIdx.x; x < M; x += blockDim.x) A[x] = w[x];
    const double2* R = stockham(A, B, tw, M, f.logM, 1, true);
    double* ob = out + (size_t)b * ldOut;
    for (int x = threadIdx.x; x < n; x += blockDim.x) {
        const size_t i = (size_t)(x + n * row) * s;
        ob[i] = R[x].x;
        ob[i + off3] = R[x].y;
    }
}

// Host side: the plan (twiddles, kernel spectra, the Z and W planes for `batch` members).
struct FftPlan {
    int n = 0, M = 0, logM = 0, batch = 0;
    host::DevMem tw, GH, Z, W;

    static int size_for(int n) {
        int M = 1;
        while (M < 2 * n - 1) M <<= 1;
        return M;
    }
    FftKs dev() const {
        FftKs f;
        f.n = n;
        f.M = M;
        f.logM = logM;
        f.tw = tw.as<double2>();
        f.GH = GH.as<double4>();
        f.Z = Z.as<double2>();
        f.W = W.as<double2>();
        return f;
    }
    int row_threads() const { return M / 4 < 64 ? 64 : (M / 4 > 256 ? 256 : M / 4); }
    int col_threads() const { return M / 2 < 64 ? 64 : (M / 2 > 256 ? 256 : M / 2); }
    size_t row_lds() const { return (size_t)3 * M * 16; }
    size_t col_lds() const { return (size_t)5 * M * 16; }
    // blocks of the column pass: (M/2 + 1) pairs x batch, padded to a multiple of 8
    int col_blocks(int batch) const { return ((M / 2 + 1) * batch + 7) / 8 * 8; }
};

// Largest transform: the column pass keeps 5M complex values in LDS (40 KiB at M = 512, fault
// patches of up to 256 x 256 nodes).
constexpr int kFftMaxM = 512;

// In-place radix-2 transform of a[0], a[st], ..., a[(M-1) st] (host, forward), twiddles from
// std::polar per index.
inline void host_fft(std::complex<double>* a, int M, int st) {
    for (int i = 1, j = 0; i < M; ++i) {
        int bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[(size_t)i * st], a[(size_t)j * st]);
    }
    for (int len = 2; len <= M; len <<= 1) {
        for (int i = 0; i < M; i += len) {
            for (int k = 0; k < len / 2; ++k) {
                const double ang = -2.0 * M_PI * (double)k / (double)len;
                const std::complex<double> w(std::cos(ang), std::sin(ang));
                std::complex<double>& u = a[(size_t)(i + k) * st];
                std::complex<double>& v = a[(size_t)(i + k + len / 2) * st];
                const std::complex<double> t = v * w;
                v = u - t;
                u = u + t;
            }
        }
    }
}

// Plans the product for an n x n fault grid from the separation kernel Tk (host copy, 4 values
// per key (da + n - 1) + (2n - 1)(db + n - 1): K22, K32, K23, K33) for `batch` members.
inline void fft_plan_build(FftPlan& p, const std::vector<double>& Tk, int n, int batch) {
    p.n = n;
    p.M = FftPlan::size_for(n);
    p.logM = 0;
    while ((1 << p.logM) < p.M) ++p.logM;
    p.batch = batch;
    const int M = p.M, Wd = 2 * n - 1;
    if (M > kFftMaxM) throw host::Error(EQD_ERR_GRID, "fault patch too large for the FFT K s (n2f > 256)");
    // the four real kernels wrapped onto the M x M torus: a[c][(db mod M) M + (da mod M)]
    std::vector<std::complex<double>> a((size_t)4 * M * M, 0.0);
    for (int db = -(n - 1); db <= n - 1; ++db)
        for (int da = -(n - 1); da <= n - 1; ++da) {
            const size_t key = (size_t)(da + n - 1) + (size_t)Wd * (db + n - 1);
            const size_t o = (size_t)((db + M) % M) * M + (size_t)((da + M) % M);
            for (int c = 0; c < 4; ++c) a[(size_t)c * M * M + o] = Tk[4 * key + c];
        }
    for (int c = 0; c < 4; ++c) {
        std::complex<double>* ac = a.data() + (size_t)c * M * M;
        for (int r = 0; r < M; ++r) host_fft(ac + (size_t)r * M, M, 1);  // along x (da)
        for (int col = 0; col < M; ++col) host_fft(ac + col, M, M);     // along y (db)
    }
    const double sc = 1.0 / ((double)M * (double)M);
    const std::complex<double> I(0.0, 1.0);
    std::vector<double> gh((size_t)M * M * 4);
    for (int kx = 0; kx < M; ++kx)
        for (int ky = 0; ky < M; ++ky) {
            const size_t o = (size_t)ky * M + kx;
            const std::complex<double> P = a[o] + I * a[(size_t)M * M + o];
            const std::complex<double> Q = a[(size_t)2 * M * M + o] + I * a[(size_t)3 * M * M + o];
            const std::complex<double> G = 0.5 * sc * (P - I * Q), H = 0.5 * sc * (P + I * Q);
            double* g = gh.data() + 4 * ((size_t)kx * M + ky);
            g[0] = G.real();
            g[1] = G.imag();
            g[2] = H.real();
            g[3] = H.imag();
        }
    std::vector<double> tw((size_t)2 * M);
    for (int j = 0; j < M; ++j) {
        const double ang = -2.0 * M_PI * (double)j / (double)M;
        tw[2 * j] = std::cos(ang);
        tw[2 * j + 1] = std::sin(ang);
    }
    p.tw.alloc((size_t)M * 16);
    p.GH.alloc(gh.size() * 8);
    p.Z.alloc((size_t)batch * n * M * 16);
    p.W.alloc((size_t)batch * n * M * 16);
    HIPCHK(hipMemcpy(p.tw.p, tw.data(), (size_t)M * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p.GH.p, gh.data(), gh.size() * 8, hipMemcpyHostToDevice));
}

// tau = K s for members [0, batch) of y_in (member stride ldY; s2 at node i, s3 at nf + i) into
// out + b ldOut + node s (tau2) and + off3 (tau3). Members whose ctl[b].done is set are skipped
// when skip_done.
template <class CtlT>
void fft_ks_launch(const FftPlan& p, hipStream_t st, const double* y_in, int64_t ldY, int nf, double* out,
                   int64_t ldOut, int s, int64_t off3, int batch, const CtlT* ctl, int skip_done) {
    const FftKs f = p.dev();
    hipLaunchKernelGGL(k_fft_rows_fwd<CtlT>, dim3(p.n, batch), dim3(p.row_threads()), p.row_lds(), st, f, y_in, ldY,
                       nf, ctl, skip_done);
    hipLaunchKernelGGL(k_fft_cols<CtlT>, dim3(p.col_blocks(batch)), dim3(p.col_threads()), p.col_lds(), st, f, batch,
                       ctl, skip_done);
    hipLaunchKernelGGL(k_fft_rows_inv<CtlT>, dim3(p.n, batch), dim3(p.row_threads()), p.row_lds(), st, f, out, ldOut,
                       s, off3, ctl, skip_done);
}

}  // namespace bp8
}  // namespace eqd
