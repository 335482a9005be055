// SBP order-4 diagonal-norm operators (oracle side; see sbp4.h for provenance).
#include "sbp4.h"
#include <stdexcept>

namespace sbp4 {

namespace {
// Boundary closures at the left end, in units of 1/h (D1) and 1/h^2 (D2).
const double kH[4] = {17.0 / 48, 59.0 / 48, 43.0 / 48, 49.0 / 48};
const double kD1[4][6] = {
    {-24.0 / 17, 59.0 / 34, -4.0 / 17, -3.0 / 34, 0, 0},
    {-1.0 / 2, 0, 1.0 / 2, 0, 0, 0},
    {4.0 / 43, -59.0 / 86, 0, 59.0 / 86, -4.0 / 43, 0},
    {3.0 / 98, 0, -59.0 / 98, 0, 32.0 / 49, -4.0 / 49},
};
const double kD2[4][6] = {
    {2, -5, 4, -1, 0, 0},
    {1, -2, 1, 0, 0, 0},
    {-4.0 / 43, 59.0 / 43, -110.0 / 43, 59.0 / 43, -4.0 / 43, 0},
    {-1.0 / 49, 0, 59.0 / 49, -118.0 / 49, 64.0 / 49, -4.0 / 49},
};
const double kD1i[5] = {1.0 / 12, -2.0 / 3, 0, 2.0 / 3, -1.0 / 12};
const double kD2i[5] = {-1.0 / 12, 4.0 / 3, -5.0 / 2, 4.0 / 3, -1.0 / 12};
const double kd[4] = {-11.0 / 6, 3, -3.0 / 2, 1.0 / 3};
}  // namespace

Ops1D build(int n, double h) {
    if (n < 9) throw std::runtime_error("SBP order 4 needs at least 9 points");
    Ops1D o;
    o.n = n;
    o.h = h;
    o.H.assign(n, h);
    o.D1.assign((size_t)n * n, 0.0);
    o.D2.assign((size_t)n * n, 0.0);
    o.dl.assign(n, 0.0);
    o.dr.assign(n, 0.0);
    for (int i = 0; i < 4; ++i) {
        o.H[i] = h * kH[i];
        o.H[n - 1 - i] = h * kH[i];
    }
    const double ih = 1.0 / h, ih2 = 1.0 / (h * h);
    for (int i = 4; i < n - 4; ++i)
        for (int t = 0; t < 5; ++t) {
            o.D1[(size_t)i * n + i - 2 + t] = kD1i[t] * ih;
            o.D2[(size_t)i * n + i - 2 + t] = kD2i[t] * ih2;
        }
    for (int r = 0; r < 4; ++r)
        for (int t = 0; t < 6; ++t) {
            // left closure, and the right end: D1 antisymmetric, D2 mirrored
            o.D1[(size_t)r * n + t] = kD1[r][t] * ih;
            o.D1[(size_t)(n - 1 - r) * n + (n - 1 - t)] = -kD1[r][t] * ih;
            o.D2[(size_t)r * n + t] = kD2[r][t] * ih2;
            o.D2[(size_t)(n - 1 - r) * n + (n - 1 - t)] = kD2[r][t] * ih2;
        }
    for (int t = 0; t < 4; ++t) {
        o.dl[t] = kd[t] * ih;
        o.dr[n - 1 - t] = -kd[t] * ih;
    }
    return o;
}

std::vector<double> matmul(const std::vector<double>& A, const std::vector<double>& B, int n) {
    std::vector<double> C((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            double a = A[(size_t)i * n + k];
            if (a == 0.0) continue;
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
        }
    return C;
}

}  // namespace sbp4
