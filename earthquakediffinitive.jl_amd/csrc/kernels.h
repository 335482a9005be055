// Device kernels of libeqd (gfx950). Included once, by eqd.hip.
//
// Batched layout: see eqd_internal.h. Grid convention for the node-parallel kernels:
// blockIdx.y = tile, blockIdx.x = group g in [0, G) with G = gridDim.x groups per tile; a
// block of 256 threads is (256/K) node slots x K columns and walks the tile's nodes with
// stride G*(256/K). G is a property of the grid (more groups for larger grids, so one tile
// alone still fills the device), never of the batch. Per-column partial sums are reduced
// in a fixed order (thread -> block slot 0 -> group index), so a column's arithmetic does
// not depend on which tile or lane holds it, nor on how many tiles are in flight.
#pragma once
#include "eqd_internal.h"

namespace eqd {

constexpr int kGroups = 64;      // groups per tile of small grids (the minimum)
constexpr int kMaxGroups = 32768;  // groups per tile of the largest grids
constexpr int kFoldAbove = 1024;   // grids with more groups sum their partials in two levels
#ifndef EQD_CANDIDATE_EVERY
#define EQD_CANDIDATE_EVERY 4
#endif
constexpr int kCandidateEvery = EQD_CANDIDATE_EVERY;
constexpr int kBlock = 256;

enum ColStatus : int { kActive = 0, kSolved = 1, kTired = 2, kZeroCurv = 3, kIdle = 4 };

// Per-column CG state (one slot per column of every tile in flight).
struct ColState {
    double rr;      // current (weighted) r.r
    double alpha, beta;
    double eps;     // atol + rtol*||b||
    double pnorm2;  // Krylov's running ||p||^2 estimate (zero-curvature test)
    double bnorm;
    int iters;
    int status;
    int since_rel;  // iterations since the last reliable update (mixed precision)
    int rel_flag;   // 1: a reliable update is due after this iteration
    int xpend;      // 1: x += alpha p of this iteration is still to be applied (64-column tiles)
    int pad_;
    int64_t itmax;
    int64_t col;    // global column id (-1 = idle slot)
};

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Block -> (tile, group) map. Workgroups are dispatched round-robin over the 8 XCDs in
// linear block order, so XCD x receives blocks b = x, x + 8, ... With T tiles of G
// groups (G = gridDim.x, a multiple of 64), the first 8*floor(T/8) tiles are dealt whole
// to XCDs (each XCD's L2 only holds its own tiles' vectors) and the groups of the remaining
// T % 8 tiles are split evenly over all 8 XCDs, so every XCD carries T/8 tiles' work instead
// of some carrying one tile more. Each XCD takes a contiguous run of those groups: group g
// works on pencils (and nodes) next to those of g +- 1, so an XCD's neighbour lines stay in
// its own L2.
__device__ __forceinline__ void tile_group(int& tile, int& g) {
    const int G = (int)gridDim.x;  // groups per tile
    const int b = blockIdx.x + G * blockIdx.y;
    const int xcd = b & 7, q = b >> 3;
    const int whole = (int)(gridDim.y >> 3) * G;  // per-XCD blocks of the dealt tiles
    if (q < whole) {
        tile = ((q / G) << 3) + xcd;
        g = q % G;
    } else {
        const int per = (int)(gridDim.y & 7u) * (G >> 3);  // remainder blocks per XCD
        const int r = xcd * per + (q - whole);              // 0 .. (T % 8) * G - 1
        tile = (int)(gridDim.y & ~7u) + r / G;
        g = r % G;
    }
}

// q = -H (reduced split-node operator) p for the node (m, y, z) of one column.
template <typename T, int K>
__device__ __forceinline__ void stencil_node(const OpView<T>& op, const T* __restrict__ in, int m, int y, int z,
                                             int lane, T out[3]) {
    const int NX = op.NX, NY = op.NY;
    auto ld = [&](int mm, int yy, int zz, int c) -> T {
        return in[((size_t)(mm + NX * (yy + NY * zz)) * 3 + c) * K + lane];
    };
    T a0 = 0, a1 = 0, a2 = 0;
    {  // x1 direction, same component (sigma_11 family) and the two others
        const int s = op.AXs.start[m], w = op.AXs.width[m];
        const T* c = op.AXs.c + (size_t)m * kTabW;
        for (int t = 0; t < w; ++t) a0 += c[t] * ld(s + t, y, z, 0);
        const int s2 = op.AXo.start[m], w2 = op.AXo.width[m];
        const T* c2 = op.AXo.c + (size_t)m * kTabW;
        for (int t = 0; t < w2; ++t) {
            a1 += c2[t] * ld(s2 + t, y, z, 1);
            a2 += c2[t] * ld(s2 + t, y, z, 2);
        }
    }
    {  // x2 direction
        const int s = op.To.start[y], w = op.To.width[y];
        const T* c = op.To.c + (size_t)y * kTabW;
        for (int t = 0; t < w; ++t) {
            a0 += c[t] * ld(m, s + t, z, 0);
            a2 += c[t] * ld(m, s + t, z, 2);
        }
        const int s2 = op.Ts.start[y], w2 = op.Ts.width[y];
        const T* c2 = op.Ts.c + (size_t)y * kTabW;
        for (int t = 0; t < w2; ++t) a1 += c2[t] * ld(m, s2 + t, z, 1);
    }
    {  // x3 direction
        const int s = op.To.start[z], w = op.To.width[z];
        const T* c = op.To.c + (size_t)z * kTabW;
        for (int t = 0; t < w; ++t) {
            a0 += c[t] * ld(m, y, s + t, 0);
            a1 += c[t] * ld(m, y, s + t, 1);
        }
        const int s2 = op.Ts.start[z], w2 = op.Ts.width[z];
        const T* c2 = op.Ts.c + (size_t)z * kTabW;
        for (int t = 0; t < w2; ++t) a2 += c2[t] * ld(m, y, s2 + t, 2);
    }
    {  // mixed x1-x2: (lam+mu) D1x D1y
        const int sx = op.MX.start[m], wx = op.MX.width[m];
        const T* cx = op.MX.c + (size_t)m * kTabW;
        const int sy = op.D1.start[y], wy = op.D1.width[y];
        const T* cy = op.D1.c + (size_t)y * kTabW;
        for (int tx = 0; tx < wx; ++tx) {
            T s0 = 0, s1 = 0;
            for (int ty = 0; ty < wy; ++ty) {
                s0 += cy[ty] * ld(sx + tx, sy + ty, z, 1);
                s1 += cy[ty] * ld(sx + tx, sy + ty, z, 0);
            }
            a0 += cx[tx] * s0;
            a1 += cx[tx] * s1;
        }
    }
    {  // mixed x1-x3
        const int sx = op.MX.start[m], wx = op.MX.width[m];
        const T* cx = op.MX.c + (size_t)m * kTabW;
        const int sz = op.D1.start[z], wz = op.D1.width[z];
        const T* cz = op.D1.c + (size_t)z * kTabW;
        for (int tx = 0; tx < wx; ++tx) {
            T s0 = 0, s2 = 0;
            for (int tz = 0; tz < wz; ++tz) {
                s0 += cz[tz] * ld(sx + tx, y, sz + tz, 2);
                s2 += cz[tz] * ld(sx + tx, y, sz + tz, 0);
            }
            a0 += cx[tx] * s0;
            a2 += cx[tx] * s2;
        }
    }
    {  // mixed x2-x3
        const int sy = op.D1s.start[y], wy = op.D1s.width[y];
        const T* cy = op.D1s.c + (size_t)y * kTabW;
        const int sz = op.D1.start[z], wz = op.D1.width[z];
        const T* cz = op.D1.c + (size_t)z * kTabW;
        for (int ty = 0; ty < wy; ++ty) {
            T s1 = 0, s2 = 0;
            for (int tz = 0; tz < wz; ++tz) {
                s1 += cz[tz] * ld(m, sy + ty, sz + tz, 2);
                s2 += cz[tz] * ld(m, sy + ty, sz + tz, 1);
            }
            a1 += cy[ty] * s1;
            a2 += cy[ty] * s2;
        }
    }
    const T hw = -(op.Hx[m] * op.Hy[y] * op.Hy[z]);
    out[0] = hw * a0;
    out[1] = hw * a1;
    out[2] = hw * a2;
}

// Block-level fixed-order reduction of one double per thread into partial[(tile*G+g)*K+lane].
template <int K>
__device__ __forceinline__ void block_partial(double acc, double* __restrict__ partial, int tile, int g) {
    constexpr int NPB = kBlock / K;
    __shared__ double sm[kBlock];
    sm[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0.0;
        for (int i = 0; i < NPB; ++i) s += sm[i * K + threadIdx.x];
        partial[((size_t)tile * gridDim.x + g) * K + threadIdx.x] = s;
    }
}

// CG stencil pass: q = A p and the weighted partial dot p.q. Tiles with no active
// column are skipped (their partials are written as zero).
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_apply_dot(OpView<T> op, int NN, const T* __restrict__ p,
                                                      T* __restrict__ q, const int* __restrict__ tile_on,
                                                      double* __restrict__ partial) {
    constexpr int NPB = kBlock / K;
    int tile, g;
    tile_group(tile, g);
    const int lane = threadIdx.x % K, slot = threadIdx.x / K;
    double acc = 0.0;
    if (tile_on[tile]) {
        const T* pt = p + (size_t)tile * NN * 3 * K;
        T* qt = q + (size_t)tile * NN * 3 * K;
        for (int base = g * NPB; base < NN; base += (int)gridDim.x * NPB) {
            int node = base + slot;
            if (K == 64) node = uni(node);
            if (node >= NN) continue;
            const int m = node % op.NX, yz = node / op.NX, y = yz % op.NY, z = yz / op.NY;
            T o[3];
            stencil_node<T, K>(op, pt, m, y, z, lane, o);
            const size_t e = (size_t)node * 3 * K + lane;
            const double w = (m == op.mface) ? 2.0 : 1.0;
            double d = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                qt[e + c * K] = o[c];
                d += (double)pt[e + c * K] * (double)o[c];
            }
            acc += w * d;
        }
    }
    block_partial<K>(acc, partial, tile, g);
}

// Reliable-update residual in FP64: r = b - A x, stored to r (type T) for flagged
// columns, with the weighted partial r.r (r.M^{-1}r with a Jacobi dinv) taken on the
// FP64 values.
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_residual(OpView<double> op, int NN, const double* __restrict__ x,
                                                     const double* __restrict__ b, T* __restrict__ r,
                                                     const ColState* __restrict__ cs, const int* __restrict__ tile_on,
                                                     double* __restrict__ partial, const double* __restrict__ dinv) {
    constexpr int NPB = kBlock / K;
    int tile, g;
    tile_group(tile, g);
    const int lane = threadIdx.x % K, slot = threadIdx.x / K;
    const bool mine = cs[(size_t)tile * K + lane].rel_flag != 0;
    double acc = 0.0;
    if (tile_on[tile]) {
        const size_t off = (size_t)tile * NN * 3 * K;
        for (int base = g * NPB; base < NN; base += (int)gridDim.x * NPB) {
            int node = base + slot;
            if (K == 64) node = uni(node);
            if (node >= NN) continue;
            const int m = node % op.NX, yz = node / op.NX, y = yz % op.NY, z = yz / op.NY;
            double o[3];
            stencil_node<double, K>(op, x + off, m, y, z, lane, o);
            const size_t e = off + (size_t)node * 3 * K + lane;
            const double w = (m == op.mface) ? 2.0 : 1.0;
            double d = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double rv = b[e + c * K] - o[c];
                if (mine) r[e + c * K] = (T)rv;
                if (dinv) d += rv * (dinv[(size_t)node * 3 + c] * rv);
                else d += rv * rv;
            }
            acc += w * d;
        }
    }
    block_partial<K>(acc, partial, tile, g);
}

// x += alpha p; r -= alpha q; partial r.r, or r.z with z = dinv r (Jacobi)
// (X = solution accumulator type)
template <typename T, typename X, int K>
__global__ __launch_bounds__(kBlock) void k_update(int NN, int NX, int mface, X* __restrict__ x,
                                                   T* __restrict__ r, const T* __restrict__ p,
                                                   const T* __restrict__ q, const ColState* __restrict__ cs,
                                                   const int* __restrict__ tile_on, double* __restrict__ partial,
                                                   const T* __restrict__ dinv) {
    constexpr int NPB = kBlock / K;
    int tile, g;
    tile_group(tile, g);
    const int lane = threadIdx.x % K, slot = threadIdx.x / K;
    const ColState& st = cs[(size_t)tile * K + lane];
    double acc = 0.0;
    if (tile_on[tile] && st.status == kActive) {
        const X a = (X)st.alpha;
        const T at = (T)st.alpha;
        const size_t off = (size_t)tile * NN * 3 * K;
        for (int base = g * NPB; base < NN; base += (int)gridDim.x * NPB) {
            const int node = base + slot;
            if (node >= NN) continue;
            const size_t e = off + (size_t)node * 3 * K + lane;
            const double w = (node % NX == mface) ? 2.0 : 1.0;
            double d = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const size_t i = e + c * K;
                x[i] += a * (X)p[i];
                T rv = r[i] - at * q[i];
                r[i] = rv;
                if (dinv) {
                    const T zv = dinv[(size_t)node * 3 + c] * rv;
                    d += (double)rv * (double)zv;
                } else {
                    d += (double)rv * (double)rv;
                }
            }
            acc += w * d;
        }
    }
    block_partial<K>(acc, partial, tile, g);
}

// p = z + beta p, z = r (no preconditioner) or dinv r (Jacobi)
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_pupdate(int NN, T* __restrict__ p, const T* __restrict__ r,
                                                    const ColState* __restrict__ cs, const int* __restrict__ tile_on,
                                                    const T* __restrict__ dinv) {
    constexpr int NPB = kBlock / K;
    int tile, g;
    tile_group(tile, g);
    const int lane = threadIdx.x % K, slot = threadIdx.x / K;
    const ColState& st = cs[(size_t)tile * K + lane];
    if (!tile_on[tile] || st.status != kActive) return;
    const T b = (T)st.beta;
    const size_t off = (size_t)tile * NN * 3 * K;
    for (int base = g * NPB; base < NN; base += (int)gridDim.x * NPB) {
        const int node = base + slot;
        if (node >= NN) continue;
        const size_t e = off + (size_t)node * 3 * K + lane;
        if (dinv) {
#pragma unroll
            for (int c = 0; c < 3; ++c) p[e + c * K] = dinv[(size_t)node * 3 + c] * r[e + c * K] + b * p[e + c * K];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) p[e + c * K] = r[e + c * K] + b * p[e + c * K];
        }
    }
}

// 16-byte vectors of the tile-interleaved layout for K = 64: a thread owns VW adjacent
// columns of one (node, component) row, 64 / VW threads cover a row.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    using type = double2;
    static constexpr int n = 2;
};
template <typename T>
__device__ __forceinline__ T vget(const typename Vec16<T>::type& v, int j) {
    if constexpr (Vec16<T>::n == 4) return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    else return j == 0 ? v.x : v.y;
}
template <typename T>
__device__ __forceinline__ void vset(typename Vec16<T>::type& v, int j, T a) {
    if constexpr (Vec16<T>::n == 4) {
        if (j == 0) v.x = a;
        else if (j == 1) v.y = a;
        else if (j == 2) v.z = a;
        else v.w = a;
    } else {
        if (j == 0) v.x = a;
        else v.y = a;
    }
}

// Streaming accesses of the 64-column vector kernels. EQD_NT: 0 = ordinary, 1 = non-temporal
// stores, 2 = non-temporal loads and stores (each vector is streamed once per kernel, so
// keeping it out of the caches leaves them to the stencil's neighbour lines; 2 measured
// fastest: 224 -> 209 ms for the dz = 50 m exact build).
#ifndef EQD_NT
#define EQD_NT 2
#endif
template <typename T>
struct NativeVec16;
template <>
struct NativeVec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct NativeVec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <typename T, typename V>
__device__ __forceinline__ V sld(const V* p) {
    if constexpr (EQD_NT >= 2) {
        using N = typename NativeVec16<T>::type;
        return __builtin_bit_cast(V, __builtin_nontemporal_load(reinterpret_cast<const N*>(p)));
    } else {
        return *p;
    }
}
template <typename T, typename V>
__device__ __forceinline__ void sst(V* p, const V& v) {
    if constexpr (EQD_NT >= 1) {
        using N = typename NativeVec16<T>::type;
        __builtin_nontemporal_store(__builtin_bit_cast(N, v), reinterpret_cast<N*>(p));
    } else {
        *p = v;
    }
}

// Per-column partials of the VW columns a thread owns, reduced over the block's rows in
// a fixed order into partial[tile][g][64].
template <int VW>
__device__ __forceinline__ void block_partial_vec(const double (&acc)[VW], double* __restrict__ partial, int tile,
                                                  int g) {
    constexpr int TPR = 64 / VW, ROWS = kBlock / TPR;
    __shared__ double sm[ROWS][64];
    const int t = threadIdx.x, c0 = (t % TPR) * VW;
#pragma unroll
    for (int j = 0; j < VW; ++j) sm[t / TPR][c0 + j] = acc[j];
    __syncthreads();
    if (t < 64) {
        double s = 0.0;
        for (int i = 0; i < ROWS; ++i) s += sm[i][t];
        partial[((size_t)tile * gridDim.x + g) * 64 + t] = s;
    }
}

// r -= alpha q and the weighted partial r.r, for 64-column tiles with 16-byte accesses.
// (x += alpha p is applied by k_pupdate64, which reads p anyway: 8 vector passes per
// iteration instead of 9.)
// PRE: Jacobi, the partial is r.z with z = dinv r (dinv per node and component, shared by
// the 64 columns of a row).
template <typename T, bool PRE>
__global__ __launch_bounds__(kBlock) void k_update64(int NN, int NX, int mface, T* __restrict__ r,
                                                     const T* __restrict__ q, const ColState* __restrict__ cs,
                                                     const int* __restrict__ tile_on, double* __restrict__ partial,
                                                     const T* __restrict__ dinv) {
    using V = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n, TPR = 64 / VW;
    int tile, g;
    tile_group(tile, g);
    const int t = threadIdx.x, c0 = (t % TPR) * VW;
    double acc[VW];
#pragma unroll
    for (int j = 0; j < VW; ++j) acc[j] = 0.0;
    if (tile_on[tile]) {
        T a[VW];
        bool act[VW];
#pragma unroll
        for (int j = 0; j < VW; ++j) {
            const ColState& st = cs[(size_t)tile * 64 + c0 + j];
            act[j] = st.status == kActive;
            a[j] = (T)st.alpha;
        }
        const size_t off = (size_t)tile * NN * 192;
        const size_t nv = (size_t)NN * 192 / VW;
        V* rv = reinterpret_cast<V*>(r + off);
        const V* qv = reinterpret_cast<const V*>(q + off);
        for (size_t v = (size_t)g * kBlock + t; v < nv; v += (size_t)gridDim.x * kBlock) {
            const int node = (int)(v * VW / 192);
            const double w = (node % NX == mface) ? 2.0 : 1.0;
            V rr = sld<T>(rv + v);
            const V qq = sld<T>(qv + v);
            T dz = T(1);
            if constexpr (PRE) dz = dinv[v * VW / 64];
#pragma unroll
            for (int j = 0; j < VW; ++j) {
                if (act[j]) {
                    const T rn = vget<T>(rr, j) - a[j] * vget<T>(qq, j);
                    vset<T>(rr, j, rn);
                    if constexpr (PRE) acc[j] += w * ((double)rn * (double)(dz * rn));
                    else acc[j] += w * ((double)rn * (double)rn);
                }
            }
            sst<T>(rv + v, rr);
        }
    }
    block_partial_vec<VW>(acc, partial, tile, g);
}

// x += alpha p (columns updated this iteration) and p = r + beta p (columns still
// active), for 64-column tiles with 16-byte accesses.
// PRE: Jacobi, p = dinv r + beta p.
template <typename T, bool PRE>
__global__ __launch_bounds__(kBlock) void k_pupdate64(int NN, T* __restrict__ x, T* __restrict__ p,
                                                      const T* __restrict__ r, const ColState* __restrict__ cs,
                                                      const T* __restrict__ dinv) {
    using V = typename Vec16<T>::type;
    constexpr int VW = Vec16<T>::n, TPR = 64 / VW;
    int tile, g;
    tile_group(tile, g);
    const int t = threadIdx.x, c0 = (t % TPR) * VW;
    T a[VW], b[VW];
    bool xu[VW], pu[VW];
    bool anyx = false, anyp = false;
#pragma unroll
    for (int j = 0; j < VW; ++j) {
        const ColState& st = cs[(size_t)tile * 64 + c0 + j];
        xu[j] = st.xpend != 0;
        pu[j] = st.status == kActive;
        a[j] = (T)st.alpha;
        b[j] = (T)st.beta;
        anyx |= xu[j];
        anyp |= pu[j];
    }
    if (!anyx && !anyp) return;
    const size_t off = (size_t)tile * NN * 192;
    const size_t nv = (size_t)NN * 192 / VW;
    V* xv = reinterpret_cast<V*>(x + off);
    V* pv = reinterpret_cast<V*>(p + off);
    const V* rv = reinterpret_cast<const V*>(r + off);
    for (size_t v = (size_t)g * kBlock + t; v < nv; v += (size_t)gridDim.x * kBlock) {
        V pp = sld<T>(pv + v);
        if (anyx) {
            V xx = sld<T>(xv + v);
#pragma unroll
            for (int j = 0; j < VW; ++j)
                if (xu[j]) vset<T>(xx, j, vget<T>(xx, j) + a[j] * vget<T>(pp, j));
            sst<T>(xv + v, xx);
        }
        if (anyp) {
            const V rr = sld<T>(rv + v);
            if constexpr (PRE) {
                const T dz = dinv[v * VW / 64];
#pragma unroll
                for (int j = 0; j < VW; ++j)
                    if (pu[j]) vset<T>(pp, j, dz * vget<T>(rr, j) + b[j] * vget<T>(pp, j));
            } else {
#pragma unroll
                for (int j = 0; j < VW; ++j)
                    if (pu[j]) vset<T>(pp, j, vget<T>(rr, j) + b[j] * vget<T>(pp, j));
            }
            sst<T>(pv + v, pp);
        }
    }
}

// Reliable update, part 1: x64 += xc, xc = 0 for flagged columns (every column when all != 0).
template <int K>
__global__ __launch_bounds__(kBlock) void k_xaccum(int NN, double* __restrict__ x, float* __restrict__ xc,
                                                   const ColState* __restrict__ cs, const int* __restrict__ tile_on,
                                                   int all, const int* __restrict__ mask) {
    constexpr int NPB = kBlock / K;
    int tile, g;
    tile_group(tile, g);
    const int lane = threadIdx.x % K, slot = threadIdx.x / K;
    if (mask) {  // exactly the masked columns (lanes retired by the serving loop)
        if (!mask[(size_t)tile * K + lane]) return;
    } else if (!all && (!tile_on[tile] || !cs[(size_t)tile * K + lane].rel_flag)) {
        return;
    }
    const size_t off = (size_t)tile * NN * 3 * K;
    for (int base = g * NPB; base < NN; base += (int)gridDim.x * NPB) {
        const int node = base + slot;
        if (node >= NN) continue;
        const size_t e = off + (size_t)node * 3 * K + lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[e + c * K] += (double)xc[e + c * K];
            xc[e + c * K] = 0.0f;
        }
    }
}

// Sum of a column's G group partials in group order (G is a multiple of 16): each chunk's
// 16 loads are issued before its adds, so the adds wait on one round trip per chunk.
template <int K>
__device__ __forceinline__ double sum_partials(const double* __restrict__ partial, int G, int tile, int lane) {
    const double* p = partial + (size_t)tile * G * K + lane;
    double s = 0.0;
    for (int g0 = 0; g0 < G; g0 += 16) {
        double v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = p[(size_t)(g0 + j) * K];
#pragma unroll
        for (int j = 0; j < 16; ++j) s += v[j];
    }
    return s;
}

// First level of the two-level partial sum of grids with many groups: block (b, tile) sums
// the partials of groups [b chunk, (b + 1) chunk) of every column, each of 16 strands taking
// every 16th group in order and the strands then added in order, into
// folded[(tile nb + b) K + lane]. k_finalize / k_init_state then sum the nb block totals (nb a
// multiple of 16). The order is fixed by the grid, like the one-level sum.
template <int K>
__global__ __launch_bounds__(K * 16) void k_fold_partials(int G, int chunk, const double* __restrict__ partial,
                                                          double* __restrict__ folded) {
    __shared__ double red[16 * K];
    const int b = blockIdx.x, tile = blockIdx.y, nb = (int)gridDim.x;
    const int lane = threadIdx.x % K, strand = threadIdx.x / K;
    const int g0 = b * chunk, g1 = min(G, g0 + chunk);
    const double* p = partial + (size_t)tile * G * K + lane;
    double acc = 0.0;
    for (int g = g0 + strand; g < g1; g += 16) acc += p[(size_t)g * K];
    red[strand * K + lane] = acc;
    __syncthreads();
    if (strand == 0) {
        double tot = 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) tot += red[j * K + lane];
        folded[((size_t)tile * nb + b) * K + lane] = tot;
    }
}

// One block per tile, K threads: per-column scalar steps of Krylov's cg!
// (src/ElasticitySplitNode.jl:379 -> Krylov 0.10.9) and the tile activity flags.
// mode 0: after the stencil (alpha, zero-curvature exit)
// mode 1: after the update (convergence test, beta, iteration count, itmax)
// mode 2: after a reliable-update residual (mixed precision only)
// Sum of a column's partials of groups [g0, g1) in group order, 16 loads in flight per round.
template <int K>
__device__ __forceinline__ double sum_partials_range(const double* __restrict__ partial, int G, int tile, int lane,
                                                     int g0, int g1) {
    const double* p = partial + (size_t)tile * G * K + lane;
    double s = 0.0;
    int g = g0;
    for (; g + 16 <= g1; g += 16) {
        double v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = p[(size_t)(g + j) * K];
#pragma unroll
        for (int j = 0; j < 16; ++j) s += v[j];
    }
    for (; g < g1; ++g) s += p[(size_t)g * K];
    return s;
}

// Strands of k_finalize: each sums a contiguous quarter of a column's group partials, the four
// sums are added in strand order, so the block's load rounds are a quarter of a one-strand sum.
constexpr int kFinStrands = 4;

// One block per tile, K x kFinStrands threads: per-column scalar steps of Krylov's cg!
// (src/ElasticitySplitNode.jl:379 -> Krylov 0.10.9) and the tile activity flags, the column's
// steps on strand 0.
// mode 0: after the stencil (alpha, zero-curvature exit)
// mode 1: after the update (convergence test, beta, iteration count, itmax)
// mode 2: after a reliable-update residual (mixed precision only)
template <int K>
__global__ __launch_bounds__(K * kFinStrands) void k_finalize(int G, int mode, int mixed, int rel_interval,
                                                              ColState* __restrict__ cs,
                                                              const double* __restrict__ partial,
                                                              int* __restrict__ tile_on, int* __restrict__ tile_rel,
                                                              int* __restrict__ active_count) {
    __shared__ double red[kFinStrands * K];
    const int tile = blockIdx.x, lane = threadIdx.x % K, strand = threadIdx.x / K;
    ColState& st = cs[(size_t)tile * K + lane];
    const bool active = st.status == kActive;  // read before strand 0 changes the state
    {
        const int per = (G + kFinStrands - 1) / kFinStrands;
        const int g0 = min(G, strand * per), g1 = min(G, g0 + per);
        red[strand * K + lane] = active ? sum_partials_range<K>(partial, G, tile, lane, g0, g1) : 0.0;
    }
    __syncthreads();
    const double meps = 2.220446049250313e-16;
    if (strand == 0) {
        if (mode == 0) st.xpend = 0;  // the previous iteration's p update has run
        if (active) {
            double s = red[lane];
#pragma unroll
            for (int j = 1; j < kFinStrands; ++j) s += red[j * K + lane];
            if (mode == 0) {
                if (s <= meps * st.pnorm2 && fabs(s) <= meps * st.pnorm2) {
                    st.status = kZeroCurv;
                } else {
                    st.alpha = st.rr / s;
                }
            } else if (mode == 1) {
                st.xpend = 1;  // updated this iteration, whatever the convergence test says
                const double rnorm = sqrt(s);
                const bool solved = rnorm <= st.eps || rnorm + 1.0 <= 1.0;
                st.iters += 1;
                st.since_rel += 1;
                st.rel_flag = 0;
                if (solved && !mixed) {
                    st.status = kSolved;
                } else {
                    const double beta = s / st.rr;
                    st.beta = beta;
                    st.pnorm2 = s + beta * beta * st.pnorm2;
                    st.rr = s;
                    // A column whose FP32 residual claims convergence is checked in FP64 only on
                    // iterations divisible by kCandidateEvery: the columns of a tile then share
                    // their checks instead of each triggering a full-tile FP64 pass. The schedule
                    // depends on the column's own iteration count only.
                    if (mixed && ((solved && st.iters % kCandidateEvery == 0) || st.since_rel >= rel_interval))
                        st.rel_flag = 1;
                    if (st.iters >= st.itmax) {
                        st.status = kTired;
                        st.rel_flag = mixed ? 1 : 0;  // fold the last correction into x
                    }
                }
            } else {
                if (st.rel_flag) {
                    const double rnorm = sqrt(s);
                    st.since_rel = 0;
                    if (rnorm <= st.eps || rnorm + 1.0 <= 1.0) st.status = kSolved;
                    else st.rr = s;
                }
            }
        } else if (mode == 2 && st.rel_flag) {
            st.since_rel = 0;  // a column that hit itmax in this iteration: its correction was folded
        }
        if (mode == 2) st.rel_flag = 0;
    }
    const int on = __syncthreads_or(strand == 0 && st.status == kActive);
    const int rel = __syncthreads_or(strand == 0 && st.rel_flag != 0);
    if (threadIdx.x == 0) {
        tile_on[tile] = on;
        if (mode == 1) tile_rel[tile] = rel;
        if (active_count && on) atomicAdd(active_count, 1);
    }
}

// Unit-slip fields on the face grid (n23 x n23 per column, two directions):
// column id < nf -> s2 = e_col, else s3 = e_(col - nf)   (src/FaultResponse.jl:249-253).
__global__ void k_unit_fields(int n23, int i2lo, int n2f, int nf, int ncols_slots, const int64_t* __restrict__ colid,
                              double* __restrict__ fields, const int* __restrict__ mask) {
    const int slot = blockIdx.x;
    if (slot >= ncols_slots) return;
    if (mask && !mask[slot]) return;
    double* f = fields + (size_t)slot * 2 * n23 * n23;
    for (int i = threadIdx.x; i < 2 * n23 * n23; i += blockDim.x) f[i] = 0.0;
    __syncthreads();
    const int64_t c = colid[slot];
    if (threadIdx.x == 0 && c >= 0) {
        const int dir = c < nf ? 0 : 1;
        const int k = (int)(dir == 0 ? c : c - nf);
        const int iy = i2lo + k % n2f, iz = i2lo + k / n2f;
        f[(size_t)dir * n23 * n23 + iy + (size_t)n23 * iz] = 1.0;
    }
}

__device__ __forceinline__ double tab_at(const Tab<double>& t, int row, int col) {
    const int j = col - t.start[row];
    return (j >= 0 && j < t.width[row]) ? t.c[(size_t)row * kTabW + j] : 0.0;
}

// b = H P (D+SAT) chi(s) on the merged grid, from slip fields on the face grid.
// chi = -+s/2 across the fault (src/FaultResponse.jl:176-185); the x1 structure of
// every term reduces to the jump coefficients Jt (tangential components, from s
// itself) and Jn (normal component, from the strike/dip derivatives of s).
// Also initialises the Krylov vectors: x = 0, xc = 0, r = b, p = z = b (dinv b with a
// Jacobi dinv).
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_rhs(RhsView rv, int NN, const double* __restrict__ fields, int nslots,
                                                double* __restrict__ b, double* __restrict__ x, T* __restrict__ r,
                                                T* __restrict__ p, float* __restrict__ xc, const int* __restrict__ mask,
                                                const double* __restrict__ dinv) {
    constexpr int NPB = kBlock / K;
    int tile, g;
    tile_group(tile, g);
    const int lane = threadIdx.x % K, slot = threadIdx.x / K;
    const int col = tile * K + lane;
    if (mask && !mask[col]) return;  // only the masked (newly admitted) columns
    const int n23 = rv.n23;
    const double* s2 = fields + (size_t)col * 2 * n23 * n23;
    const double* s3 = s2 + (size_t)n23 * n23;
    const size_t off = (size_t)tile * NN * 3 * K;
    for (int base = g * NPB; base < NN; base += (int)gridDim.x * NPB) {
        const int node = base + slot;
        if (node >= NN) continue;
        const int m = node % rv.NX, yz = node / rv.NX, y = yz % rv.NY, z = yz / rv.NY;
        const int iy = y + 1, iz = z + 1;  // full face indices
        double v[3] = {0.0, 0.0, 0.0};
        if (col < nslots) {
            const double H = rv.Hx[m] * rv.Hy[y] * rv.Hy[z];
            const double jt = rv.Jt[m], jn = rv.Jn[m];
            if (jt != 0.0) {
                v[1] = H * jt * s2[iy + (size_t)n23 * iz];
                v[2] = H * jt * s3[iy + (size_t)n23 * iz];
            }
            if (jn != 0.0) {
                double d = 0.0;
                const int sy = rv.D1full.start[y], wy = rv.D1full.width[y];
                for (int t = 0; t < wy; ++t) d += rv.D1full.c[(size_t)y * kTabW + t] * s2[(sy + t) + (size_t)n23 * iz];
                const int sz = rv.D1full.start[z], wz = rv.D1full.width[z];
                for (int t = 0; t < wz; ++t) d += rv.D1full.c[(size_t)z * kTabW + t] * s3[iy + (size_t)n23 * (sz + t)];
                v[0] = H * jn * d;
            }
        }
        const size_t e = off + (size_t)node * 3 * K + lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            b[e + c * K] = v[c];
            x[e + c * K] = 0.0;
            r[e + c * K] = (T)v[c];
            p[e + c * K] = dinv ? (T)(dinv[(size_t)node * 3 + c] * v[c]) : (T)v[c];
            if (xc) xc[e + c * K] = 0.0f;
        }
    }
}

// Initial per-column state from ||b|| (weighted): Krylov's setup before its loop.
template <int K>
__global__ void k_init_state(int G, ColState* __restrict__ cs, const double* __restrict__ partial,
                             const int64_t* __restrict__ colid, double rtol, double atol, int64_t itmax,
                             int* __restrict__ tile_on, const int* __restrict__ mask) {
    const int tile = blockIdx.x, lane = threadIdx.x;
    ColState& st = cs[(size_t)tile * K + lane];
    // Columns outside the mask (already in flight) keep their state. No early return: the
    // tile's activity below is one reduction over every lane of the block.
    if (!mask || mask[(size_t)tile * K + lane]) {
        const double s = sum_partials<K>(partial, G, tile, lane);
        st.col = colid[(size_t)tile * K + lane];
        st.rr = s;
        st.pnorm2 = s;
        st.bnorm = sqrt(s);
        st.eps = atol + rtol * st.bnorm;
        st.alpha = st.beta = 0.0;
        st.iters = 0;
        st.since_rel = 0;
        st.rel_flag = 0;
        st.xpend = 0;
        st.pad_ = 0;
        st.itmax = itmax;
        if (st.col < 0) st.status = kIdle;
        else if (s == 0.0 || st.bnorm <= st.eps) st.status = kSolved;  // x = 0 is the answer
        else st.status = kActive;
    }
    const int on = __syncthreads_or(st.status == kActive);
    if (lane == 0) tile_on[tile] = on;
}

// Weighted partial dot of one vector with itself (used for ||b||), or v.(dinv v) with a
// Jacobi dinv (Krylov's initial sqrt(b.M^{-1}b)).
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_selfdot(int NN, int NX, int mface, const T* __restrict__ v,
                                                    double* __restrict__ partial, const double* __restrict__ dinv) {
    constexpr int NPB = kBlock / K;
    int tile, g;
    tile_group(tile, g);
    const int lane = threadIdx.x % K, slot = threadIdx.x / K;
    const size_t off = (size_t)tile * NN * 3 * K;
    double acc = 0.0;
    for (int base = g * NPB; base < NN; base += (int)gridDim.x * NPB) {
        const int node = base + slot;
        if (node >= NN) continue;
        const size_t e = off + (size_t)node * 3 * K + lane;
        double d = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double vv = (double)v[e + c * K];
            if (dinv) d += vv * (dinv[(size_t)node * 3 + c] * vv);
            else d += vv * vv;
        }
        acc += ((node % NX == mface) ? 2.0 : 1.0) * d;
    }
    block_partial<K>(acc, partial, tile, g);
}

// x1-inner tiles (xinner.h): nodes per column quad of a (line, component) block. Lines of up to
// 16 nodes pad to 16, longer ones to a multiple of 8, so every column quad spans whole 128-byte
// lines of FP32 (and 256-byte lines of FP64).
__host__ __device__ inline int xi_nxb(int NX) { return NX <= 16 ? 16 : (NX + 7) & ~7; }

// Elements per tile vector of an x1-inner grid.
__host__ __device__ inline size_t xi_tile_elems(int NX, int NY) { return (size_t)NY * NY * 3 * 64 * xi_nxb(NX); }

// Traction change on Omega_f from the merged solution u and the slip fields:
// dtau_c = 1/2 (T_c U- + T_c U+), U = P u + chi     (src/FaultResponse.jl:201-209)
// Writes K[row, col] with row = (c-1)*nf + a + n2f*b (column-major, leading dim ldK).
// XI: u in the x1-inner layout of 64-column tiles (see xinner.h).
template <int K, bool XI = false>
__global__ void k_extract(ExtractView ev, int NN, const double* __restrict__ u, const double* __restrict__ fields,
                          const int64_t* __restrict__ colid, double* __restrict__ Kout, int64_t ldK,
                          int64_t col_offset, const int* __restrict__ mask) {
    const int tile = blockIdx.y;
    const int lane = threadIdx.x % K;
    const int k = blockIdx.x * (blockDim.x / K) + threadIdx.x / K;  // Omega_f node
    if (k >= ev.nf) return;
    if (mask && !mask[(size_t)tile * K + lane]) return;
    const int64_t col = colid[(size_t)tile * K + lane];
    if (col < 0) return;
    const int n23 = ev.n23, NX = ev.NX, NY = ev.NY;
    const int iy = ev.i2lo + k % ev.n2f, iz = ev.i2lo + k / ev.n2f;
    const bool yin = iy >= 1 && iy <= n23 - 2, zin = iz >= 1 && iz <= n23 - 2;
    const double* f = fields + ((size_t)tile * K + lane) * 2 * n23 * n23;
    const int B = xi_nxb(NX);  // x1 nodes per column quad of an x1-inner block
    const double* ut = u + (XI ? xi_tile_elems(NX, NY) : (size_t)NN * 3 * K) * tile;
    auto ld = [&](int m, int y, int z, int c) {
        if constexpr (XI)
            return ut[(size_t)((y + NY * z) * 3 + c) * 64 * B + (lane >> 2) * 4 * B + 4 * m + (lane & 3)];
        else
            return ut[((size_t)(m + NX * (y + NY * z)) * 3 + c) * K + lane];
    };
    double out[2];
    for (int d = 0; d < 2; ++d) {
        const int c = d + 1;
        double v = ev.Jext * f[(size_t)d * n23 * n23 + iy + (size_t)n23 * iz];
        if (yin && zin)
            for (int m = 0; m < NX; ++m) v += ev.Dn[m] * ld(m, iy - 1, iz - 1, c);
        double tg = 0.0;
        const int row = d == 0 ? iy : iz;
        const int s = ev.D1ext.start[row], w = ev.D1ext.width[row];
        if (d == 0 ? zin : yin)
            for (int t = 0; t < w; ++t) {
                const double cc = ev.D1ext.c[(size_t)row * kTabW + t];
                tg += cc * (d == 0 ? ld(ev.mface, s + t, iz - 1, 0) : ld(ev.mface, iy - 1, s + t, 0));
            }
        out[d] = v + ev.mu * tg;
    }
    double* Kc = Kout + (size_t)(col - col_offset) * ldK;
    Kc[k] = out[0];
    Kc[ev.nf + k] = out[1];
}

}  // namespace eqd

#include "pencil.h"
#include "xinner.h"

namespace eqd {

// Plain single-vector kernels (one column, K = 1) reuse the batched ones.

}  // namespace eqd
