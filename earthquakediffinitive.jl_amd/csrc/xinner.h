// x1-inner tile layout and its kernels (64-column tiles of the mixed-precision engine on grids
// whose merged x1 lines have 9 to 64 nodes). Included by kernels.h after pencil.h.
//
// Layout. A tile's vector holds, for every x2-x3 line (y, z) and component c, one block of
// 16 column quads x B x1 nodes x 4 columns, B = 16 for lines of at most 16 nodes and B = NX
// rounded up to a multiple of 8 for longer ones (nodes m >= NX zero padding), so that every
// column quad's run of nodes is whole 128-byte lines:
//     element (m, y, z, c, column j) = ((y + NY z) 3 + c) 64 B + (j / 4) 4 B + 4 m + j % 4.
// A 16-byte access holds four columns of one node, so 16 lanes with consecutive nodes read
// 256 contiguous bytes. The FP64 vectors b and x use the same element order.
//
// Stencil. Lanes are (x1 node n = lane % 16, column quad lane / 16): a wave holds 16 columns (one
// slab) of 16 nodes, so one 16-byte load per lane (32 bytes in FP64) fetches a line for the
// wave's columns. A wave marches the x3 planes of a unit of Z pencils (x2 row y, planes z0 ..
// z0 + Z - 1): a line loaded once serves every pencil of the unit that needs it, and the mixed
// x2-x3 term is sum-factorised per plane. The x1 operators act across the 16 lanes of a row: on
// the matrix cores in FP32, through DPP row shifts in FP64 (each lane holding the band row of
// its own node). The four waves of a block take four consecutive units of one slab, and the
// groups of a tile take its slabs in turn (slab-major), so the lines in flight on an XCD stay
// within one slab of one tile.
//
// Lines of more than 16 nodes (k_xl_pencil): the four waves of a block are the x1 segments of
// 16 nodes of one 16-column slab, and the x1 bands read their neighbours across segments from
// an LDS copy of the unit's lines.
//
// Vector kernels. A workgroup walks (line, component) blocks g, g + G, ...; thread t owns the
// slots of column quad t / 16 and nodes t % 16 + 16 s, so its columns (hence its per-column
// sums) are fixed.
#pragma once
#include <utility>

#ifndef EQD_XI_Z
#define EQD_XI_Z 3  // pencils per unit of the FP32 stencil
#endif
// The stencil takes the tiles in the reverse of the vector kernels' order, so it starts on the
// tiles the p update wrote last (dz = 50 m build 157.1 -> 154.8 ms; temporal instead of
// non-temporal p stores: 160.7 ms).
#ifndef EQD_XI_REV
#define EQD_XI_REV 1
#endif
#ifndef EQD_XI_WAVES
#define EQD_XI_WAVES 3
#endif
// EQD_XI_SLAB: the four waves of a block take four units of one 16-column slab (slab-major
// order over a tile, so the slab's x2-x3 planes stay in L2) instead of the four slabs of one
// unit. EQD_XI_SPREAD: those four units G/4 apart instead of consecutive.
#ifndef EQD_XI_SLAB
#define EQD_XI_SLAB 1
#endif
#ifndef EQD_XI_SPREAD
#define EQD_XI_SPREAD 0
#endif
// Units (x2 row y, x3 group of Z pencils) are ordered in bands of EQD_XI_YB rows: x3 groups
// of a band in turn, its rows fastest, so a unit's x3 neighbours run EQD_XI_YB units later.
#ifndef EQD_XI_YB
#define EQD_XI_YB 32
#endif
// Waves per block of the FP32 stencil (slab-major order): consecutive units of one slab.
#ifndef EQD_XI_UPD_U
#define EQD_XI_UPD_U 2  // (line, component) blocks per step of the r update
#endif
#ifndef EQD_XI_BW
#define EQD_XI_BW 4
#endif

namespace eqd {

template <typename T>
using xv4 = T __attribute__((ext_vector_type(4)));

// Four columns of one node through a buffer resource: 16 bytes (FP32) or 32 bytes (FP64).
template <typename T>
__device__ __forceinline__ xv4<T> xload(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(xv4<T>, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    } else {
        typedef unsigned int u8v __attribute__((ext_vector_type(8)));
        const auto lo = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
        const auto hi = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16, soff, 0);
        const u8v w = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        return __builtin_bit_cast(xv4<T>, w);
    }
}

// Value of lane n + D of the same 16-lane row (0 past the row's ends): DPP row_shl / row_shr.
template <int D>
__device__ __forceinline__ int xnb32(int v) {
    constexpr int ctrl = D > 0 ? (0x100 + D) : (0x110 - D);
    return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xF, 0xF, true);
}
template <int D, typename T>
__device__ __forceinline__ T xnb(T v) {
    if constexpr (D == 0) {
        return v;
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, xnb32<D>(__builtin_bit_cast(int, v)));
    } else {
        typedef int i2v __attribute__((ext_vector_type(2)));
        const i2v w = __builtin_bit_cast(i2v, v);
        const i2v s = {xnb32<D>(w[0]), xnb32<D>(w[1])};
        return __builtin_bit_cast(T, s);
    }
}

// x1 bands in source form: t[R + d] of lane m is the coefficient with which node m feeds node
// m - d, so sum_d b[n][d] v(n + d) = sum_d shift_d(t[d] v), and a line and an accumulated
// derivative that share a shift share one DPP move.
// o += sum_{d=-RA..RA} shift_d(ta[d] va + tb[d] vb), tb reaching RB <= RA.
template <int RA, int RB, int D = -RA, typename T>
__device__ __forceinline__ T xpair(const T* ta, const T* tb, T va, T vb, T o) {
    if constexpr (D > RA) {
        return o;
    } else {
        T t = ta[RA + D] * va;
        if constexpr (D >= -RB && D <= RB) t = tb[RB + D] * vb + t;
        o += xnb<D>(t);
        // FP64: the running sum is materialised after every tap, so each shifted product is
        // consumed at once instead of all of them being held for a deferred sum (register spills)
        if constexpr (sizeof(T) == 8) asm volatile("" : "+v"(o));
        return xpair<RA, RB, D + 1>(ta, tb, va, vb, o);
    }
}

template <int RA, int RB, typename T>
__device__ __forceinline__ void xpair4(const T* ta, const T* tb, const xv4<T>& va, const xv4<T>& vb, xv4<T>& o) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        o[e] = xpair<RA, RB>(ta, tb, va[e], vb[e], o[e]);
        EQD_FENCE();
    }
}

// FP32 x1 bands on the matrix cores. For the 16 columns of a wave, out^T (16 columns x 16 nodes)
// += V^T (16 x 16) Band^T is four v_mfma_f32_16x16x4f32 steps per band: the accumulator lane
// layout (lane = node + 16 (column / 4), register = column % 4) is the x1-inner one, so the
// accumulated x2 / x3 terms enter as C directly. The A operand (lane = column + 16 k, the value
// of node 4 s + k at step s) comes from a transpose through a per-wave LDS tile, in which node
// n sits at position 4 (n % 4) + n / 4 of its column's row, so each lane reads its four steps
// with one 16-byte load. The B operand (lane = node j + 16 k: Band[j][4 s + k]) is constant.
constexpr int kXtRow = 20;  // floats per column row of the transpose tile (16 + 4: bank spread)

__device__ __forceinline__ float xband_at(const float* band, int R, int j, int m) {
    const int d = m - j;
    return (d >= -R && d <= R) ? band[j * (2 * R + 1) + R + d] : 0.0f;
}

// A operands of one input (lane layout: node n, column quad cq, four columns in registers).
__device__ __forceinline__ xv4<float> xt_transpose(float* tile, const xv4<float>& v, int n, int cq, int lane) {
    const int pos = 4 * (n & 3) + (n >> 2);
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[(4 * cq + r) * kXtRow + pos] = v[r];
    return *reinterpret_cast<const xv4<float>*>(tile + (lane & 15) * kXtRow + 4 * (lane >> 4));
}

__device__ __forceinline__ xv4<float> xt_mfma(const xv4<float>& at, const float (&bt)[4], xv4<float> c) {
#pragma unroll
    for (int st = 0; st < 4; ++st) c = __builtin_amdgcn_mfma_f32_16x16x4f32(at[st], bt[st], c, 0, 0, 0);
    return c;
}

// Per-unit state shared by the plane steps of k_xi_pencil.
template <typename T, int Z>
struct XiUnit {
    const PencilView<T>& op;
    __amdgpu_buffer_rsrc_t rs;
    int voff, blk, NY, y, z0, nk;  // blk: elements of a (line, component) block
    const T* Tsy;
    const T* Toy;
    T cTs[4], cTo[4], cD1[4], cDs[4];  // x2 band taps of row y at dy = -2, -1, 1, 2
    int rr[4];                         // rows y + dy, clamped
    __device__ int line(int yy, int zz, int c) const { return ((yy + NY * zz) * 3 + c) * blk * (int)sizeof(T); }
    __device__ xv4<T> ld(int yy, int zz, int c) const { return xload<T>(rs, voff, line(yy, zz, c)); }
};

// Terms of a plane that read row y (c0, c1, c2) and components 1 of rows y +- 1, 2 (v1, with
// their D1y sum w1): phase A of xi_plane.
template <typename T, int Z, int P>
__device__ __forceinline__ void xi_terms_a(const XiUnit<T, Z>& U, const xv4<T>& c0, const xv4<T>& c1,
                                           const xv4<T>& c2, const xv4<T> (&v1)[4], xv4<T> (&o)[Z][3],
                                           xv4<T> (&a)[Z][3]) {
    xv4<T> w1 = U.cDs[0] * v1[0];
#pragma unroll
    for (int j = 1; j < 4; ++j) w1 += U.cDs[j] * v1[j];
#pragma unroll
    for (int k = 0; k < Z; ++k) {
        if (k >= U.nk) continue;
        const int zk = U.z0 + k;
        const int dz = P - kRyO - k;  // compile-time after unrolling
        const T* Tsz = U.op.Ts + (size_t)zk * (2 * kRyS + 1) + kRyS;
        const T* Toz = U.op.To + (size_t)zk * (2 * kRyO + 1) + kRyO;
        const T* D1z = U.op.D1 + (size_t)zk * (2 * kRyO + 1) + kRyO;
        if (dz == 0) {
            // own lines: the centre taps of the x2 and x3 bands, and the x2 terms of component 1
            const T cy0 = ucoef(U.Toy), cy1 = ucoef(U.Tsy), cz0 = ucoef(Toz), cz2 = ucoef(Tsz);
            o[k][0] += (cy0 + cz0) * c0;
            o[k][1] += (cy1 + cz0) * c1;
            o[k][2] += (cy0 + cz2) * c2;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                o[k][1] += U.cTs[j] * v1[j];
                a[k][0] += U.cD1[j] * v1[j];
            }
        } else if (dz >= -kRyO && dz <= kRyO) {
            // x3 terms of row y and the mixed x2-x3 term of component 1 (the D1 rows of interior
            // nodes have a zero centre tap, so dz = 0 adds nothing to the latter)
            const T tS = ucoef(Tsz + dz), tO = ucoef(Toz + dz), tD = ucoef(D1z + dz);
            o[k][2] += tS * c2;
            a[k][0] += tD * c2;
            o[k][0] += tO * c0;
            a[k][2] += tD * c0;
            o[k][1] += tO * c1;
            o[k][2] += tD * w1;
        } else if (dz >= -kRyS && dz <= kRyS) {
            const T tS = ucoef(Tsz + dz);  // wide closure rows of the x3 band
            if (tS != T(0)) o[k][2] += tS * c2;
        }
    }
}

// Terms of a plane that read components 2 (v2, D1y sum w2) and 0 (v0, on the pencils' own
// planes) of rows y +- 1, 2: phase B of xi_plane.
template <typename T, int Z, int P>
__device__ __forceinline__ void xi_terms_b(const XiUnit<T, Z>& U, const xv4<T> (&v2)[4], const xv4<T> (&v0)[4],
                                           xv4<T> (&o)[Z][3], xv4<T> (&a)[Z][3]) {
    xv4<T> w2 = U.cDs[0] * v2[0];
#pragma unroll
    for (int j = 1; j < 4; ++j) w2 += U.cDs[j] * v2[j];
#pragma unroll
    for (int k = 0; k < Z; ++k) {
        if (k >= U.nk) continue;
        const int zk = U.z0 + k;
        const int dz = P - kRyO - k;
        if (dz == 0) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                o[k][0] += U.cTo[j] * v0[j];
                a[k][1] += U.cD1[j] * v0[j];
                o[k][2] += U.cTo[j] * v2[j];
            }
        } else if (dz >= -kRyO && dz <= kRyO) {
            const T tD = ucoef(U.op.D1 + (size_t)zk * (2 * kRyO + 1) + kRyO + dz);
            o[k][1] += tD * w2;
        }
    }
}

// One x3 plane zp = z0 - 2 + P of a unit: load its lines (row y: every component; rows y +- 1, 2:
// components 1 and 2, and component 0 when a pencil of the unit lies on the plane) and apply
// them to each pencil k of the unit, the x3 offset dz = P - 2 - k being a compile-time constant.
// The (lam+mu) D1y sums of components 1 and 2 are formed once per plane and D1z applies them
// to each pencil (mixed x2-x3 term). The x1 bands of the pencils' own lines act at the end of
// the unit (k_xi_pencil), together with the mixed x1 terms. FP64 loads the plane in two phases
// (row y and component 1, then components 2 and 0) to halve the registers held by lines.
template <typename T, int Z, int P>
__device__ __forceinline__ void xi_plane(const XiUnit<T, Z>& U, xv4<T> (&o)[Z][3], xv4<T> (&a)[Z][3]) {
    EQD_FENCE();
    const int zp = U.z0 - kRyO + P;
    if (zp < 0 || zp >= U.NY) return;
    constexpr int KP = P - kRyO;  // the unit's pencil on this plane, if 0 <= KP < Z
    const bool own = KP >= 0 && KP < Z && KP < U.nk;
    const xv4<T> c0 = U.ld(U.y, zp, 0), c1 = U.ld(U.y, zp, 1), c2 = U.ld(U.y, zp, 2);
    xv4<T> v1[4], v2[4], v0[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) v1[j] = U.ld(U.rr[j], zp, 1);
    if constexpr (sizeof(T) == 8) {
        EQD_FENCE();
        xi_terms_a<T, Z, P>(U, c0, c1, c2, v1, o, a);
        EQD_FENCE();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) v2[j] = U.ld(U.rr[j], zp, 2);
    if (own) {
#pragma unroll
        for (int j = 0; j < 4; ++j) v0[j] = U.ld(U.rr[j], zp, 0);
    }
    EQD_FENCE();
    if constexpr (sizeof(T) == 4) xi_terms_a<T, Z, P>(U, c0, c1, c2, v1, o, a);
    xi_terms_b<T, Z, P>(U, v2, v0, o, a);
}

// Unit u of a grid of NY rows and nzg x3 groups of Z pencils: its row y and first plane z0.
template <int Z>
__device__ __forceinline__ void xi_unit(int u, int NY, int nzg, int& y, int& z0) {
    const int band = u / (EQD_XI_YB * nzg), r = u - band * (EQD_XI_YB * nzg);
    const int yb = min(EQD_XI_YB, NY - band * EQD_XI_YB);  // rows of this band
    const int zg = r / yb;
    y = band * EQD_XI_YB + (r - zg * yb);
    z0 = zg * Z;
}

template <typename T, int Z, int... P>
__device__ __forceinline__ void xi_planes(const XiUnit<T, Z>& U, xv4<T> (&o)[Z][3], xv4<T> (&a)[Z][3],
                                          std::integer_sequence<int, P...>) {
    (xi_plane<T, Z, P>(U, o, a), ...);
}

template <typename T, int MODE, int Z>
constexpr int xi_waves() {
    return sizeof(T) == 8 ? 3 : EQD_XI_WAVES;
}

// MODE 0: q = A in and the weighted partial in.q per column (T = float).
// MODE 1: reliable-update residual r = b - A x (T = double), stored as R for the flagged
//         columns, with the weighted partial r.r (r.(dinv r) with a Jacobi dinv).
// Grid: (G, tiles) with G a multiple of 64; block g takes units g, g + G, ...
template <typename T, typename R, int MODE, int Z>
__global__ __launch_bounds__(64 * (MODE == 0 ? EQD_XI_BW : 4)) __attribute__((amdgpu_waves_per_eu(xi_waves<T, MODE, Z>()))) void k_xi_pencil(
    PencilView<T> op, const T* __restrict__ in, T* __restrict__ q, const double* __restrict__ b,
    R* __restrict__ rout, const ColState* __restrict__ cs, const int* __restrict__ tile_on,
    double* __restrict__ partial, const double* __restrict__ dinv) {
    int tile, g;
    tile_group(tile, g);
#if EQD_XI_REV
    tile = (int)gridDim.y - 1 - tile;
#endif
    if (!tile_on[tile]) return;
#if EQD_XI_SLAB
    if constexpr (MODE == 1) {
        // the FP64 residual: a block whose 16-column slab holds no column flagged for a reliable
        // update leaves at once (its partial slots zeroed; only flagged columns read theirs)
        const int l = threadIdx.x & 63, sl = g / ((int)gridDim.x >> 2);
        if (!__any(l < 16 && cs[(size_t)tile * 64 + 16 * sl + l].rel_flag != 0)) {
            if (threadIdx.x < 64) partial[((size_t)tile * gridDim.x + g) * 64 + threadIdx.x] = 0.0;
            return;
        }
    }
#endif
    constexpr int BW = MODE == 0 ? EQD_XI_BW : 4;  // waves per block
    static_assert(BW == 4 || EQD_XI_SLAB, "other block sizes take the slab-major order");
    constexpr int WS = 2 * kRxS + 1, WO = 2 * kRxO + 1;
    __shared__ T xs[16 * WS], xo[16 * WO], xm[16 * WO];
    // per-wave transpose tiles of the matrix-core x1 bands (FP32): six inputs per pencil
    __shared__ __attribute__((aligned(16))) float xtile[sizeof(T) == 4 ? BW * 6 * 16 * kXtRow : 1];
    __shared__ double sm[BW][64];
    // source form of the x1 bands: entry (m, d) = band row m - d, tap d (0 off the 16 rows)
    for (int i = threadIdx.x; i < 16 * WS; i += 64 * BW) {
        const int m = i / WS, d = i % WS - kRxS, r = m - d;
        xs[i] = (r >= 0 && r < 16) ? op.AXs[r * WS + kRxS + d] : T(0);
    }
    for (int i = threadIdx.x; i < 16 * WO; i += 64 * BW) {
        const int m = i / WO, d = i % WO - kRxO, r = m - d;
        xo[i] = (r >= 0 && r < 16) ? op.AXo[r * WO + kRxO + d] : T(0);
        xm[i] = (r >= 0 && r < 16) ? op.MX[r * WO + kRxO + d] : T(0);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = lane & 15, cq = lane >> 4;
    const int NX = op.NX, NY = op.NY;
    const bool live = n < NX;
    float bAs[4] = {}, bAo[4] = {}, bMx[4] = {};  // B operands: Band[lane % 16][4 s + lane / 16]
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            const int m = 4 * st + (lane >> 4);
            bAs[st] = xband_at(op.AXs, kRxS, n, m);
            bAo[st] = xband_at(op.AXo, kRxO, n, m);
            bMx[st] = xband_at(op.MX, kRxO, n, m);
        }
    }
    const size_t tfl = xi_tile_elems(NX, NY);
    const size_t off = (size_t)tile * tfl;
#if EQD_XI_SLAB
    const int G4 = (int)gridDim.x >> 2;
    const int slab = g / G4, gs = g % G4;
#else
    const int slab = wv;
#endif
    const int col0 = 16 * slab + 4 * cq;  // the lane's columns col0 .. col0 + 3
    const int lo = ((4 * slab + cq) * 16 + n) * 4;  // their element offset in a (line, component) block
    bool mine[4] = {true, true, true, true};
    if constexpr (MODE == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) mine[i] = cs[(size_t)tile * 64 + col0 + i].rel_flag != 0;
    }
    const T mw = (live && n == op.mface) ? T(2) : T(1);
    const T hx = live ? op.Hx[n] : T(0);
    XiUnit<T, Z> U{op, eqd_rsrc(in + off, (uint32_t)(tfl * sizeof(T))), lo * (int)sizeof(T), 1024, NY, 0, 0, 0,
                   nullptr, nullptr, {}, {}, {}, {}, {}};
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int nzg = (NY + Z - 1) / Z, nunits = NY * nzg;
#if EQD_XI_SLAB
    for (int i = 0;; ++i) {
        const int u0 = EQD_XI_SPREAD ? gs + G4 * (wv + BW * i) : BW * (gs + G4 * i) + wv;
        if (u0 >= nunits) break;
#else
    for (int u0 = g; u0 < nunits; u0 += (int)gridDim.x) {
#endif
        const int u = uni(u0);
        int y, z0;
        xi_unit<Z>(u, NY, nzg, y, z0);
        U.y = y;
        U.z0 = z0;
        U.nk = min(Z, NY - z0);
        U.Tsy = op.Ts + (size_t)y * (2 * kRyS + 1) + kRyS;
        U.Toy = op.To + (size_t)y * (2 * kRyO + 1) + kRyO;
        const T* D1y = op.D1 + (size_t)y * (2 * kRyO + 1) + kRyO;
        const T* D1sy = op.D1s + (size_t)y * (2 * kRyO + 1) + kRyO;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int dy = j < 2 ? j - 2 : j - 1;  // -2, -1, 1, 2
            U.cTs[j] = ucoef(U.Tsy + dy);
            U.cTo[j] = ucoef(U.Toy + dy);
            U.cD1[j] = ucoef(D1y + dy);
            U.cDs[j] = ucoef(D1sy + dy);
            U.rr[j] = rclamp(y + dy, NY);
        }
        xv4<T> o[Z][3], a[Z][3];
#pragma unroll
        for (int k = 0; k < Z; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                o[k][c] = xv4<T>{T(0), T(0), T(0), T(0)};
                a[k][c] = xv4<T>{T(0), T(0), T(0), T(0)};
            }
        // planes z0 - 2 .. z0 + Z + 1
        xi_planes<T, Z>(U, o, a, std::make_integer_sequence<int, Z + 2 * kRyO>{});
        EQD_FENCE();
        // closure reaches beyond the interior bands: x2 rows y +- 3, 4 (Ts of component 1) on the
        // pencils' planes and x3 planes z0 - 4, z0 - 3, z0 + Z + 2, z0 + Z + 3 (Ts of component 2)
        const int nk = U.nk;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int dy = j < 2 ? j - 4 : j + 1;  // -4, -3, 3, 4
            const T c = ucoef(U.Tsy + dy);
            if (c != T(0)) {
                const int yy = rclamp(y + dy, NY);
#pragma unroll
                for (int k = 0; k < Z; ++k)
                    if (k < nk) o[k][1] += c * U.ld(yy, z0 + k, 1);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int zp = j < 2 ? z0 - 4 + j : z0 + Z + j;  // z0 - 4, z0 - 3, z0 + Z + 2, z0 + Z + 3
            if (zp < 0 || zp >= NY) continue;
            bool any = false;
            T ck[Z];
#pragma unroll
            for (int k = 0; k < Z; ++k) {
                const int dz = zp - (z0 + k);
                ck[k] = (k < nk && dz >= -kRyS && dz <= kRyS)
                            ? ucoef(op.Ts + (size_t)(z0 + k) * (2 * kRyS + 1) + kRyS + dz)
                            : T(0);
                any = any || ck[k] != T(0);
            }
            if (any) {
                const xv4<T> v = U.ld(y, zp, 2);
#pragma unroll
                for (int k = 0; k < Z; ++k) o[k][2] += ck[k] * v;
            }
        }
        // the own lines again (cache hits): their x1 bands together with (lam+mu) D1x of the
        // accumulated x2 / x3 derivatives; weights, stores and the partial dot
        const T hy = op.Hy[y];
#pragma unroll
        for (int k = 0; k < Z; ++k) {
            if (k >= nk) continue;
            const int zk = z0 + k;
            const xv4<T> p0 = U.ld(y, zk, 0), p1 = U.ld(y, zk, 1), p2 = U.ld(y, zk, 2);
            if constexpr (sizeof(T) == 4) {
                float* xt = xtile + wv * 6 * 16 * kXtRow;
                const xv4<float> t0 = xt_transpose(xt, p0, n, cq, lane);
                const xv4<float> t1 = xt_transpose(xt + 16 * kXtRow, p1, n, cq, lane);
                const xv4<float> t2 = xt_transpose(xt + 2 * 16 * kXtRow, p2, n, cq, lane);
                const xv4<float> u0 = xt_transpose(xt + 3 * 16 * kXtRow, a[k][0], n, cq, lane);
                const xv4<float> u1 = xt_transpose(xt + 4 * 16 * kXtRow, a[k][1], n, cq, lane);
                const xv4<float> u2 = xt_transpose(xt + 5 * 16 * kXtRow, a[k][2], n, cq, lane);
                o[k][0] = xt_mfma(u0, bMx, xt_mfma(t0, bAs, o[k][0]));
                o[k][1] = xt_mfma(u1, bMx, xt_mfma(t1, bAo, o[k][1]));
                o[k][2] = xt_mfma(u2, bMx, xt_mfma(t2, bAo, o[k][2]));
            } else {
                // band offsets made opaque here, so the coefficients are read from LDS where they are
                // used instead of being held in registers across the unit loop
                int os = n * WS, oo = n * WO;
                asm volatile("" : "+v"(os), "+v"(oo));
                xpair4<kRxS, kRxO>(xs + os, xm + oo, p0, a[k][0], o[k][0]);
                xpair4<kRxO, kRxO>(xo + oo, xm + oo, p1, a[k][1], o[k][1]);
                xpair4<kRxO, kRxO>(xo + oo, xm + oo, p2, a[k][2], o[k][2]);
            }
            const T hw = -(hx * (hy * op.Hy[zk]));
            const size_t e = off + (size_t)(y + NY * zk) * 3 * 1024 + lo;
            double d[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const xv4<T> v = hw * o[k][c];
                if constexpr (MODE == 0) {
                    const xv4<T> pv = c == 0 ? p0 : (c == 1 ? p1 : p2);
                    // padding lanes store zeros: every 256-byte segment of q is written whole
                    __builtin_nontemporal_store(live ? v : xv4<T>{T(0), T(0), T(0), T(0)},
                                                (xv4<T>*)(q + e + c * 1024));
#pragma unroll
                    for (int i = 0; i < 4; ++i) d[i] += (double)pv[i] * (double)v[i];
                } else {
                    const double dz = dinv ? dinv[(size_t)(n + NX * (y + NY * zk)) * 3 + c] : 1.0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const double rv = b[e + c * 1024 + i] - (double)v[i];
                        if (live && mine[i]) rout[e + c * 1024 + i] = (R)rv;
                        d[i] += rv * (dz * rv);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] += (double)mw * d[i];
        }
    }
    // per column: sum over the 16 node lanes of a row (fixed butterfly), one slot per column
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double s = live ? acc[i] : 0.0;
        s += __shfl_xor(s, 8, 16);
        s += __shfl_xor(s, 4, 16);
        s += __shfl_xor(s, 2, 16);
        s += __shfl_xor(s, 1, 16);
        if (n == 0) sm[wv][4 * cq + i] = s;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int j = threadIdx.x;
#if EQD_XI_SLAB
        // the four waves' sums in a fixed order; the other slabs' columns get 0
        double v = 0.0;
        if ((j >> 4) == slab) {
            v = sm[0][j & 15];
#pragma unroll
            for (int w = 1; w < BW; ++w) v += sm[w][j & 15];
        }
#else
        const double v = sm[j >> 4][j & 15];
#endif
        partial[((size_t)tile * gridDim.x + g) * 64 + j] = v;
    }
}

// ---------------------------------------------------------------------------------------------
// Lines of 17 to 64 x1 nodes. A block's four waves are the x1 segments (nodes 16 w ..
// 16 w + 15) of one 16-column slab and march the same unit of Z pencils. The y/z terms are those
// of k_xi_pencil; the x1 bands need neighbours across segments, so at the end of a pencil every
// wave copies its own lines and accumulated derivatives into an LDS line (with zero rows past
// the ends), and each lane applies its node's band rows to the neighbours it reads there.
constexpr int kXlHalo = 8;                   // zero rows on each side of the LDS line (>= the x1 reach)
constexpr int kXlRows = 64 + 2 * kXlHalo;    // rows of one LDS line: up to 64 nodes + halos

template <typename T, int R>
__device__ __forceinline__ void xl_band(const T* rowc, const T* line, int m, int cq, xv4<T>& o) {
#pragma unroll
    for (int d = -R; d <= R; ++d) {
        const T c = rowc[R + d];
        const xv4<T> v = *reinterpret_cast<const xv4<T>*>(line + ((kXlHalo + m + d) * 16 + 4 * cq));
        o += c * v;
    }
}

template <typename T, typename R, int MODE, int Z>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 3))) void k_xl_pencil(
    PencilView<T> op, const T* __restrict__ in, T* __restrict__ q, const double* __restrict__ b,
    R* __restrict__ rout, const ColState* __restrict__ cs, const int* __restrict__ tile_on,
    double* __restrict__ partial, const double* __restrict__ dinv) {
    int tile, g;
    tile_group(tile, g);
#if EQD_XI_REV
    tile = (int)gridDim.y - 1 - tile;
#endif
    if (!tile_on[tile]) return;
    constexpr int WS = 2 * kRxS + 1, WO = 2 * kRxO + 1;
    const int NX = op.NX, NY = op.NY;
    // band rows of the 64 node slots (zero past NX), and the six LDS lines of a pencil
    __shared__ T bS[64 * WS], bO[64 * WO], bM[64 * WO];
    __shared__ __attribute__((aligned(16))) T xl[6 * kXlRows * 16];
    __shared__ double sm[4][16];
    for (int i = threadIdx.x; i < 64 * WS; i += kBlock) bS[i] = i / WS < NX ? op.AXs[i] : T(0);
    for (int i = threadIdx.x; i < 64 * WO; i += kBlock) {
        bO[i] = i / WO < NX ? op.AXo[i] : T(0);
        bM[i] = i / WO < NX ? op.MX[i] : T(0);
    }
    for (int i = threadIdx.x; i < 6 * kXlRows * 16; i += kBlock) xl[i] = T(0);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int n = lane & 15, cq = lane >> 4;
    const int S = (NX + 15) >> 4;        // segments of the line
    const bool wlive = wv < S;            // the wave has nodes
    const int m = 16 * wv + n;            // the lane's node
    const bool live = m < NX;
    const int mc = live ? m : NX - 1;
    const int G4 = (int)gridDim.x >> 2;
    const int sl = g / G4;                // the block's 16-column slab (slab-major over the tile)
    const int qd = 4 * sl + cq;           // the lane's column quad
    const size_t tfl = xi_tile_elems(NX, NY);
    const size_t off = (size_t)tile * tfl;
    const int B = xi_nxb(NX);
    const int blk = 64 * B;
    const int lo = qd * 4 * B + 4 * mc;  // the lane's element offset in a (line, component) block
    const bool st = m < B;               // lanes m < B store (zeros past NX): whole 128-byte lines
    bool mine[4] = {true, true, true, true};
    if constexpr (MODE == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) mine[i] = cs[(size_t)tile * 64 + 4 * qd + i].rel_flag != 0;
    }
    const T mw = (live && m == op.mface) ? T(2) : T(1);
    const T hx = live ? op.Hx[m] : T(0);
    XiUnit<T, Z> U{op, eqd_rsrc(in + off, (uint32_t)(tfl * sizeof(T))), lo * (int)sizeof(T), blk, NY, 0, 0, 0,
                   nullptr, nullptr, {}, {}, {}, {}, {}};
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int nzg = (NY + Z - 1) / Z, nunits = NY * nzg;
    for (int u0 = g % G4; u0 < nunits; u0 += G4) {
        const int u = uni(u0);
        int y, z0;
        xi_unit<Z>(u, NY, nzg, y, z0);
        U.y = y;
        U.z0 = z0;
        U.nk = min(Z, NY - z0);
        U.Tsy = op.Ts + (size_t)y * (2 * kRyS + 1) + kRyS;
        U.Toy = op.To + (size_t)y * (2 * kRyO + 1) + kRyO;
        const T* D1y = op.D1 + (size_t)y * (2 * kRyO + 1) + kRyO;
        const T* D1sy = op.D1s + (size_t)y * (2 * kRyO + 1) + kRyO;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int dy = j < 2 ? j - 2 : j - 1;
            U.cTs[j] = ucoef(U.Tsy + dy);
            U.cTo[j] = ucoef(U.Toy + dy);
            U.cD1[j] = ucoef(D1y + dy);
            U.cDs[j] = ucoef(D1sy + dy);
            U.rr[j] = rclamp(y + dy, NY);
        }
        xv4<T> o[Z][3], a[Z][3];
#pragma unroll
        for (int k = 0; k < Z; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                o[k][c] = xv4<T>{T(0), T(0), T(0), T(0)};
                a[k][c] = xv4<T>{T(0), T(0), T(0), T(0)};
            }
        const int nk = U.nk;
        if (wlive) {
            xi_planes<T, Z>(U, o, a, std::make_integer_sequence<int, Z + 2 * kRyO>{});
            EQD_FENCE();
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int dy = j < 2 ? j - 4 : j + 1;
                const T c = ucoef(U.Tsy + dy);
                if (c != T(0)) {
                    const int yy = rclamp(y + dy, NY);
#pragma unroll
                    for (int k = 0; k < Z; ++k)
                        if (k < nk) o[k][1] += c * U.ld(yy, z0 + k, 1);
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int zp = j < 2 ? z0 - 4 + j : z0 + Z + j;
                if (zp < 0 || zp >= NY) continue;
                bool any = false;
                T ck[Z];
#pragma unroll
                for (int k = 0; k < Z; ++k) {
                    const int dz = zp - (z0 + k);
                    ck[k] = (k < nk && dz >= -kRyS && dz <= kRyS)
                                ? ucoef(op.Ts + (size_t)(z0 + k) * (2 * kRyS + 1) + kRyS + dz)
                                : T(0);
                    any = any || ck[k] != T(0);
                }
                if (any) {
                    const xv4<T> v = U.ld(y, zp, 2);
#pragma unroll
                    for (int k = 0; k < Z; ++k) o[k][2] += ck[k] * v;
                }
            }
        }
        const T hy = op.Hy[y];
#pragma unroll
        for (int k = 0; k < Z; ++k) {
            if (k >= nk) continue;
            const int zk = z0 + k;
            xv4<T> p0{}, p1{}, p2{};
            if (wlive) {
                p0 = U.ld(y, zk, 0);
                p1 = U.ld(y, zk, 1);
                p2 = U.ld(y, zk, 2);
                if (live) {
                    const int r0 = (kXlHalo + m) * 16 + 4 * cq;
                    *reinterpret_cast<xv4<T>*>(xl + 0 * kXlRows * 16 + r0) = p0;
                    *reinterpret_cast<xv4<T>*>(xl + 1 * kXlRows * 16 + r0) = p1;
                    *reinterpret_cast<xv4<T>*>(xl + 2 * kXlRows * 16 + r0) = p2;
                    *reinterpret_cast<xv4<T>*>(xl + 3 * kXlRows * 16 + r0) = a[k][0];
                    *reinterpret_cast<xv4<T>*>(xl + 4 * kXlRows * 16 + r0) = a[k][1];
                    *reinterpret_cast<xv4<T>*>(xl + 5 * kXlRows * 16 + r0) = a[k][2];
                }
            }
            __syncthreads();  // the unit's lines are complete in LDS
            if (wlive) {
                const int ms = live ? m : 0;
                xl_band<T, kRxS>(bS + ms * WS, xl + 0 * kXlRows * 16, ms, cq, o[k][0]);
                xl_band<T, kRxO>(bM + ms * WO, xl + 3 * kXlRows * 16, ms, cq, o[k][0]);
                EQD_FENCE();
                xl_band<T, kRxO>(bO + ms * WO, xl + 1 * kXlRows * 16, ms, cq, o[k][1]);
                xl_band<T, kRxO>(bM + ms * WO, xl + 4 * kXlRows * 16, ms, cq, o[k][1]);
                EQD_FENCE();
                xl_band<T, kRxO>(bO + ms * WO, xl + 2 * kXlRows * 16, ms, cq, o[k][2]);
                xl_band<T, kRxO>(bM + ms * WO, xl + 5 * kXlRows * 16, ms, cq, o[k][2]);
            }
            __syncthreads();  // every wave has read them before the next pencil's copies
            if (!wlive) continue;
            const T hw = -(hx * (hy * op.Hy[zk]));
            const size_t e = off + (size_t)(y + NY * zk) * 3 * blk + lo;  // node mc
            const int es = 4 * (m - mc);  // to the lane's own node (padding lanes m < B)
            double d[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const xv4<T> v = hw * o[k][c];
                if constexpr (MODE == 0) {
                    const xv4<T> pv = c == 0 ? p0 : (c == 1 ? p1 : p2);
                    if (st)
                        __builtin_nontemporal_store(live ? v : xv4<T>{T(0), T(0), T(0), T(0)},
                                                    (xv4<T>*)(q + e + es + c * blk));
#pragma unroll
                    for (int i = 0; i < 4; ++i) d[i] += (double)pv[i] * (double)v[i];
                } else {
                    const double dz = dinv ? dinv[(size_t)(mc + NX * (y + NY * zk)) * 3 + c] : 1.0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const double rv = b[e + c * blk + i] - (double)v[i];
                        if (live && mine[i]) rout[e + c * blk + i] = (R)rv;
                        d[i] += rv * (dz * rv);
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[i] += (double)mw * d[i];
        }
    }
    // per column: the 16 node lanes of a row (fixed butterfly), then the four segments in order
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double v = live ? acc[i] : 0.0;
        v += __shfl_xor(v, 8, 16);
        v += __shfl_xor(v, 4, 16);
        v += __shfl_xor(v, 2, 16);
        v += __shfl_xor(v, 1, 16);
        if (n == 0) sm[wv][4 * cq + i] = v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int j = threadIdx.x;
        partial[((size_t)tile * gridDim.x + g) * 64 + j] =
            (j >> 4) == sl ? ((sm[0][j & 15] + sm[1][j & 15]) + sm[2][j & 15]) + sm[3][j & 15] : 0.0;
    }
}

// Slot geometry of the vector kernels: thread t owns column quad t / 16 and nodes t % 16 + 16 s
// (s < ceil(B / 16)) of every (line, component) block its workgroup visits (g, g + G, ...).
struct XiSlot {
    int n, cq, B, S, blk;
};
__device__ __forceinline__ XiSlot xi_slot(int NX) {
    const int B = xi_nxb(NX);
    return XiSlot{(int)(threadIdx.x & 15), (int)(threadIdx.x >> 4), B, (B + 15) >> 4, 64 * B};
}

// Per-column partials of a vector kernel: the 16 node threads of a column quad (one 16-lane
// row of a wave) in a fixed butterfly, then partial[tile][g][64].
__device__ __forceinline__ void xi_partial(double (&acc)[4], double* __restrict__ partial, int tile, int g) {
    const int cq = threadIdx.x >> 4, m = threadIdx.x & 15;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double s = acc[i];
        s += __shfl_xor(s, 8, 16);
        s += __shfl_xor(s, 4, 16);
        s += __shfl_xor(s, 2, 16);
        s += __shfl_xor(s, 1, 16);
        if (m == 0) partial[((size_t)tile * gridDim.x + g) * 64 + 4 * cq + i] = s;
    }
}

// r -= alpha q and the weighted partial r.r (r.(dinv r) with a Jacobi dinv), FP32 vectors.
// Padding nodes (16-node blocks of shorter lines) take part with zeros, so every 256-byte
// segment is written whole.
template <bool PRE>
__global__ __launch_bounds__(kBlock) void k_xi_update(int NX, int NY, int mface, float* __restrict__ r,
                                                      const float* __restrict__ q, const ColState* __restrict__ cs,
                                                      const int* __restrict__ tile_on, double* __restrict__ partial,
                                                      const float* __restrict__ dinv) {
    using V = xv4<float>;
    int tile, g;
    tile_group(tile, g);
    const XiSlot sl = xi_slot(NX);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (tile_on[tile]) {
        float a[4];
        bool act[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const ColState& st = cs[(size_t)tile * 64 + 4 * sl.cq + j];
            act[j] = st.status == kActive;
            a[j] = (float)st.alpha;
        }
        const size_t off = xi_tile_elems(NX, NY) * tile + (size_t)sl.cq * 4 * sl.B;
        const int nbl = NY * NY * 3;
        const int G = (int)gridDim.x;
        // EQD_XI_UPD_U (line, component) blocks per step, all loaded before any is stored; a
        // thread whose four columns are all done leaves r alone
        const bool anya = act[0] || act[1] || act[2] || act[3];
        for (int bl0 = g; anya && bl0 < nbl; bl0 += EQD_XI_UPD_U * G) {
            for (int s = 0; s < sl.S; ++s) {
                const int m = sl.n + 16 * s;
                if (m >= sl.B) break;
                const bool pad = m >= NX;
                const double w = m == mface ? 2.0 : 1.0;
                V rr[EQD_XI_UPD_U], qq[EQD_XI_UPD_U];
#pragma unroll
                for (int u = 0; u < EQD_XI_UPD_U; ++u) {
                    const int bl = bl0 + u * G;
                    if (bl >= nbl) continue;
                    rr[u] = __builtin_nontemporal_load(reinterpret_cast<const V*>(r + off + (size_t)bl * sl.blk + 4 * m));
                    qq[u] = __builtin_nontemporal_load(reinterpret_cast<const V*>(q + off + (size_t)bl * sl.blk + 4 * m));
                }
#pragma unroll
                for (int u = 0; u < EQD_XI_UPD_U; ++u) {
                    const int bl = bl0 + u * G;
                    if (bl >= nbl) continue;
                    float dz = 1.0f;
                    if constexpr (PRE) {
                        if (!pad) dz = dinv[(size_t)(m + NX * (bl / 3)) * 3 + bl % 3];
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (act[j]) {
                            const float rn = pad ? 0.0f : rr[u][j] - a[j] * qq[u][j];
                            rr[u][j] = rn;
                            acc[j] += w * ((double)rn * (double)(dz * rn));
                        }
                    }
                    __builtin_nontemporal_store(rr[u], reinterpret_cast<V*>(r + off + (size_t)bl * sl.blk + 4 * m));
                }
            }
        }
    }
    xi_partial(acc, partial, tile, g);
}

// x_c += alpha p (columns updated this iteration) and p = r + beta p (dinv r + beta p with a
// Jacobi dinv) for the columns still active, FP32 vectors. Columns flagged for a reliable update
// in an active tile then fold x_c into the FP64 x and clear it (k_xi_xaccum's step, the same
// arithmetic, without another pass over x_c).
template <bool PRE>
__global__ __launch_bounds__(kBlock) void k_xi_pupdate(int NX, int NY, float* __restrict__ x, float* __restrict__ p,
                                                       const float* __restrict__ r, const ColState* __restrict__ cs,
                                                       const float* __restrict__ dinv, double* __restrict__ x64,
                                                       const int* __restrict__ tile_on) {
    using V = xv4<float>;
    using V64 = xv4<double>;
    int tile, g;
    tile_group(tile, g);
    const XiSlot sl = xi_slot(NX);
    float a[4], be[4];
    bool xu[4], pu[4], fo[4];
    bool anyx = false, anyp = false, anyf = false;
    const bool ton = x64 && tile_on[tile];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const ColState& st = cs[(size_t)tile * 64 + 4 * sl.cq + j];
        xu[j] = st.xpend != 0;
        pu[j] = st.status == kActive;
        fo[j] = ton && st.rel_flag;
        a[j] = (float)st.alpha;
        be[j] = (float)st.beta;
        anyx |= xu[j];
        anyp |= pu[j];
        anyf |= fo[j];
    }
    if (!anyx && !anyp && !anyf) return;
    const size_t off = xi_tile_elems(NX, NY) * tile + (size_t)sl.cq * 4 * sl.B;
    const int nbl = NY * NY * 3;
    for (int bl = g; bl < nbl; bl += (int)gridDim.x) {
        for (int s = 0; s < sl.S; ++s) {
            const int m = sl.n + 16 * s;
            if (m >= sl.B) break;
            const bool pad = m >= NX;
            const size_t e = off + (size_t)bl * sl.blk + 4 * m;
            V* pv = reinterpret_cast<V*>(p + e);
            V pp = __builtin_nontemporal_load(pv);
            if (anyx || anyf) {
                V* xv = reinterpret_cast<V*>(x + e);
                V xx = __builtin_nontemporal_load(xv);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (xu[j]) xx[j] = pad ? 0.0f : xx[j] + a[j] * pp[j];
                if (anyf) {
                    V64* wv = reinterpret_cast<V64*>(x64 + e);
                    V64 ww = *wv;
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (fo[j]) {
                            ww[j] += (double)xx[j];
                            xx[j] = 0.0f;
                        }
                    *wv = ww;
                }
                __builtin_nontemporal_store(xx, xv);
            }
            if (anyp) {
                const V rr = __builtin_nontemporal_load(reinterpret_cast<const V*>(r + e));
                float dz = 1.0f;
                if constexpr (PRE) {
                    if (!pad) dz = dinv[(size_t)(m + NX * (bl / 3)) * 3 + bl % 3];
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (pu[j]) pp[j] = pad ? 0.0f : (PRE ? dz * rr[j] : rr[j]) + be[j] * pp[j];
                __builtin_nontemporal_store(pp, pv);
            }
        }
    }
}

// Reliable update, part 1: x64 += xc, xc = 0 for the flagged columns (every column when all !=
// 0; exactly the masked columns with a mask). Padding nodes hold zeros in x and xc.
__global__ __launch_bounds__(kBlock) void k_xi_xaccum(int NX, int NY, double* __restrict__ x, float* __restrict__ xc,
                                                      const ColState* __restrict__ cs,
                                                      const int* __restrict__ tile_on, int all,
                                                      const int* __restrict__ mask) {
    int tile, g;
    tile_group(tile, g);
    const XiSlot sl = xi_slot(NX);
    bool f[4];
    bool any = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const size_t col = (size_t)tile * 64 + 4 * sl.cq + j;
        f[j] = mask ? mask[col] != 0 : (all || (tile_on[tile] && cs[col].rel_flag));
        any |= f[j];
    }
    if (!any) return;
    const size_t off = xi_tile_elems(NX, NY) * tile + (size_t)sl.cq * 4 * sl.B;
    const int nbl = NY * NY * 3;
    for (int bl = g; bl < nbl; bl += (int)gridDim.x) {
        for (int s = 0; s < sl.S; ++s) {
            const int m = sl.n + 16 * s;
            if (m >= sl.B) break;
            const size_t e = off + (size_t)bl * sl.blk + 4 * m;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (f[j]) {
                    x[e + j] += (double)xc[e + j];
                    xc[e + j] = 0.0f;
                }
            }
        }
    }
}

// Weighted partial dot of b with itself per column (b.(dinv b) with a Jacobi dinv), FP64.
__global__ __launch_bounds__(kBlock) void k_xi_selfdot(int NX, int NY, int mface, const double* __restrict__ v,
                                                       double* __restrict__ partial,
                                                       const double* __restrict__ dinv) {
    int tile, g;
    tile_group(tile, g);
    const XiSlot sl = xi_slot(NX);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const size_t off = xi_tile_elems(NX, NY) * tile + (size_t)sl.cq * 4 * sl.B;
    const int nbl = NY * NY * 3;
    for (int bl = g; bl < nbl; bl += (int)gridDim.x) {
        for (int s = 0; s < sl.S; ++s) {
            const int m = sl.n + 16 * s;
            if (m >= sl.B) break;
            if (m >= NX) continue;
            const double w = m == mface ? 2.0 : 1.0;
            const double dz = dinv ? dinv[(size_t)(m + NX * (bl / 3)) * 3 + bl % 3] : 1.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double vv = v[off + (size_t)bl * sl.blk + 4 * m + j];
                acc[j] += w * (vv * (dz * vv));
            }
        }
    }
    xi_partial(acc, partial, tile, g);
}

// b = H P (D+SAT) chi(s) and the Krylov start (x = 0, xc = 0, r = b, p = b or dinv b) for the
// masked (or all) columns, as k_rhs, in the x1-inner layout: thread slots of four columns.
__global__ __launch_bounds__(kBlock) void k_xi_rhs(RhsView rv, const double* __restrict__ fields, int nslots,
                                                   double* __restrict__ b, double* __restrict__ x,
                                                   float* __restrict__ r, float* __restrict__ p,
                                                   float* __restrict__ xc, const int* __restrict__ mask,
                                                   const double* __restrict__ dinv) {
    int tile, g;
    tile_group(tile, g);
    const int NX = rv.NX, NY = rv.NY, n23 = rv.n23;
    const XiSlot sl = xi_slot(NX);
    bool on[4];
    bool any = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = tile * 64 + 4 * sl.cq + j;
        on[j] = !mask || mask[col];
        any |= on[j];
    }
    if (!any) return;
    const size_t off = xi_tile_elems(NX, NY) * tile + (size_t)sl.cq * 4 * sl.B;
    const int nbl = NY * NY * 3;
    for (int bl = g; bl < nbl; bl += (int)gridDim.x) {
        const int c = bl % 3, yz = bl / 3;
        const int y = yz % NY, z = yz / NY, iy = y + 1, iz = z + 1;
        for (int s = 0; s < sl.S; ++s) {
            const int m = sl.n + 16 * s;
            if (m >= sl.B) break;
            const size_t e = off + (size_t)bl * sl.blk + 4 * m;
            double vals[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!on[j]) continue;
                const int col = tile * 64 + 4 * sl.cq + j;
                double val = 0.0;
                if (m < NX && col < nslots) {
                    const double* s2 = fields + (size_t)col * 2 * n23 * n23;
                    const double* s3 = s2 + (size_t)n23 * n23;
                    const double H = rv.Hx[m] * rv.Hy[y] * rv.Hy[z];
                    if (c > 0) {
                        const double jt = rv.Jt[m];
                        if (jt != 0.0) val = H * jt * (c == 1 ? s2 : s3)[iy + (size_t)n23 * iz];
                    } else {
                        const double jn = rv.Jn[m];
                        if (jn != 0.0) {
                            double d = 0.0;
                            const int sy = rv.D1full.start[y], wy = rv.D1full.width[y];
                            for (int t = 0; t < wy; ++t)
                                d += rv.D1full.c[(size_t)y * kTabW + t] * s2[(sy + t) + (size_t)n23 * iz];
                            const int sz = rv.D1full.start[z], wz = rv.D1full.width[z];
                            for (int t = 0; t < wz; ++t)
                                d += rv.D1full.c[(size_t)z * kTabW + t] * s3[iy + (size_t)n23 * (sz + t)];
                            val = H * jn * d;
                        }
                    }
                }
                vals[j] = val;
            }
            const double dz = (m < NX && dinv) ? dinv[(size_t)(m + NX * yz) * 3 + c] : 1.0;
            if (on[0] && on[1] && on[2] && on[3]) {
                // the four columns' entries are 32 (FP64) and 16 (FP32) contiguous bytes
                *reinterpret_cast<xv4<double>*>(b + e) = xv4<double>{vals[0], vals[1], vals[2], vals[3]};
                *reinterpret_cast<xv4<double>*>(x + e) = xv4<double>{0.0, 0.0, 0.0, 0.0};
                *reinterpret_cast<xv4<float>*>(r + e) =
                    xv4<float>{(float)vals[0], (float)vals[1], (float)vals[2], (float)vals[3]};
                *reinterpret_cast<xv4<float>*>(p + e) =
                    dinv ? xv4<float>{(float)(dz * vals[0]), (float)(dz * vals[1]), (float)(dz * vals[2]),
                                      (float)(dz * vals[3])}
                         : xv4<float>{(float)vals[0], (float)vals[1], (float)vals[2], (float)vals[3]};
                *reinterpret_cast<xv4<float>*>(xc + e) = xv4<float>{0.0f, 0.0f, 0.0f, 0.0f};
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!on[j]) continue;
                    b[e + j] = vals[j];
                    x[e + j] = 0.0;
                    r[e + j] = (float)vals[j];
                    p[e + j] = dinv ? (float)(dz * vals[j]) : (float)vals[j];
                    xc[e + j] = 0.0f;
                }
            }
        }
    }
}

}  // namespace eqd
