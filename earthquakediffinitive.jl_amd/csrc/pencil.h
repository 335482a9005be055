// Pencil stencil for tiles of 64 columns. Included by kernels.h.
//
// One wavefront = one x1 pencil segment of S nodes at fixed (x2, x3); lanes = the 64
// right-hand sides of a tile. Every coefficient is wave-uniform: the x2/x3 bands come
// through the scalar cache, the x1 bands from LDS (broadcast reads). The x1 operators
// are applied from a register window; the x2/x3 operators are uniform-coefficient line
// updates, so one loaded line serves every output component that needs it. The mixed
// x1-x2 and x1-x3 terms are sum-factorised: the x2 (x3) derivative is accumulated over
// the x1 window first and the x1 band is applied once.
//
// The tile's vector is read through one buffer resource, so neighbour-line loads need
// no per-load address registers. Register discipline: every component index is a template parameter (no dynamic
// indexing of register arrays), the loops over neighbour lines are not unrolled (one
// line's loads in flight per wave; occupancy supplies the rest), and scheduling fences
// keep one output row's coefficients from being hoisted into the next.
//
// The inner loops are branch-free: x1 positions outside [0, NX) read a clamped (valid)
// address and meet a zero band coefficient (the tables only couple interior columns),
// and the x1 bands are zero-padded to nseg*S rows so the unused outputs of the last
// segment stay in bounds.
#pragma once

#ifndef EQD_PENCIL_WAVES
#define EQD_PENCIL_WAVES 4
#endif
// q is streamed to HBM with non-temporal stores: the update kernel reads it once, and the
// caches are left to the neighbour lines of p.
#ifndef EQD_NT_Q
#define EQD_NT_Q 1
#endif
// x2 rows per block of the pencil order on large grids (see PencilView::by)
#ifndef EQD_PENCIL_BY
#define EQD_PENCIL_BY 8
#endif

namespace eqd {

// Pencil segment length: 16 nodes in FP32, 8 in FP64 (register budget).
template <typename T>
constexpr int pencil_s() { return sizeof(T) == 8 ? 8 : 16; }

#define EQD_FENCE() __builtin_amdgcn_sched_barrier(0)

// Unroll factor of the loops over neighbour lines (lines in flight per wave).
#ifndef EQD_LINE_UNROLL
#define EQD_LINE_UNROLL 1
#endif

// Buffer addressing: one resource per tile vector (base and size in scalar registers),
// a per-lane byte offset in one VGPR and the wave-uniform element offset in an SGPR,
// so a neighbour-line load costs no address registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t eqd_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// x1 band tables staged in LDS by k_pencil (AXs, AXo, MX for every padded row).
template <typename T>
struct XBands {
    const T* AXs;
    const T* AXo;
    const T* MX;
};

// F ("full line"): one segment covers the whole x1 line (NX <= S), so the x1 windows need
// no halo: columns outside [0, NX) only ever meet zero band coefficients.
template <typename T, int S, bool F>
struct PencilCtx {
    const PencilView<T>& op;
    const XBands<T>& xb;
    __amdgpu_buffer_rsrc_t rs;
    int voff;  // lane * sizeof(T)
    int m0, y, z;
    static constexpr int C = 3 * 64 * (int)sizeof(T);  // byte stride between consecutive x1 nodes
    // byte offset of node (0, yy, zz), component c
    __device__ int line(int yy, int zz, int c) const {
        return ((op.NX * (yy + op.NY * zz)) * 3 + c) * 64 * (int)sizeof(T);
    }
    __device__ int xc(int mm) const { return mm < 0 ? 0 : (mm >= op.NX ? op.NX - 1 : mm); }
    __device__ T ld(int lineoff, int mm) const { return bload<T>(rs, voff, lineoff + xc(mm) * C); }
};

// Wave-uniform read of a band coefficient through the constant address space, so it is
// a scalar load (counted on lgkmcnt): a vector load here would share vmcnt with the
// data loads and every wait for a coefficient would drain all loads in flight.
template <typename T>
__device__ __forceinline__ T ucoef(const T* p) {
    using CP = const __attribute__((address_space(4))) T*;
    return *(CP)p;
}

// x1 term of component Cc: AXs for component 0, AXo for 1 and 2.
template <typename T, int S, bool F, int Cc>
__device__ __forceinline__ void pen_x(const PencilCtx<T, S, F>& P, T (&o)[3][S]) {
    constexpr int RB = Cc == 0 ? kRxS : kRxO;
    constexpr int R = F ? 0 : kRxS, W = S + 2 * R;
    EQD_FENCE();
    const int L = P.line(P.y, P.z, Cc);
    T w[W];
#pragma unroll
    for (int j = 0; j < W; ++j) w[j] = P.ld(L, P.m0 - R + j);
    const T* band = Cc == 0 ? P.xb.AXs : P.xb.AXo;
    // The centre taps of the x2 and x3 terms act on this same line (the pencil's own
    // nodes): they are applied here instead of loading the line again.
    const T cy = Cc == 1 ? ucoef(P.op.Ts + (size_t)P.y * (2 * kRyS + 1) + kRyS)
                         : ucoef(P.op.To + (size_t)P.y * (2 * kRyO + 1) + kRyO);
    const T cz = Cc == 2 ? ucoef(P.op.Ts + (size_t)P.z * (2 * kRyS + 1) + kRyS)
                         : ucoef(P.op.To + (size_t)P.z * (2 * kRyO + 1) + kRyO);
    const T cyz = cy + cz;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        // A real (uniform) branch per row: keeps the compiler from hoisting every row's
        // coefficient reads to the top of the block (which needs hundreds of registers).
        if (P.m0 + i < P.op.NX) {
            const T* b = band + (size_t)(P.m0 + i) * (2 * RB + 1);
            T a = cyz * w[i + R];
#pragma unroll
            for (int d = -RB; d <= RB; ++d)
                if (i + R + d >= 0 && i + R + d < W) a += b[RB + d] * w[i + R + d];
            o[Cc][i] += a;
        }
    }
}

// Neighbour rows outside [0, NY) are clamped to a valid row: their band coefficients
// are zero (the tables only couple interior rows), so the loads stay branch-free.
__device__ __forceinline__ int rclamp(int r, int n) { return r < 0 ? 0 : (r >= n ? n - 1 : r); }

// Mixed x1-x2 (AX = 0) or x1-x3 (AX = 1) as a pass of its own: o[DST] += MX (D1 along AX of
// component SRC). Used when the pencil has x1 halos (the fused form would spill there).
template <typename T, int S, bool F, int AX, int SRC, int DST>
__device__ __forceinline__ void pen_mixed_x(const PencilCtx<T, S, F>& P, T (&o)[3][S]) {
    constexpr int R = F ? 0 : kRxO, W = S + 2 * R;
    EQD_FENCE();
    T a[W];
#pragma unroll
    for (int j = 0; j < W; ++j) a[j] = T(0);
    const int row = AX == 0 ? P.y : P.z;
    const T* bd = P.op.D1 + (size_t)row * (2 * kRyO + 1);
#pragma unroll EQD_LINE_UNROLL
    for (int d = -kRyO; d <= kRyO; ++d) {
        const T cd = ucoef(bd + kRyO + d);
        if (cd != T(0)) {  // uniform: interior D1 rows have a zero centre tap
            const int rr = rclamp(row + d, P.op.NY);
            const int L = AX == 0 ? P.line(rr, P.z, SRC) : P.line(P.y, rr, SRC);
            T v[W];
#pragma unroll
            for (int j = 0; j < W; ++j) v[j] = P.ld(L, P.m0 - R + j);
            EQD_FENCE();  // issue the whole line before the first use
#pragma unroll
            for (int j = 0; j < W; ++j) a[j] += cd * v[j];
        }
    }
#pragma unroll
    for (int i = 0; i < S; ++i) {
        if (P.m0 + i < P.op.NX) {  // per-row branch, as in pen_x
            const T* b = P.xb.MX + (size_t)(P.m0 + i) * (2 * kRxO + 1);
            T s = T(0);
#pragma unroll
            for (int d = -kRxO; d <= kRxO; ++d)
                if (i + R + d >= 0 && i + R + d < W) s += b[kRxO + d] * a[i + R + d];
            o[DST][i] += s;
        }
    }
}

// (lam+mu) D1x applied to an accumulated x1 window: o[DST][i] += sum_dx MX[m][dx] a[i + R + dx].
template <typename T, int S, bool F, int DST, int W>
__device__ __forceinline__ void apply_mx(const PencilCtx<T, S, F>& P, const T (&a)[W], T (&o)[3][S]) {
    constexpr int R = (W - S) / 2;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        if (P.m0 + i < P.op.NX) {  // per-row branch, as in pen_x
            const T* b = P.xb.MX + (size_t)(P.m0 + i) * (2 * kRxO + 1);
            T s = T(0);
#pragma unroll
            for (int d = -kRxO; d <= kRxO; ++d)
                if (i + R + d >= 0 && i + R + d < W) s += b[kRxO + d] * a[i + R + d];
            o[DST][i] += s;
        }
    }
}

// x2 (AX = 0) or x3 (AX = 1) terms: the second-derivative bands (Ts on the component along
// AX, To on the other two) and, from the same lines, the mixed x1-x2 / x1-x3 terms
// (lam+mu) D1x D1_AX: the D1_AX sum of a source component is accumulated over its lines
// and D1x is applied once at the end. Source components: the component along AX feeds
// o[0]; component 0 feeds o[AX == 0 ? 1 : 2]. The centre taps (d = 0) are in pen_x.
// Lines that feed a mixed term are loaded with the x1 halo of D1x when the pencil has one.
template <typename T, int S, bool F, int AX>
__device__ __forceinline__ void pen_yz(const PencilCtx<T, S, F>& P, T (&o)[3][S]) {
    constexpr int SC = AX == 0 ? 1 : 2, O1 = 0, O2 = AX == 0 ? 2 : 1;
    constexpr int DA = 0, DB = AX == 0 ? 1 : 2;
    constexpr int RM = F ? 0 : kRxO, WM = S + 2 * RM;  // x1 window of the mixed term
    EQD_FENCE();
    const int row = AX == 0 ? P.y : P.z;
    const T* bs = P.op.Ts + (size_t)row * (2 * kRyS + 1);
    const T* bo = P.op.To + (size_t)row * (2 * kRyO + 1);
    const T* bd = P.op.D1 + (size_t)row * (2 * kRyO + 1);
    auto lines = [&](int rr, int c) { return AX == 0 ? P.line(rr, P.z, c) : P.line(P.y, rr, c); };
    T a[WM];
    // component SC: Ts, and D1 into the mixed accumulator for o[DA]
#pragma unroll
    for (int j = 0; j < WM; ++j) a[j] = T(0);
#pragma unroll EQD_LINE_UNROLL
    for (int d = -kRyS; d <= kRyS; ++d) {
        if (d == 0) continue;
        const T cs = ucoef(bs + kRyS + d);
        const T cd = (d >= -kRyO && d <= kRyO) ? ucoef(bd + kRyO + d) : T(0);
        const int L = lines(rclamp(row + d, P.op.NY), SC);
        if (cd != T(0)) {  // uniform branches: skip what vanishes on this row
            T v[WM];
#pragma unroll
            for (int j = 0; j < WM; ++j) v[j] = P.ld(L, P.m0 - RM + j);
            EQD_FENCE();
#pragma unroll
            for (int i = 0; i < S; ++i) o[SC][i] += cs * v[i + RM];
#pragma unroll
            for (int j = 0; j < WM; ++j) a[j] += cd * v[j];
        } else if (cs != T(0)) {
            T v[S];
#pragma unroll
            for (int i = 0; i < S; ++i) v[i] = P.ld(L, P.m0 + i);
            EQD_FENCE();
#pragma unroll
            for (int i = 0; i < S; ++i) o[SC][i] += cs * v[i];
        }
    }
    apply_mx<T, S, F, DA, WM>(P, a, o);
    EQD_FENCE();
    // components O1 (= 0) and O2: To, and D1 of component 0 into the accumulator for o[DB]
#pragma unroll
    for (int j = 0; j < WM; ++j) a[j] = T(0);
#pragma unroll EQD_LINE_UNROLL
    for (int d = -kRyO; d <= kRyO; ++d) {
        if (d == 0) continue;
        const T co = ucoef(bo + kRyO + d);
        const T cd = ucoef(bd + kRyO + d);
        if (co != T(0) || cd != T(0)) {
            const int rr = rclamp(row + d, P.op.NY);
            const int La = lines(rr, O1), Lb = lines(rr, O2);
            T va[WM], vb[S];
#pragma unroll
            for (int j = 0; j < WM; ++j) va[j] = P.ld(La, P.m0 - RM + j);
#pragma unroll
            for (int i = 0; i < S; ++i) vb[i] = P.ld(Lb, P.m0 + i);
            EQD_FENCE();
#pragma unroll
            for (int i = 0; i < S; ++i) {
                o[O1][i] += co * va[i + RM];
                o[O2][i] += co * vb[i];
            }
#pragma unroll
            for (int j = 0; j < WM; ++j) a[j] += cd * va[j];
        }
    }
    apply_mx<T, S, F, DB, WM>(P, a, o);
}

// Mixed x2-x3: (lam+mu) D1y D1z, coupling components 1 and 2.
template <typename T, int S, bool F>
__device__ __forceinline__ void pen_mixed_yz(const PencilCtx<T, S, F>& P, T (&o)[3][S]) {
    EQD_FENCE();
    const T* by = P.op.D1s + (size_t)P.y * (2 * kRyO + 1);
    const T* bz = P.op.D1 + (size_t)P.z * (2 * kRyO + 1);
#pragma unroll EQD_LINE_UNROLL
    for (int t = 0; t < (2 * kRyO + 1) * (2 * kRyO + 1); ++t) {
        const int dy = t / (2 * kRyO + 1) - kRyO, dz = t % (2 * kRyO + 1) - kRyO;
        const T cc = ucoef(by + kRyO + dy) * ucoef(bz + kRyO + dz);
        if (cc != T(0)) {  // 9 of the 25 combinations vanish on interior rows
            const int yy = rclamp(P.y + dy, P.op.NY), zz = rclamp(P.z + dz, P.op.NY);
            const int L1 = P.line(yy, zz, 1), L2 = P.line(yy, zz, 2);
            T v1[S], v2[S];
#pragma unroll
            for (int i = 0; i < S; ++i) {
                v2[i] = P.ld(L2, P.m0 + i);
                v1[i] = P.ld(L1, P.m0 + i);
            }
            EQD_FENCE();
#pragma unroll
            for (int i = 0; i < S; ++i) {
                o[1][i] += cc * v2[i];
                o[2][i] += cc * v1[i];
            }
        }
    }
}

template <typename T, int S, bool F>
__device__ __forceinline__ void pencil_apply(const PencilView<T>& op, const XBands<T>& xb, __amdgpu_buffer_rsrc_t rs,
                                             int m0, int y, int z, int lane, T (&o)[3][S]) {
    const PencilCtx<T, S, F> P{op, xb, rs, lane * (int)sizeof(T), m0, y, z};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < S; ++i) o[c][i] = T(0);
    pen_x<T, S, F, 0>(P, o);
    pen_x<T, S, F, 1>(P, o);
    pen_x<T, S, F, 2>(P, o);
    pen_yz<T, S, F, 0>(P, o);  // + out0 += (lam+mu) D1x D1y u1, out1 += (lam+mu) D1x D1y u0
    pen_yz<T, S, F, 1>(P, o);  // + out0 += (lam+mu) D1x D1z u2, out2 += (lam+mu) D1x D1z u0
    pen_mixed_yz<T, S, F>(P, o);
}

// MODE 0: q = A p with the weighted partial p.q.
// MODE 1: reliable-update residual r = b - A x (FP64 in, r stored as R for flagged
//         lanes) with the weighted partial r.r of the FP64 residual (r.(dinv r) when a
//         Jacobi dinv is given).
// Waves per SIMD a pencil instantiation is built for: full-line FP32 pencils keep 4, pencils
// with x1 halos 3, and an FP64 pencil wider than 8 nodes 2 (its accumulators alone take 90 VGPRs).
template <typename T, bool F, int SW>
constexpr int pencil_waves() {
    return (sizeof(T) == 8 && SW > 8) ? 2 : (F ? EQD_PENCIL_WAVES : EQD_PENCIL_WAVES - 1);
}

// SW: pencil width (default: pencil_s<T>()); a full-line pencil of an x1 line shorter than that
// can be instantiated narrower so that no load falls past the line's end, and an FP64 full-line
// pencil wider than pencil_s<double>() replaces two 8-node segments with halos.
template <typename T, typename R, int MODE, bool F, int SW = pencil_s<T>()>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(pencil_waves<T, F, SW>()))) void k_pencil(
    PencilView<T> op, int NN, const T* __restrict__ in, T* __restrict__ q, const T* __restrict__ b,
    R* __restrict__ rout, const ColState* __restrict__ cs, const int* __restrict__ tile_on,
    double* __restrict__ partial, const double* __restrict__ dinv) {
    constexpr int S = SW;
    int tile, g;
    tile_group(tile, g);
    // A block of an inactive tile leaves before staging anything: k_finalize reads the
    // partials of active columns only, and every column of an inactive tile is inactive.
    if (!tile_on[tile]) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    extern __shared__ unsigned char eqd_lds[];
    T* lds = reinterpret_cast<T*>(eqd_lds);
    const int rows = op.nseg * S;
    const int nS = rows * (2 * kRxS + 1), nO = rows * (2 * kRxO + 1);
    for (int i = threadIdx.x; i < nS; i += kBlock) lds[i] = op.AXs[i];
    for (int i = threadIdx.x; i < nO; i += kBlock) {
        lds[nS + i] = op.AXo[i];
        lds[nS + nO + i] = op.MX[i];
    }
    __syncthreads();
    const XBands<T> xb{lds, lds + nS, lds + nS + nO};
    double acc = 0.0;
    {
        const bool mine = MODE == 0 || cs[(size_t)tile * 64 + lane].rel_flag != 0;
        const size_t off = (size_t)tile * NN * 3 * 64;
        const __amdgpu_buffer_rsrc_t rs = eqd_rsrc(in + off, (uint32_t)((size_t)NN * 3 * 64 * sizeof(T)));
        const int np = op.NY * op.NY * op.nseg;
        for (int pi = g * 4 + wv; pi < np; pi += (int)gridDim.x * 4) {
            const int pidx = uni(pi);
            // pencil order: x2 blocks of op.by rows one after another; within a block the
            // x1 segment fastest, then x2, then x3 (op.by = NY: whole x2 rows)
            const int full = op.by * op.NY * op.nseg;
            const int yb = pidx / full;
            int rem = pidx - yb * full;
            const int by = min(op.by, op.NY - yb * op.by);
            const int seg = rem % op.nseg;
            rem /= op.nseg;
            const int y = yb * op.by + rem % by, z = rem / by;
            const int m0 = seg * S;
            T o[3][S];
            pencil_apply<T, S, F>(op, xb, rs, m0, y, z, lane, o);
            EQD_FENCE();
            const T hyz = op.Hy[y] * op.Hy[z];
#pragma unroll
            for (int i = 0; i < S; ++i) {
                const int m = m0 + i;
                if (m < op.NX) {
                    const T hw = -(op.Hx[m] * hyz);
                    const size_t e = off + ((size_t)(m + op.NX * (y + op.NY * z)) * 3) * 64 + lane;
                    double d = 0.0;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const T v = hw * o[c][i];
                        if (MODE == 0) {
#if EQD_NT_Q
                            __builtin_nontemporal_store(v, q + e + c * 64);
#else
                            q[e + c * 64] = v;
#endif
                            d += (double)in[e + c * 64] * (double)v;
                        } else {
                            const double rv = (double)b[e + c * 64] - (double)v;
                            if (mine) rout[e + c * 64] = (R)rv;
                            if (dinv) d += rv * (dinv[(size_t)(m + op.NX * (y + op.NY * z)) * 3 + c] * rv);
                            else d += rv * rv;
                        }
                    }
                    acc += (m == op.mface ? 2.0 : 1.0) * d;
                }
            }
        }
    }
    block_partial<64>(acc, partial, tile, g);
}

// Dynamic LDS bytes k_pencil needs for the x1 bands of a view (pencil width S).
template <typename T, int S = pencil_s<T>()>
inline size_t pencil_lds_bytes(const PencilView<T>& op) {
    const size_t rows = (size_t)op.nseg * S;
    return rows * ((2 * kRxS + 1) + 2 * (2 * kRxO + 1)) * sizeof(T);
}

}  // namespace eqd
