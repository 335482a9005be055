// libeqd host orchestration and C ABI (see include/eqd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eqd.h"

// FP32 pencils per workgroup that set the groups per tile of a grid (see eqd_operator_create)
#ifndef EQD_PENCILS_PER_GROUP
#define EQD_PENCILS_PER_GROUP 4
#endif
#include "eqd_internal.h"
#include "kernels.h"

using namespace eqd;

#include "host_common.h"

using namespace eqd::host;

namespace {

void validate_grid(const eqd_grid_desc& g) {
    if (g.n1 < 9 || g.n23 < 9)
        throw Error(EQD_ERR_GRID, "SBP order 4 needs at least 9 points per dimension (n1 = " + std::to_string(g.n1) +
                                      ", n23 = " + std::to_string(g.n23) + ")");
    if (!(g.L_normal > 0) || !(g.L_fault > 0) || !(g.mu > 0)) throw Error(EQD_ERR_ARG, "L_normal, L_fault, mu must be > 0");
    // The tile kernels address a 64-column tile vector through a buffer resource with 32-bit
    // byte offsets: the largest one (FP64, 3 components per stored node) must stay below 2^31 B.
    const int64_t NN = (int64_t)(2 * g.n1 - 3) * (g.n23 - 2) * (g.n23 - 2);
    const int64_t tile_bytes = NN * 3 * 64 * (int64_t)sizeof(double);
    if (tile_bytes > (int64_t)INT32_MAX)
        throw Error(EQD_ERR_GRID, "grid too large: " + std::to_string(NN) + " stored nodes make a 64-column FP64 tile of " +
                                      std::to_string(tile_bytes) + " B, above the 2^31 B a tile vector may span (at most " +
                                      std::to_string((int64_t)INT32_MAX / (3 * 64 * 8)) + " stored nodes)");
}

// Omega_f from l_f as src/FaultResponse.jl:114-126 finds it.
void omega_range(const eqd_grid_desc& g, int& lo, int& cnt) {
    const double h2 = 2.0 * g.L_fault / (g.n23 - 1);
    const double tol = 1e-9 * std::max(g.l_f, 1.0);
    lo = -1;
    cnt = 0;
    for (int i = 0; i < g.n23; ++i) {
        double x = -g.L_fault + i * h2;
        if (std::fabs(x) <= g.l_f + tol) {
            if (lo < 0) lo = i;
            ++cnt;
        }
    }
    if (cnt == 0) throw Error(EQD_ERR_GRID, "no fault nodes inside Omega_f");
    double xlo = -g.L_fault + lo * h2, xhi = -g.L_fault + (lo + cnt - 1) * h2;
    if (!(std::fabs(xlo + g.l_f) <= tol && std::fabs(xhi - g.l_f) <= tol))
        throw Error(EQD_ERR_GRID, "the elastic grid has no nodes exactly on x = +-l_f; pick a spacing that divides l_f");
}

}  // namespace

struct eqd_op {
    eqd_grid_desc g{};
    HostTables tb;
    int NN = 0, nf = 0, n2f = 0, i2lo = 0;
    int64_t ntot = 0, N = 0;
    std::vector<std::unique_ptr<DevMem>> mem;
    OpView<double> v64{};
    OpView<float> v32{};
    RhsView rv{};
    ExtractView ev{};
    std::atomic<int> refs{1};

    template <typename V>
    V* up(const std::vector<V>& h) {
        mem.emplace_back(new DevMem(h.size() * sizeof(V)));
        HIPCHK(hipMemcpy(mem.back()->p, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
        return mem.back()->as<V>();
    }
    template <typename V>
    Tab<V> tab(const HostTab& h) {
        std::vector<V> c(h.c.begin(), h.c.end());
        return Tab<V>{up(h.start), up(h.width), up(c)};
    }
    template <typename V>
    OpView<V> view() {
        std::vector<V> hx(tb.Hx.begin(), tb.Hx.end()), hy(tb.Hy.begin(), tb.Hy.end());
        return OpView<V>{tb.NX, tb.NY, tb.mface, tab<V>(tb.AXs), tab<V>(tb.AXo), tab<V>(tb.MX), tab<V>(tb.Ts),
                         tab<V>(tb.To), tab<V>(tb.D1), tab<V>(tb.D1s), up(hx), up(hy)};
    }
    template <typename V>
    PencilView<V> pview() {
        constexpr int S = pencil_s<V>();
        const int nseg = (tb.NX + S - 1) / S;
        auto band = [&](const HostTab& t, int R) {  // x1 bands padded with zero rows to nseg*S
            std::vector<double> b = build_band(t, R);
            if (t.nrows == tb.NX) b.resize((size_t)nseg * S * (2 * R + 1), 0.0);
            return up(std::vector<V>(b.begin(), b.end()));
        };
        std::vector<V> hx(tb.Hx.begin(), tb.Hx.end()), hy(tb.Hy.begin(), tb.Hy.end());
        // Pencil order: a grid whose x3 planes are larger than an XCD's L2 (NY > 48) walks x2
        // blocks of EQD_PENCIL_BY rows, so the pencils in flight on an XCD share their x2 and
        // x3 neighbour lines (dz = 25 m: stencil 2.03 -> 1.47 ms per tile); smaller grids walk
        // whole x2 rows.
        const int by = tb.NY > 48 ? EQD_PENCIL_BY : tb.NY;
        return PencilView<V>{tb.NX, tb.NY, tb.mface, nseg, by, band(tb.AXs, kRxS),
                             band(tb.AXo, kRxO), band(tb.MX, kRxO), band(tb.Ts, kRyS), band(tb.To, kRyO),
                             band(tb.D1, kRyO), band(tb.D1s, kRyO), up(hx), up(hy)};
    }
    PencilView<double> p64{};
    PencilView<float> p32{};
    // x1 lines of 15 nodes (n1 = 9): the FP64 view as one 15-wide segment instead of two 8-node
    // segments with x1 halos (p64's bands are padded to 16 rows, the pencil reads 15)
    PencilView<double> p64w{};
    bool p64w_ok = false;
    // x1 lines of 9 to 64 nodes: the 64-column tiles of the mixed-precision engine use the
    // x1-inner layout and its kernels (xinner.h)
    bool xi = false;
    int G = kGroups;  // workgroups per tile of the node-parallel kernels (set at creation)
    int nfold = 0;    // > 0: block totals of the two-level partial sum (G > kFoldAbove)
    int fchunk = 0;   // groups per first-level block
    int64_t lin(int i1, int i2, int i3) const { return i1 + (int64_t)g.n1 * (i2 + (int64_t)g.n23 * i3); }

    // Jacobi M^{-1} on the merged unknowns (node-major, component fastest), on the host and
    // in HBM in both precisions; built by the first solver that asks for :jacobi.
    std::mutex jac_mu;
    std::vector<double> dinv_h;
    const double* dinv64 = nullptr;
    const float* dinv32 = nullptr;
    void ensure_jacobi() {
        std::lock_guard<std::mutex> lk(jac_mu);
        if (dinv64) return;
        std::vector<double> d = merged_diagonal(tb);
        for (double& v : d) {
            // every stored unknown is off the far field, so its diagonal is positive; a
            // non-positive entry means A is not SPD (src/ElasticitySplitNode.jl:352-357)
            if (!(v > 0)) throw Error(EQD_ERR_ARG, "diag(A) has non-positive entries; A is not PSD");
            v = 1.0 / v;
        }
        std::vector<float> f(d.begin(), d.end());
        const float* d32 = up(f);
        dinv64 = up(d);
        dinv32 = d32;
        dinv_h.swap(d);
    }
    template <typename V>
    const V* dinv() const {
        if constexpr (std::is_same<V, float>::value) return dinv32; else return dinv64;
    }
};

namespace {

// Map between the reference's full-space vector and the merged interior grid.
// merged (m, y, z, c): plane M = m + 1; M < n1-1 on g-, M > n1-1 on g+, M == n1-1 the face.
void reduce_full(const eqd_op& op, const double* full, double* red /* NN*3, node-major comp */, bool check) {
    const int n1 = op.g.n1, NX = op.tb.NX, NY = op.tb.NY;
    const int64_t N = op.N;
    double scale = 0.0;
    if (check)
        for (int64_t i = 0; i < op.ntot; ++i) scale = std::max(scale, std::fabs(full[i]));
    for (int z = 0; z < NY; ++z)
        for (int y = 0; y < NY; ++y)
            for (int m = 0; m < NX; ++m) {
                const int M = m + 1;
                const int64_t node = m + (int64_t)NX * (y + (int64_t)NY * z);
                for (int c = 0; c < 3; ++c) {
                    double v;
                    if (M < n1 - 1) v = full[c * N + op.lin(M, y + 1, z + 1)];
                    else if (M > n1 - 1) v = full[3 * N + c * N + op.lin(M - (n1 - 1), y + 1, z + 1)];
                    else {
                        double a = full[c * N + op.lin(n1 - 1, y + 1, z + 1)];
                        double b = full[3 * N + c * N + op.lin(0, y + 1, z + 1)];
                        if (check && std::fabs(a - b) > 1e-12 * scale)
                            throw Error(EQD_ERR_RANGE, "rhs is not in range(P): fault pair values differ");
                        v = 0.5 * (a + b);
                    }
                    red[node * 3 + c] = v;
                }
            }
    if (check) {  // far-field entries must vanish
        const int n2 = op.g.n23;
        for (int side = 0; side < 2; ++side)
            for (int i3 = 0; i3 < n2; ++i3)
                for (int i2 = 0; i2 < n2; ++i2)
                    for (int i1 = 0; i1 < n1; ++i1) {
                        bool ff = i2 == 0 || i2 == n2 - 1 || i3 == 0 || i3 == n2 - 1 || (side == 0 ? i1 == 0 : i1 == n1 - 1);
                        if (!ff) continue;
                        for (int c = 0; c < 3; ++c)
                            if (std::fabs(full[side * 3 * N + c * N + op.lin(i1, i2, i3)]) > 1e-12 * scale)
                                throw Error(EQD_ERR_RANGE, "rhs is not in range(P): far-field entries are nonzero");
                    }
    }
}

void embed_full(const eqd_op& op, const double* red, double* full) {
    const int n1 = op.g.n1, NX = op.tb.NX, NY = op.tb.NY;
    const int64_t N = op.N;
    std::fill(full, full + op.ntot, 0.0);
    for (int z = 0; z < NY; ++z)
        for (int y = 0; y < NY; ++y)
            for (int m = 0; m < NX; ++m) {
                const int M = m + 1;
                const int64_t node = m + (int64_t)NX * (y + (int64_t)NY * z);
                for (int c = 0; c < 3; ++c) {
                    const double v = red[node * 3 + c];
                    if (M <= n1 - 1) full[c * N + op.lin(M, y + 1, z + 1)] = v;
                    if (M >= n1 - 1) full[3 * N + c * N + op.lin(M - (n1 - 1), y + 1, z + 1)] = v;
                }
            }
}

// Slip vectors on Omega_f -> n23 x n23 face fields (two directions).
void slip_fields(const eqd_op& op, const double* s2, const double* s3, double* f) {
    const int n23 = op.g.n23;
    std::fill(f, f + 2 * (size_t)n23 * n23, 0.0);
    for (int k = 0; k < op.nf; ++k) {
        const int iy = op.i2lo + k % op.n2f, iz = op.i2lo + k / op.n2f;
        f[iy + (size_t)n23 * iz] = s2[k];
        f[(size_t)n23 * n23 + iy + (size_t)n23 * iz] = s3[k];
    }
}

}  // namespace

// Per-solver device workspace for one batch shape.
struct Workspace {
    int K = 0, ntiles = 0, prec = -1;
    int used_tiles = 0;  // > 0: the current batch fills only tiles [0, used_tiles) (set by run_columns)
    int live_tiles() const { return used_tiles > 0 ? used_tiles : ntiles; }
    size_t nvec = 0;
    DevMem b, x, r, p, q, xc, fields, partial, folded, cs, tile_on, tile_rel, colid, outpos;
    int* h_tile_on = nullptr;
    hipGraphExec_t exec = nullptr;
    int exec_iters = 0;
    ~Workspace() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (h_tile_on) (void)hipHostFree(h_tile_on);
    }
};

// Per-kernel event timing for measurement runs (kernel ids as eqd_solver_kernel_timing).
struct KTimer {
    static constexpr int kKinds = 6;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, size_t>> pending;  // kernel id, index of its begin event
    size_t used = 0;
    int cur_active = 0;
    int cur_rel = 0;  // tiles the FP64 residual pass (kernel 5) covered in the current iteration
    double ms[kKinds] = {};
    int64_t launches[kKinds] = {}, active[kKinds] = {};
    hipEvent_t next() {
        if (used == pool.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[used++];
    }
    void begin(int id, hipStream_t st) {
        pending.push_back({id, used});
        HIPCHK(hipEventRecord(next(), st));
    }
    void end(int, hipStream_t st) { HIPCHK(hipEventRecord(next(), st)); }
    // after a stream sync: fold the pending pairs into the totals; returns their sum (ms)
    double flush() {
        double tot = 0;
        for (auto& pr : pending) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, pool[pr.second], pool[pr.second + 1]));
            ms[pr.first] += t;
            launches[pr.first] += 1;
            active[pr.first] += pr.first == 5 ? cur_rel : cur_active;
            tot += t;
        }
        pending.clear();
        used = 0;
        return tot;
    }
    void reset() {
        for (int i = 0; i < kKinds; ++i) ms[i] = 0, launches[i] = active[i] = 0;
    }
    ~KTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

struct eqd_solver {
    std::unique_ptr<KTimer> kt{new KTimer()};
    eqd_op* op = nullptr;
    eqd_cg_opts o{};
    eqd_report st{};
    hipStream_t stream = nullptr;
    std::unique_ptr<Workspace> ws;
    double last_ms = 0;
    int64_t last_launches = 0, last_tile_iters = 0;
    std::vector<int> last_iters;  // per column of the last batch (slot order)
    int tuned_k[2] = {0, 0};      // autotuned tile width per precision (0 = not yet tuned)
    int last_k = 0;               // tile width of the last build
};

void eqd::host::solver_fault_dims(const eqd_solver* s, int& nf, int& n2f, int& n3f) {
    nf = s->op->nf;
    n2f = n3f = s->op->n2f;
}

namespace {

size_t bytes_per_column(const eqd_op& op, int prec) {
    const size_t n = (prec == EQD_PREC_MIXED && op.xi) ? xi_tile_elems(op.tb.NX, op.tb.NY) / 64 : (size_t)op.NN * 3;
    const size_t t = prec == EQD_PREC_MIXED ? 4 : 8;
    return n * (8 + 8 + 3 * t + (prec == EQD_PREC_MIXED ? 4 : 0)) + 2 * (size_t)op.g.n23 * op.g.n23 * 8;
}

void ensure_ws(eqd_solver* s, int K, int ntiles, int prec) {
    Workspace* w = s->ws.get();
    if (w && w->K == K && w->ntiles == ntiles && w->prec == prec) return;
    s->ws.reset();
    s->ws.reset(new Workspace());
    w = s->ws.get();
    const eqd_op& op = *s->op;
    w->K = K;
    w->ntiles = ntiles;
    w->prec = prec;
    w->nvec = (K == 64 && prec == EQD_PREC_MIXED && op.xi) ? (size_t)ntiles * xi_tile_elems(op.tb.NX, op.tb.NY)
                                                           : (size_t)ntiles * op.NN * 3 * K;
    const size_t slots = (size_t)ntiles * K;
    const size_t t = prec == EQD_PREC_MIXED ? sizeof(float) : sizeof(double);
    w->b.alloc(w->nvec * 8);
    w->x.alloc(w->nvec * 8);
    w->r.alloc(w->nvec * t);
    w->p.alloc(w->nvec * t);
    w->q.alloc(w->nvec * t);
    if (prec == EQD_PREC_MIXED) w->xc.alloc(w->nvec * sizeof(float));
    w->fields.alloc(slots * 2 * op.g.n23 * op.g.n23 * 8);
    w->partial.alloc((size_t)ntiles * op.G * K * 8);
    if (op.nfold) w->folded.alloc((size_t)ntiles * op.nfold * K * 8);
    w->cs.alloc(slots * sizeof(ColState));
    w->tile_on.alloc(ntiles * sizeof(int));
    w->tile_rel.alloc(ntiles * sizeof(int));
    w->colid.alloc(slots * 8);
    w->outpos.alloc(slots * 8);
    HIPCHK(hipHostMalloc((void**)&w->h_tile_on, ntiles * sizeof(int), hipHostMallocDefault));
}

template <typename T, int K>
struct Engine {
    // The workspace holds 64-column mixed-precision tiles of an x1-inner grid (xinner.h). Depends
    // on the workspace, not on T: extraction runs as Engine<double, K> for both precisions.
    static bool xi(const eqd_solver* s) {
        return K == 64 && s->op->xi && s->ws && s->ws->prec == EQD_PREC_MIXED;
    }

    // x += xc, xc = 0 for every column (mask == nullptr) or exactly the masked ones.
    static void xaccum_all(eqd_solver* s, hipStream_t st, dim3 grid, const int* mask) {
        Workspace& w = *s->ws;
        if (xi(s))
            hipLaunchKernelGGL(k_xi_xaccum, grid, dim3(kBlock), 0, st, s->op->tb.NX, s->op->tb.NY, w.x.as<double>(),
                               w.xc.as<float>(), w.cs.as<ColState>(), w.tile_on.as<int>(), 1, mask);
        else
            hipLaunchKernelGGL((k_xaccum<K>), grid, dim3(kBlock), 0, st, s->op->NN, w.x.as<double>(),
                               w.xc.as<float>(), w.cs.as<ColState>(), w.tile_on.as<int>(), 1, mask);
    }

    // One CG iteration for every tile in flight (device-side alpha/beta). With a
    // timer, every launch is bracketed by events (measurement runs, no graph).
    static void iteration(eqd_solver* s, hipStream_t st, KTimer* kt = nullptr) {
        Workspace& w = *s->ws;
        const eqd_op& op = *s->op;
        const bool mixed = std::is_same<T, float>::value;
        const int R = s->o.reliable_interval > 0 ? s->o.reliable_interval : 48;
        dim3 grid(s->op->G, w.ntiles);
        T* r = w.r.as<T>();
        T* p = w.p.as<T>();
        T* q = w.q.as<T>();
        ColState* cs = w.cs.as<ColState>();
        int* on = w.tile_on.as<int>();
        int* rel = w.tile_rel.as<int>();
        double* part = w.partial.as<double>();
        // the partials k_finalize reads: the groups' own, or their first-level block totals
        const double* fpart = op.nfold ? w.folded.as<double>() : part;
        const int fG = op.nfold ? op.nfold : op.G;
        auto fold = [&] {
            if (op.nfold)
                hipLaunchKernelGGL((k_fold_partials<K>), dim3(op.nfold, w.ntiles), dim3(K * 16), 0, st, op.G,
                                   op.fchunk, (const double*)part, w.folded.as<double>());
        };
        OpView<T> v;
        if constexpr (std::is_same<T, float>::value) v = op.v32; else v = op.v64;
        const bool jac = s->o.precond == EQD_PRECOND_JACOBI;
        const T* dinv = jac ? op.dinv<T>() : nullptr;
        const double* dinv64 = jac ? op.dinv64 : nullptr;
        auto timed = [&](int id, auto&& launch) {
            if (kt) kt->begin(id, st);
            launch();
            if (kt) kt->end(id, st);
        };
        timed(0, [&] {
            if constexpr (K == 64) {
                PencilView<T> pv, pvw{};
                bool wide = false;
                if constexpr (std::is_same<T, float>::value) pv = op.p32; else {
                    pv = op.p64;
                    pvw = op.p64w;
                    wide = op.p64w_ok;
                }
                // x1 lines of at most 15 nodes (n1 = 9: the c1, c2 grids) take a 15-wide pencil, so
                // that no neighbour-line load falls past the line's end (stencil 559 -> 535 us)
                if constexpr (std::is_same<T, float>::value) {
                    if (xi(s)) {
                        if (op.tb.NX <= 16)
                            hipLaunchKernelGGL((k_xi_pencil<float, float, 0, EQD_XI_Z>), grid, dim3(64 * EQD_XI_BW), 0, st,
                                               op.p32, p, q, (const double*)nullptr, (float*)nullptr, cs, on, part,
                                               (const double*)nullptr);
                        else
                            hipLaunchKernelGGL((k_xl_pencil<float, float, 0, EQD_XI_Z>), grid, dim3(kBlock), 0, st,
                                               op.p32, p, q, (const double*)nullptr, (float*)nullptr, cs, on, part,
                                               (const double*)nullptr);
                        return;
                    }
                }
                if (pv.nseg == 1 && std::is_same<T, float>::value && pv.NX <= 15)
                    hipLaunchKernelGGL((k_pencil<T, T, 0, true, 15>), grid, dim3(kBlock), pencil_lds_bytes(pv), st,
                                       pv, op.NN, p, q, (const T*)nullptr, (T*)nullptr, cs, on, part,
                                       (const double*)nullptr);
                else if (wide)  // FP64 on 15-node lines: one segment instead of two with x1 halos
                    hipLaunchKernelGGL((k_pencil<T, T, 0, true, 15>), grid, dim3(kBlock),
                                       (pencil_lds_bytes<T, 15>(pvw)), st, pvw, op.NN, p, q, (const T*)nullptr,
                                       (T*)nullptr, cs, on, part, (const double*)nullptr);
                else if (pv.nseg == 1)
                    hipLaunchKernelGGL((k_pencil<T, T, 0, true>), grid, dim3(kBlock), pencil_lds_bytes(pv), st, pv,
                                       op.NN, p, q, (const T*)nullptr, (T*)nullptr, cs, on, part,
                                       (const double*)nullptr);
                else
                    hipLaunchKernelGGL((k_pencil<T, T, 0, false>), grid, dim3(kBlock), pencil_lds_bytes(pv), st, pv,
                                       op.NN, p, q, (const T*)nullptr, (T*)nullptr, cs, on, part,
                                       (const double*)nullptr);
            } else {
                hipLaunchKernelGGL((k_apply_dot<T, K>), grid, dim3(kBlock), 0, st, v, op.NN, p, q, on, part);
            }
        });
        timed(3, [&] {
            fold();
            hipLaunchKernelGGL((k_finalize<K>), dim3(w.ntiles), dim3(K * kFinStrands), 0, st, fG, 0, (int)mixed, R, cs, fpart, on, rel,
                               nullptr);
        });
        timed(1, [&] {
            T* xu = nullptr;
            if constexpr (std::is_same<T, float>::value) xu = w.xc.as<float>();
            else xu = w.x.as<double>();
            if constexpr (K == 64 && std::is_same<T, float>::value) {
                if (xi(s)) {
                    if (dinv)
                        hipLaunchKernelGGL((k_xi_update<true>), grid, dim3(kBlock), 0, st, op.tb.NX, op.tb.NY,
                                           op.tb.mface, r, q, cs, on, part, (const float*)dinv);
                    else
                        hipLaunchKernelGGL((k_xi_update<false>), grid, dim3(kBlock), 0, st, op.tb.NX, op.tb.NY,
                                           op.tb.mface, r, q, cs, on, part, (const float*)nullptr);
                    return;
                }
            }
            if constexpr (K == 64) {
                if (dinv)
                    hipLaunchKernelGGL((k_update64<T, true>), grid, dim3(kBlock), 0, st, op.NN, op.tb.NX, op.tb.mface,
                                       r, q, cs, on, part, dinv);
                else
                    hipLaunchKernelGGL((k_update64<T, false>), grid, dim3(kBlock), 0, st, op.NN, op.tb.NX,
                                       op.tb.mface, r, q, cs, on, part, dinv);
            } else {
                hipLaunchKernelGGL((k_update<T, T, K>), grid, dim3(kBlock), 0, st, op.NN, op.tb.NX, op.tb.mface, xu,
                                   r, p, q, cs, on, part, dinv);
            }
        });
        timed(3, [&] {
            fold();
            hipLaunchKernelGGL((k_finalize<K>), dim3(w.ntiles), dim3(K * kFinStrands), 0, st, fG, 1, (int)mixed, R, cs, fpart, on, rel,
                               nullptr);
        });
        timed(2, [&] {
            T* xu = nullptr;
            if constexpr (std::is_same<T, float>::value) xu = w.xc.as<float>();
            else xu = w.x.as<double>();
            if constexpr (K == 64 && std::is_same<T, float>::value) {
                if (xi(s)) {
                    if (dinv)
                        hipLaunchKernelGGL((k_xi_pupdate<true>), grid, dim3(kBlock), 0, st, op.tb.NX, op.tb.NY, xu, p,
                                           r, cs, (const float*)dinv, w.x.as<double>(), (const int*)rel);
                    else
                        hipLaunchKernelGGL((k_xi_pupdate<false>), grid, dim3(kBlock), 0, st, op.tb.NX, op.tb.NY, xu,
                                           p, r, cs, (const float*)nullptr, w.x.as<double>(), (const int*)rel);
                    return;
                }
            }
            if constexpr (K == 64) {
                if (dinv)
                    hipLaunchKernelGGL((k_pupdate64<T, true>), grid, dim3(kBlock), 0, st, op.NN, xu, p, r, cs, dinv);
                else
                    hipLaunchKernelGGL((k_pupdate64<T, false>), grid, dim3(kBlock), 0, st, op.NN, xu, p, r, cs, dinv);
            } else {
                hipLaunchKernelGGL((k_pupdate<T, K>), grid, dim3(kBlock), 0, st, op.NN, p, r, cs, on, dinv);
            }
        });
        if (mixed) {
            timed(4, [&] {
                if (xi(s))
                    return;  // folded into k_xi_pupdate
                else
                    hipLaunchKernelGGL((k_xaccum<K>), grid, dim3(kBlock), 0, st, op.NN, w.x.as<double>(),
                                       w.xc.as<float>(), cs, rel, 0, (const int*)nullptr);
            });
            timed(5, [&] {
                if constexpr (K == 64)
                {
                    const double* x64 = w.x.as<double>();
                    const double* b64 = w.b.as<double>();
                    if (xi(s) && op.tb.NX <= 16)
                        hipLaunchKernelGGL((k_xi_pencil<double, T, 1, 1>), grid, dim3(kBlock), 0, st, op.p64, x64,
                                           (double*)nullptr, b64, r, cs, rel, part, dinv64);
                    else if (xi(s))
                        hipLaunchKernelGGL((k_xl_pencil<double, T, 1, 1>), grid, dim3(kBlock), 0, st, op.p64, x64,
                                           (double*)nullptr, b64, r, cs, rel, part, dinv64);
                    else if (op.p64.nseg == 1)
                        hipLaunchKernelGGL((k_pencil<double, T, 1, true>), grid, dim3(kBlock),
                                           pencil_lds_bytes(op.p64), st, op.p64, op.NN, x64, (double*)nullptr, b64, r,
                                           cs, rel, part, dinv64);
                    else if (op.p64w_ok)  // one 15-wide segment, no x1 halos
                        hipLaunchKernelGGL((k_pencil<double, T, 1, true, 15>), grid, dim3(kBlock),
                                           (pencil_lds_bytes<double, 15>(op.p64w)), st, op.p64w, op.NN, x64,
                                           (double*)nullptr, b64, r, cs, rel, part, dinv64);
                    else
                        hipLaunchKernelGGL((k_pencil<double, T, 1, false>), grid, dim3(kBlock),
                                           pencil_lds_bytes(op.p64), st, op.p64, op.NN, x64, (double*)nullptr, b64, r,
                                           cs, rel, part, dinv64);
                }
                else
                    hipLaunchKernelGGL((k_residual<T, K>), grid, dim3(kBlock), 0, st, op.v64, op.NN, w.x.as<double>(),
                                       w.b.as<double>(), r, cs, rel, part, dinv64);
            });
            timed(3, [&] {
                fold();
                hipLaunchKernelGGL((k_finalize<K>), dim3(w.ntiles), dim3(K * kFinStrands), 0, st, fG, 2, 1, R, cs, fpart, on, rel,
                                   nullptr);
            });
        }
        HIPCHK(hipGetLastError());
    }

    // Initial state once b, x, r, p (and xc) hold the start of every column.
    static void init_state(eqd_solver* s, hipStream_t st, const int* mask = nullptr) {
        Workspace& w = *s->ws;
        const eqd_op& op = *s->op;
        // b of the tiles the batch fills; the others hold idle columns, whose state ignores it
        dim3 grid(s->op->G, w.live_tiles());
        if (xi(s))
            hipLaunchKernelGGL(k_xi_selfdot, grid, dim3(kBlock), 0, st, op.tb.NX, op.tb.NY, op.tb.mface,
                               w.b.as<double>(), w.partial.as<double>(),
                               s->o.precond == EQD_PRECOND_JACOBI ? op.dinv64 : nullptr);
        else
            hipLaunchKernelGGL((k_selfdot<double, K>), grid, dim3(kBlock), 0, st, op.NN, op.tb.NX, op.tb.mface,
                               w.b.as<double>(), w.partial.as<double>(),
                               s->o.precond == EQD_PRECOND_JACOBI ? op.dinv64 : nullptr);
        const int64_t itmax = s->o.itmax > 0 ? s->o.itmax : 2 * op.ntot;
        if (op.nfold)
            hipLaunchKernelGGL((k_fold_partials<K>), dim3(op.nfold, w.live_tiles()), dim3(K * 16), 0, st, op.G, op.fchunk,
                               (const double*)w.partial.as<double>(), w.folded.as<double>());
        hipLaunchKernelGGL((k_init_state<K>), dim3(w.ntiles), dim3(K), 0, st, op.nfold ? op.nfold : op.G,
                           w.cs.as<ColState>(), op.nfold ? w.folded.as<double>() : w.partial.as<double>(),
                           w.colid.as<int64_t>(), s->o.rtol, s->o.atol, itmax,
                           w.tile_on.as<int>(), mask);
        HIPCHK(hipGetLastError());
    }

    // Measurement run: no graph, events around every kernel, tile activity read
    // before each iteration so the stencil's active tiles are known per launch.
    static void solve_timed(eqd_solver* s) {
        Workspace& w = *s->ws;
        hipStream_t st = s->stream;
        KTimer& kt = *s->kt;
        const int64_t itmax = s->o.itmax > 0 ? s->o.itmax : 2 * s->op->ntot;
        int64_t iters = 0, tile_iters = 0;
        double ms = 0;
        for (;;) {
            HIPCHK(hipMemcpyAsync(w.h_tile_on, w.tile_on.p, w.ntiles * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            int act = 0;
            for (int t = 0; t < w.ntiles; ++t) act += w.h_tile_on[t] != 0;
            if (!act) break;
            if (iters > itmax + 1) throw Error(EQD_ERR_DEVICE, "CG iteration guard exceeded");
            kt.cur_active = act;
            iteration(s, st, &kt);
            HIPCHK(hipStreamSynchronize(st));
            // tile_rel keeps the flags that the iteration's FP64 residual pass acted on
            kt.cur_rel = 0;
            if (std::is_same<T, float>::value) {
                std::vector<int> rel(w.ntiles);
                HIPCHK(hipMemcpy(rel.data(), w.tile_rel.p, w.ntiles * sizeof(int), hipMemcpyDeviceToHost));
                for (int v : rel) kt.cur_rel += v != 0;
            }
            ms += kt.flush();
            ++iters;
            tile_iters += act;
        }
        if (std::is_same<T, float>::value) {
            dim3 grid(s->op->G, w.live_tiles());
            xaccum_all(s, st, grid, nullptr);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));
        s->last_ms = ms;
        s->last_launches = iters;
        s->last_tile_iters = tile_iters;
    }

    // The HIP graph of `gi` CG iterations over the workspace (captured once per shape).
    static int ensure_graph(eqd_solver* s) {
        Workspace& w = *s->ws;
        hipStream_t st = s->stream;
        const int gi = s->o.graph_iters > 0 ? s->o.graph_iters : 8;
        if (!w.exec || w.exec_iters != gi) {
            if (w.exec) HIPCHK(hipGraphExecDestroy(w.exec));
            w.exec = nullptr;
            hipGraph_t graph;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < gi; ++i) iteration(s, st);
            HIPCHK(hipStreamEndCapture(st, &graph));
            HIPCHK(hipGraphInstantiate(&w.exec, graph, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(graph));
            w.exec_iters = gi;
        }
        return gi;
    }

    // Device time (ms) of `iters` CG iterations on the current workspace (autotuning).
    static double trial(eqd_solver* s, int iters) {
        hipStream_t st = s->stream;
        for (int i = 0; i < 2; ++i) iteration(s, st);  // warm-up
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, st));
        for (int i = 0; i < iters; ++i) iteration(s, st);
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return ms;
    }

    // One replay of the iteration graph (the serving loop's step).
    static void replay(eqd_solver* s) {
        ensure_graph(s);
        HIPCHK(hipGraphLaunch(s->ws->exec, s->stream));
    }

    // Iterate until no tile has an active column; graphs of `gi` iterations.
    static void solve(eqd_solver* s) {
        if (s->o.timing) return solve_timed(s);
        Workspace& w = *s->ws;
        hipStream_t st = s->stream;
        const int gi = ensure_graph(s);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, st));
        const int64_t itmax = s->o.itmax > 0 ? s->o.itmax : 2 * s->op->ntot;
        int64_t replays = 0, tile_iters = 0;
        for (;;) {
            HIPCHK(hipMemcpyAsync(w.h_tile_on, w.tile_on.p, w.ntiles * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            int act = 0;
            for (int t = 0; t < w.ntiles; ++t) act += w.h_tile_on[t] != 0;
            if (!act) break;
            if (replays * gi > itmax + gi) throw Error(EQD_ERR_DEVICE, "CG iteration guard exceeded");
            HIPCHK(hipGraphLaunch(w.exec, st));
            ++replays;
            tile_iters += (int64_t)act * gi;
        }
        if (std::is_same<T, float>::value) {  // fold any correction still held in FP32
            dim3 grid(s->op->G, w.live_tiles());
            xaccum_all(s, st, grid, nullptr);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        s->last_ms = ms;
        s->last_launches = replays * gi;
        s->last_tile_iters = tile_iters;
    }

    // The same iterations as solve(), launched one by one instead of through a captured graph
    // and without events or pinned host memory: tile activity is polled into a host word every
    // `gi` iterations (the allocation-free entry point's loop).
    static void solve_direct(eqd_solver* s) {
        Workspace& w = *s->ws;
        hipStream_t st = s->stream;
        const int gi = s->o.graph_iters > 0 ? s->o.graph_iters : 8;
        const int64_t itmax = s->o.itmax > 0 ? s->o.itmax : 2 * s->op->ntot;
        int64_t chunks = 0;
        for (;;) {
            std::vector<int> on(w.ntiles, 0);
            HIPCHK(hipMemcpyAsync(on.data(), w.tile_on.p, w.ntiles * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            int act = 0;
            for (int v : on) act += v != 0;
            if (!act) break;
            if (chunks * gi > itmax + gi) throw Error(EQD_ERR_DEVICE, "CG iteration guard exceeded");
            for (int i = 0; i < gi; ++i) iteration(s, st);
            ++chunks;
        }
        if (std::is_same<T, float>::value) {  // fold any correction still held in FP32
            dim3 grid(s->op->G, w.live_tiles());
            xaccum_all(s, st, grid, nullptr);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));
        s->last_launches = chunks * gi;
    }

    // Plain engine: one column, host-side alpha/beta (device->host read of the partial
    // sums every half iteration), no graphs -- the one-RHS baseline.
    static void solve_plain(eqd_solver* s) {
        static_assert(K == 1, "plain engine is single-column");
        Workspace& w = *s->ws;
        const eqd_op& op = *s->op;
        hipStream_t st = s->stream;
        dim3 grid(op.G, 1);
        std::vector<double> part(op.G);
        const double* dinv = s->o.precond == EQD_PRECOND_JACOBI ? op.dinv64 : nullptr;
        ColState h;
        // the state was initialised on the solver's (non-blocking) stream
        HIPCHK(hipMemcpyAsync(&h, w.cs.p, sizeof(ColState), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, st));
        const double meps = 2.220446049250313e-16;
        int one = 1;
        int64_t launches = 0;
        HIPCHK(hipMemcpy(w.tile_on.p, &one, sizeof(int), hipMemcpyHostToDevice));
        auto sum = [&]() {
            HIPCHK(hipMemcpyAsync(part.data(), w.partial.p, op.G * 8, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            double acc = 0.0;
            for (double v : part) acc += v;
            return acc;
        };
        while (h.status == kActive) {
            hipLaunchKernelGGL((k_apply_dot<double, 1>), grid, dim3(kBlock), 0, st, op.v64, op.NN,
                               w.p.as<double>(), w.q.as<double>(), w.tile_on.as<int>(), w.partial.as<double>());
            ++launches;
            const double pq = sum();
            if (pq <= meps * h.pnorm2 && std::fabs(pq) <= meps * h.pnorm2) {
                h.status = kZeroCurv;
                break;
            }
            h.alpha = h.rr / pq;
            HIPCHK(hipMemcpyAsync(w.cs.p, &h, sizeof(ColState), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL((k_update<double, double, 1>), grid, dim3(kBlock), 0, st, op.NN, op.tb.NX, op.tb.mface,
                               w.x.as<double>(), w.r.as<double>(), w.p.as<double>(), w.q.as<double>(),
                               w.cs.as<ColState>(), w.tile_on.as<int>(), w.partial.as<double>(), dinv);
            const double rr = sum();
            const double rnorm = std::sqrt(rr);
            h.iters += 1;
            if (rnorm <= h.eps || rnorm + 1.0 <= 1.0) {
                h.status = kSolved;
                break;
            }
            h.beta = rr / h.rr;
            h.pnorm2 = rr + h.beta * h.beta * h.pnorm2;
            h.rr = rr;
            if (h.iters >= h.itmax) {
                h.status = kTired;
                break;
            }
            HIPCHK(hipMemcpyAsync(w.cs.p, &h, sizeof(ColState), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL((k_pupdate<double, 1>), grid, dim3(kBlock), 0, st, op.NN, w.p.as<double>(),
                               w.r.as<double>(), w.cs.as<ColState>(), w.tile_on.as<int>(), dinv);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipMemcpyAsync(w.cs.p, &h, sizeof(ColState), hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        s->last_ms = ms;
        s->last_launches = launches;
        s->last_tile_iters = h.iters;
    }

    // Builds b (and the Krylov start vectors) from the slip fields already on the device.
    static void rhs_from_fields(eqd_solver* s, const int* mask = nullptr) {
        Workspace& w = *s->ws;
        const eqd_op& op = *s->op;
        dim3 grid(s->op->G, w.live_tiles());
        if (xi(s))
            hipLaunchKernelGGL(k_xi_rhs, grid, dim3(kBlock), 0, s->stream, op.rv, w.fields.as<double>(), w.ntiles * K,
                               w.b.as<double>(), w.x.as<double>(), w.r.as<float>(), w.p.as<float>(),
                               w.xc.as<float>(), mask, s->o.precond == EQD_PRECOND_JACOBI ? op.dinv64 : nullptr);
        else
            hipLaunchKernelGGL((k_rhs<T, K>), grid, dim3(kBlock), 0, s->stream, op.rv, op.NN, w.fields.as<double>(),
                               w.ntiles * K, w.b.as<double>(), w.x.as<double>(), w.r.as<T>(), w.p.as<T>(),
                               std::is_same<T, float>::value ? w.xc.as<float>() : nullptr, mask,
                               s->o.precond == EQD_PRECOND_JACOBI ? op.dinv64 : nullptr);
        HIPCHK(hipGetLastError());
    }

    // Folds the FP32 correction of the masked columns into x (mixed precision).
    static void fold(eqd_solver* s, const int* mask) {
        if constexpr (std::is_same<T, float>::value) {
            Workspace& w = *s->ws;
            dim3 grid(s->op->G, w.ntiles);
            xaccum_all(s, s->stream, grid, mask);
            HIPCHK(hipGetLastError());
        }
    }

    static void extract(eqd_solver* s, double* Kdev, int64_t ldK, const int* mask = nullptr) {
        Workspace& w = *s->ws;
        const eqd_op& op = *s->op;
        const int npb = kBlock / K;
        dim3 grid((op.nf + npb - 1) / npb, w.ntiles);
        if (xi(s))
            hipLaunchKernelGGL((k_extract<K, true>), grid, dim3(kBlock), 0, s->stream, op.ev, op.NN, w.x.as<double>(),
                               w.fields.as<double>(), w.outpos.as<int64_t>(), Kdev, ldK, (int64_t)0, mask);
        else
            hipLaunchKernelGGL((k_extract<K>), grid, dim3(kBlock), 0, s->stream, op.ev, op.NN, w.x.as<double>(),
                               w.fields.as<double>(), w.outpos.as<int64_t>(), Kdev, ldK, (int64_t)0, mask);
        HIPCHK(hipGetLastError());
    }
};

template <int K>
struct Dispatch {
    static void solve(eqd_solver* s) {
        if (s->o.engine == EQD_ENGINE_PLAIN) {
            if constexpr (K == 1) Engine<double, 1>::solve_plain(s);
            else throw Error(EQD_ERR_ARG, "the plain engine runs one column (tile_k = 1)");
        } else if (s->ws->prec == EQD_PREC_MIXED) Engine<float, K>::solve(s);
        else Engine<double, K>::solve(s);
    }
    static void init(eqd_solver* s, const int* mask = nullptr) {
        if (s->ws->prec == EQD_PREC_MIXED) Engine<float, K>::init_state(s, s->stream, mask);
        else Engine<double, K>::init_state(s, s->stream, mask);
    }
    static void rhs(eqd_solver* s, const int* mask = nullptr) {
        if (s->ws->prec == EQD_PREC_MIXED) Engine<float, K>::rhs_from_fields(s, mask);
        else Engine<double, K>::rhs_from_fields(s, mask);
    }
    static void fold(eqd_solver* s, const int* mask) {
        if (s->ws->prec == EQD_PREC_MIXED) Engine<float, K>::fold(s, mask);
    }
    static void replay(eqd_solver* s) {
        if (s->ws->prec == EQD_PREC_MIXED) Engine<float, K>::replay(s);
        else Engine<double, K>::replay(s);
    }
    static double trial(eqd_solver* s, int iters) {
        return s->ws->prec == EQD_PREC_MIXED ? Engine<float, K>::trial(s, iters) : Engine<double, K>::trial(s, iters);
    }
    static void extract(eqd_solver* s, double* Kdev, int64_t ldK, const int* mask = nullptr) {
        Engine<double, K>::extract(s, Kdev, ldK, mask);
    }
};

template <typename F>
void with_k(int K, F&& f) {
    switch (K) {
        case 1: f(Dispatch<1>()); break;
        case 8: f(Dispatch<8>()); break;
        case 16: f(Dispatch<16>()); break;
        case 32: f(Dispatch<32>()); break;
        case 64: f(Dispatch<64>()); break;
        default: throw Error(EQD_ERR_ARG, "tile_k must be 1, 8, 16, 32 or 64");
    }
}

// Tile width of a build: 1 for the plain engine, the requested width, or 0 (autotune).
int effective_k(const eqd_solver* s, int) {
    if (s->o.engine == EQD_ENGINE_PLAIN) return 1;
    return s->o.tile_k;
}

// Solves a batch of columns whose slip fields are produced by `fill` into
// ws->fields (host or device), then extracts tractions into Kdev[:, outpos].
// Columns [0, split) and [split, n) run as separate batch sequences (split < 0: one sequence).
template <typename Fill>
void run_columns(eqd_solver* s, const std::vector<int64_t>& colid, const std::vector<int64_t>& outpos, Fill&& fill,
                 double* Kdev, int64_t ldK, int split = -1) {
    const int n = (int)colid.size();
    if (split < 0 || split > n) split = n;
    const int prec = s->o.engine == EQD_ENGINE_PLAIN ? EQD_PREC_FP64 : s->o.precision;
    int per = s->o.max_cols_in_flight > 0 ? s->o.max_cols_in_flight : n;
    {  // bound by device memory
        size_t freeb = 0, totb = 0;
        HIPCHK(hipMemGetInfo(&freeb, &totb));
        size_t cap = (size_t)(0.6 * (double)freeb / (double)bytes_per_column(*s->op, prec));
        per = (int)std::max<size_t>(1, std::min<size_t>((size_t)per, cap));
    }
    if (s->o.engine == EQD_ENGINE_PLAIN) per = 1;
    int K = effective_k(s, n);
    if (K == 0) {  // autotune on this batch: 16 CG iterations of its first columns per candidate width
        int& tuned = s->tuned_k[prec == EQD_PREC_MIXED ? 1 : 0];
        if (!tuned) {
            const int nb = std::min({n, per, 40 * 64});
            double best = 0.0;
            for (int k : {8, 16, 32, 64}) {
                const int nt = (nb + k - 1) / k;
                ensure_ws(s, k, nt, prec);
                Workspace& w = *s->ws;
                const int sl = nt * k;
                std::vector<int64_t> cid(sl, -1), opos(sl, -1);
                for (int i = 0; i < sl && i < nb; ++i) {
                    cid[i] = colid[i];
                    opos[i] = outpos[i];
                }
                HIPCHK(hipMemcpyAsync(w.colid.p, cid.data(), sl * 8, hipMemcpyHostToDevice, s->stream));
                HIPCHK(hipMemcpyAsync(w.outpos.p, opos.data(), sl * 8, hipMemcpyHostToDevice, s->stream));
                fill(w, 0, sl);
                double ms = 0.0;
                with_k(k, [&](auto d) {
                    decltype(d)::rhs(s);
                    decltype(d)::init(s);
                    ms = decltype(d)::trial(s, 16);
                });
                if (!tuned || ms < best) {
                    best = ms;
                    tuned = k;
                }
            }
        }
        K = tuned;
    }
    s->last_k = K;
    // tile_group balances any tile count over the 8 XCDs, so no padding tiles
    const int ntiles = (std::min(per, std::max(split, n - split)) + K - 1) / K;
    ensure_ws(s, K, ntiles, prec);
    Workspace& w = *s->ws;
    const int slots = ntiles * K;
    s->last_iters.assign(n, 0);
    s->kt->reset();
    double ms = 0;
    int64_t launches = 0, tile_iters = 0;
    std::vector<std::pair<int, int>> batches;  // [base, end) of each batch
    for (int b = 0; b < split; b += slots) batches.emplace_back(b, std::min(b + slots, split));
    for (int b = split; b < n; b += slots) batches.emplace_back(b, std::min(b + slots, n));
    struct UsedTiles {  // the tiles a batch fills; every tile again afterwards
        Workspace& w;
        ~UsedTiles() { w.used_tiles = 0; }
    } used_guard{w};
    for (const auto& bt : batches) {
        const int base = bt.first, end = bt.second;
        w.used_tiles = (end - base + K - 1) / K;
        std::vector<int64_t> cid(slots, -1), opos(slots, -1);
        for (int i = 0; i < slots && base + i < end; ++i) {
            cid[i] = colid[base + i];
            opos[i] = outpos[base + i];
        }
        HIPCHK(hipMemcpyAsync(w.colid.p, cid.data(), slots * 8, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(w.outpos.p, opos.data(), slots * 8, hipMemcpyHostToDevice, s->stream));
        fill(w, base, slots);
        with_k(K, [&](auto d) {
            decltype(d)::rhs(s);
            decltype(d)::init(s);
            decltype(d)::solve(s);
            decltype(d)::extract(s, Kdev, ldK);
        });
        HIPCHK(hipStreamSynchronize(s->stream));
        ms += s->last_ms;
        launches += s->last_launches;
        tile_iters += s->last_tile_iters;
        std::vector<ColState> hs(slots);
        HIPCHK(hipMemcpy(hs.data(), w.cs.p, slots * sizeof(ColState), hipMemcpyDeviceToHost));
        for (int i = 0; i < slots && base + i < end; ++i) {
            const ColState& c = hs[i];
            s->st.solves += 1;
            s->st.iterations += c.iters;
            s->st.max_iterations = std::max<int64_t>(s->st.max_iterations, c.iters);
            s->last_iters[base + i] = c.iters;
            if (c.status != kSolved) {
                s->st.unconverged += 1;
                if (s->st.unconverged == 1)
                    std::fprintf(stderr,
                                 "warning: CG did not converge on a split-node solve after %d iterations; "
                                 "raise itmax or loosen rtol\n",
                                 c.iters);
            }
        }
    }
    s->last_ms = ms;
    s->last_launches = launches;
    s->last_tile_iters = tile_iters;
}

void unit_fill(eqd_solver* s, Workspace& w, int slots) {
    const eqd_op& op = *s->op;
    hipLaunchKernelGGL(k_unit_fields, dim3(slots), dim3(256), 0, s->stream, op.g.n23, op.i2lo, op.n2f, op.nf, slots,
                       w.colid.as<int64_t>(), w.fields.as<double>(), (const int*)nullptr);
    HIPCHK(hipGetLastError());
}

__global__ void k_toeplitz_fill(int nf, int n2, int n3, const double* __restrict__ ker, const int* __restrict__ have,
                                double* __restrict__ K, int64_t ldK) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nf * nf) return;
    const int i = (int)(idx % nf), j = (int)(idx / nf);
    const int ai = i % n2, bi = i / n2, aj = j % n2, bj = j / n2;
    const int W2 = 2 * n2 - 1;
    const int key = (ai - aj + n2 - 1) + W2 * (bi - bj + n3 - 1);
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
    if (have[key]) {
        v0 = ker[4 * key];
        v1 = ker[4 * key + 1];
        v2 = ker[4 * key + 2];
        v3 = ker[4 * key + 3];
    }
    K[i + ldK * j] = v0;
    K[nf + i + ldK * j] = v1;
    K[i + ldK * (nf + j)] = v2;
    K[nf + i + ldK * (nf + j)] = v3;
}

void stiffness(eqd_solver* s, int mode, const int* cols, int ncols, double* Kdev, int64_t ldK) {
    const eqd_op& op = *s->op;
    const int nf = op.nf;
    if (mode == EQD_STIFFNESS_EXACT) {
        std::vector<int64_t> cid, opos;
        if (!cols) {
            if (ldK < 2 * nf) throw Error(EQD_ERR_ARG, "ldK must be >= 2nf");
            for (int i = 0; i < 2 * nf; ++i) {
                cid.push_back(i);
                opos.push_back(i);
            }
        } else {
            if (ncols < 0) throw Error(EQD_ERR_ARG, "ncols must be >= 0");
            if (ldK < 2 * nf) throw Error(EQD_ERR_ARG, "ldK must be >= 2nf");
            for (int i = 0; i < ncols; ++i) {
                if (cols[i] < 0 || cols[i] >= 2 * nf) throw Error(EQD_ERR_ARG, "column index out of range");
                cid.push_back(cols[i]);
                opos.push_back(i);
            }
        }
        if (cid.empty()) return;
        int split = -1;
        // A unit slip s2 on face row iz, or s3 on face column iy, outside the interior range
        // [1, n23 - 2] gives b = HP(D+SAT) chi = 0 (k_rhs reads s2 only on interior rows and s3
        // only on interior columns): such columns converge at iteration 0. They run as a batch
        // of their own after the others, so they hold no slots in tiles that iterate and the
        // iterating batch has fewer tiles (l_f = L_fault at dz = 50 m: 132 of the 2178 columns,
        // 32 tiles instead of 35). Each column's result does not depend on its tile or batch,
        // so K is unchanged.
        {
            const int n23 = op.g.n23;
            auto zero_rhs = [&](int64_t c) {
                const int k = (int)(c < nf ? c : c - nf);
                const int f = op.i2lo + (c < nf ? k / op.n2f : k % op.n2f);
                return f < 1 || f > n23 - 2;
            };
            std::vector<int64_t> c2, o2;
            size_t nz = 0;
            for (size_t i = 0; i < cid.size(); ++i) nz += !zero_rhs(cid[i]);
            for (int pass = 0; pass < 2; ++pass)
                for (size_t i = 0; i < cid.size(); ++i)
                    if (zero_rhs(cid[i]) == (pass == 1)) {
                        c2.push_back(cid[i]);
                        o2.push_back(opos[i]);
                    }
            cid.swap(c2);
            opos.swap(o2);
            split = (int)nz;
        }
        run_columns(s, cid, opos, [&](Workspace& w, int, int slots) { unit_fill(s, w, slots); }, Kdev, ldK, split);
        return;
    }
    if (mode != EQD_STIFFNESS_TOEPLITZ) throw Error(EQD_ERR_ARG, "mode must be EXACT or TOEPLITZ");
    if (cols) throw Error(EQD_ERR_ARG, "the Toeplitz build produces the whole K (cols must be NULL)");
    const int n2 = op.n2f, n3 = op.n2f;
    if (nf != n2 * n3) throw Error(EQD_ERR_GRID, "the Toeplitz expansion assumes Omega_f forms a full n2 x n3 grid");
    if (ldK < 2 * nf) throw Error(EQD_ERR_ARG, "ldK must be >= 2nf");
    // sources in priority order: centre, then the four corners (src/FaultResponse.jl:351-353)
    const int ac = (n2 + 1) / 2 - 1, bc = (n3 + 1) / 2 - 1;
    const int src[5] = {ac + n2 * bc, 0, n2 - 1, n2 * (n3 - 1), n2 - 1 + n2 * (n3 - 1)};
    std::vector<int64_t> cid, opos;
    for (int k = 0; k < 5; ++k) {
        cid.push_back(src[k]);
        opos.push_back(2 * k);
        cid.push_back(nf + src[k]);
        opos.push_back(2 * k + 1);
    }
    DevMem cols10((size_t)2 * nf * 10 * 8);
    run_columns(s, cid, opos, [&](Workspace& w, int, int slots) { unit_fill(s, w, slots); }, cols10.as<double>(),
                2 * nf);
    std::vector<double> hc((size_t)2 * nf * 10);
    HIPCHK(hipMemcpy(hc.data(), cols10.p, hc.size() * 8, hipMemcpyDeviceToHost));
    const int W2 = 2 * n2 - 1, W3 = 2 * n3 - 1;
    std::vector<int> have((size_t)W2 * W3, 0);
    std::vector<double> ker((size_t)W2 * W3 * 4, 0.0);
    for (int k = 0; k < 5; ++k) {
        const int as = src[k] % n2, bs = src[k] / n2;
        const double* c2 = &hc[(size_t)(2 * k) * 2 * nf];
        const double* c3 = &hc[(size_t)(2 * k + 1) * 2 * nf];
        for (int i = 0; i < nf; ++i) {
            const int ai = i % n2, bi = i / n2;
            const size_t key = (size_t)(ai - as + n2 - 1) + (size_t)W2 * (bi - bs + n3 - 1);
            if (have[key]) continue;  // the earlier source wins
            have[key] = 1;
            ker[4 * key] = c2[i];
            ker[4 * key + 1] = c2[nf + i];
            ker[4 * key + 2] = c3[i];
            ker[4 * key + 3] = c3[nf + i];
        }
    }
    DevMem dker(ker.size() * 8), dhave(have.size() * sizeof(int));
    HIPCHK(hipMemcpy(dker.p, ker.data(), ker.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dhave.p, have.data(), have.size() * sizeof(int), hipMemcpyHostToDevice));
    const int64_t tot = (int64_t)nf * nf;
    hipLaunchKernelGGL(k_toeplitz_fill, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s->stream, nf, n2, n3,
                       dker.as<double>(), dhave.as<int>(), Kdev, ldK);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
}

}  // namespace

// ---------------------------------------------------------------------------------
// Serving loop: continuous batching of stiffness-column and forward-solve requests.
//
// Each grid added to a server owns an operator, a solver and a resident pool of 64-column
// tiles sized by the server's HBM budget. A step retires the columns that finished (their
// FP32 correction folded into x, tractions extracted into a per-slot result arena and
// copied to the request), admits queued requests into the freed lanes (slip fields, RHS and
// Krylov start built for those lanes only), and replays the CG iteration graph once. A
// column's arithmetic depends on its own lane only, so its result is the same alone and in
// any mix of requests.

struct ServeRequest {
    int grid = 0;
    int kind = 0;  // 0: unit-slip columns, 1: forward solve
    std::vector<int64_t> cols;
    std::vector<double> fields;  // kind 1: slip fields on the face grid
    int ncols = 0, pending = 0;
    std::vector<double> result;  // 2nf x ncols, column-major
    std::vector<int> iters;
    int unconverged = 0;
    bool done = false;
};

struct ServeGrid {
    eqd_op* op = nullptr;
    eqd_solver* s = nullptr;
    int ntiles = 0, slots = 0;
    std::vector<int64_t> slot_req, colid;
    std::vector<int> slot_idx;
    DevMem arena, mask_fresh, mask_unit, mask_retire;
    std::deque<int64_t> queue;
    std::vector<ColState> hcs;
    int64_t busy = 0;
    hipEvent_t ev = nullptr;  // recorded after each replay: the grid steps again once it has passed
    bool inflight = false;
    ~ServeGrid() {
        if (ev) (void)hipEventDestroy(ev);
        if (s) eqd_solver_destroy(s);
        if (op) eqd_operator_destroy(op);
    }
};

struct eqd_server {
    eqd_cg_opts o{};
    size_t budget = 0, used = 0;
    std::vector<std::unique_ptr<ServeGrid>> grids;
    std::map<int64_t, ServeRequest> reqs;
    std::deque<int64_t> completed;
    int64_t next_id = 1;
};

namespace {

void serve_check_grid(const eqd_server* sv, int grid) {
    if (grid < 0 || grid >= (int)sv->grids.size()) throw Error(EQD_ERR_ARG, "unknown grid id");
}

// Retire finished columns of one grid, admit queued requests into free lanes, replay once.
// A grid whose previous replay is still running is left alone (grids run on their own
// streams, so a large grid does not hold back the steps of small ones).
int64_t serve_grid_step(eqd_server* sv, ServeGrid& G) {
    if (G.inflight) {
        const hipError_t q = hipEventQuery(G.ev);
        if (q == hipErrorNotReady) return G.busy;
        HIPCHK(q);
        G.inflight = false;
    }
    eqd_solver* s = G.s;
    Workspace& w = *s->ws;
    const eqd_op& op = *G.op;
    const int n2 = 2 * op.nf;
    std::vector<int> retire(G.slots, 0), fresh(G.slots, 0), unit(G.slots, 0);
    bool any_retire = false, any_change = false;
    if (G.busy > 0) {
        HIPCHK(hipMemcpyAsync(G.hcs.data(), w.cs.p, G.slots * sizeof(ColState), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int i = 0; i < G.slots; ++i)
            if (G.slot_req[i] >= 0 && G.hcs[i].status != kActive) {
                retire[i] = 1;
                any_retire = true;
            }
    }
    if (any_retire) {
        HIPCHK(hipMemcpyAsync(G.mask_retire.p, retire.data(), G.slots * sizeof(int), hipMemcpyHostToDevice,
                              s->stream));
        with_k(64, [&](auto d) {
            decltype(d)::fold(s, G.mask_retire.as<int>());
            decltype(d)::extract(s, G.arena.as<double>(), n2, G.mask_retire.as<int>());
        });
        std::vector<int64_t> finished;
        for (int i = 0; i < G.slots; ++i) {
            if (!retire[i]) continue;
            finished.push_back(G.slot_req[i]);
            ServeRequest& r = sv->reqs.at(G.slot_req[i]);
            const int j = G.slot_idx[i];
            HIPCHK(hipMemcpyAsync(r.result.data() + (size_t)j * n2, G.arena.as<double>() + (size_t)i * n2, n2 * 8,
                                  hipMemcpyDeviceToHost, s->stream));
            const ColState& c = G.hcs[i];
            r.iters[j] = c.iters;
            s->st.solves += 1;
            s->st.iterations += c.iters;
            s->st.max_iterations = std::max<int64_t>(s->st.max_iterations, c.iters);
            if (c.status != kSolved) {
                r.unconverged += 1;
                s->st.unconverged += 1;
            }
            G.slot_req[i] = -1;
            G.colid[i] = -1;
            G.busy -= 1;
            fresh[i] = 1;  // re-initialised below: idle unless a new column takes the lane
            any_change = true;
        }
        HIPCHK(hipStreamSynchronize(s->stream));  // the result copies have landed
        for (int64_t id : finished) {
            ServeRequest& r = sv->reqs.at(id);
            if (--r.pending == 0) {
                r.done = true;
                sv->completed.push_back(id);
            }
        }
    }
    // admission, whole requests in arrival order
    std::vector<int> freelist;
    for (int i = 0; i < G.slots; ++i)
        if (G.slot_req[i] < 0) freelist.push_back(i);
    size_t fi = 0;
    std::vector<std::pair<int, const double*>> fwd;
    while (!G.queue.empty()) {
        ServeRequest& r = sv->reqs.at(G.queue.front());
        if ((size_t)r.ncols > freelist.size() - fi) break;
        const int64_t id = G.queue.front();
        G.queue.pop_front();
        for (int j = 0; j < r.ncols; ++j) {
            const int i = freelist[fi++];
            G.slot_req[i] = id;
            G.slot_idx[i] = j;
            fresh[i] = 1;
            if (r.kind == 0) {
                G.colid[i] = r.cols[j];
                unit[i] = 1;
            } else {
                G.colid[i] = (int64_t)1 << 40;  // not a unit column: fields come from the request
                fwd.emplace_back(i, r.fields.data());
            }
            G.busy += 1;
            any_change = true;
        }
    }
    if (any_change) {
        HIPCHK(hipMemcpyAsync(w.colid.p, G.colid.data(), G.slots * 8, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(G.mask_fresh.p, fresh.data(), G.slots * sizeof(int), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(G.mask_unit.p, unit.data(), G.slots * sizeof(int), hipMemcpyHostToDevice, s->stream));
        const size_t fsz = 2 * (size_t)op.g.n23 * op.g.n23;
        for (auto& f : fwd)
            HIPCHK(hipMemcpyAsync(w.fields.as<double>() + (size_t)f.first * fsz, f.second, fsz * 8,
                                  hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(k_unit_fields, dim3(G.slots), dim3(256), 0, s->stream, op.g.n23, op.i2lo, op.n2f, op.nf,
                           G.slots, w.colid.as<int64_t>(), w.fields.as<double>(), G.mask_unit.as<int>());
        HIPCHK(hipGetLastError());
        with_k(64, [&](auto d) {
            decltype(d)::rhs(s, G.mask_fresh.as<int>());
            decltype(d)::init(s, G.mask_fresh.as<int>());
        });
        // host copies of request fields must outlive the async copies above
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    if (G.busy > 0) {
        with_k(64, [&](auto d) { decltype(d)::replay(s); });
        HIPCHK(hipEventRecord(G.ev, s->stream));
        G.inflight = true;
    }
    return G.busy;
}

}  // namespace

extern "C" {

const char* eqd_version(void) { return "libeqd 0.2.0 (gfx950)"; }

int64_t eqd_tables_export(const eqd_grid_desc* g, int which, int* start, int* width, double* coef) {
    int64_t n = -1;
    eqd_status rc = guard([&] {
        if (!g) throw Error(EQD_ERR_ARG, "null argument");
        validate_grid(*g);
        HostTables t = build_tables(g->n1, g->n23, g->L_normal, g->L_fault, g->lambda, g->mu);
        const HostTab* tabs[9] = {&t.AXs, &t.AXo, &t.MX, &t.Ts, &t.To, &t.D1, &t.D1s, &t.D1full, &t.D1ext};
        const std::vector<double>* vecs[5] = {&t.Hx, &t.Hy, &t.Jt, &t.Jn, &t.Dn};
        if (which >= 0 && which < 9) {
            const HostTab& h = *tabs[which];
            n = h.nrows;
            if (start) std::memcpy(start, h.start.data(), h.nrows * sizeof(int));
            if (width) std::memcpy(width, h.width.data(), h.nrows * sizeof(int));
            if (coef) std::memcpy(coef, h.c.data(), h.c.size() * sizeof(double));
        } else if (which >= 100 && which < 105) {
            const std::vector<double>& v = *vecs[which - 100];
            n = (int64_t)v.size();
            if (coef) std::memcpy(coef, v.data(), v.size() * sizeof(double));
        } else if (which == 105) {
            n = 1;
            if (coef) coef[0] = t.Jext;
        } else {
            throw Error(EQD_ERR_ARG, "unknown table id");
        }
    });
    return rc == EQD_OK ? n : -1;
}

size_t eqd_last_error(char* buf, size_t len) {
    if (buf && len) std::snprintf(buf, len, "%s", g_err.c_str());
    return g_err.size();
}

void eqd_cg_opts_default(eqd_cg_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->rtol = 1e-10;
    o->atol = 0.0;
    o->itmax = 0;
    o->precond = EQD_PRECOND_NONE;
    o->precision = EQD_PREC_FP64;
    o->engine = EQD_ENGINE_BATCHED;
    o->tile_k = 0;
    o->max_cols_in_flight = 0;
    o->reliable_interval = 48;
    o->graph_iters = 8;
}

eqd_status eqd_operator_create(const eqd_grid_desc* g, eqd_op** out) {
    return guard([&] {
        if (!g || !out) throw Error(EQD_ERR_ARG, "null argument");
        *out = nullptr;
        validate_grid(*g);
        int lo = 0, cnt = 0;
        omega_range(*g, lo, cnt);
        check_device();
        std::unique_ptr<eqd_op> op(new eqd_op());
        op->g = *g;
        op->tb = build_tables(g->n1, g->n23, g->L_normal, g->L_fault, g->lambda, g->mu);
        op->NN = op->tb.NX * op->tb.NY * op->tb.NY;
        op->N = (int64_t)g->n1 * g->n23 * g->n23;
        op->ntot = 6 * op->N;
        op->i2lo = lo;
        op->n2f = cnt;
        op->nf = cnt * cnt;
        op->v64 = op->view<double>();
        op->v32 = op->view<float>();
        op->p64 = op->pview<double>();
        op->p32 = op->pview<float>();
        {
            const char* sv = std::getenv("EQD_STENCIL");
            // 9 <= NX keeps the FP64 bands of short lines at 16 rows (two 8-node segments), as the
            // kernels read them; lines of more than 64 nodes keep the node-major pencils. x1-inner
            // tiles pad the lines (to 16 nodes, or a multiple of 8): their FP64 tile vector must also
            // stay below the 2^31 B of the buffer resource's offsets, else the grid keeps the
            // node-major layout
            const bool fits = (int64_t)xi_tile_elems(op->tb.NX, op->tb.NY) * 8 <= (int64_t)INT32_MAX;
            op->xi = op->tb.NX >= 9 && op->tb.NX <= 64 && fits && !(sv && std::strcmp(sv, "pencil") == 0);
        }
        if (op->tb.NX > pencil_s<double>() && op->tb.NX <= 15) {
            op->p64w = op->p64;
            op->p64w.nseg = 1;
            op->p64w_ok = true;
        }
        // Groups per tile: about 4 FP32 pencils per workgroup (one per wave), at least 64 and a
        // multiple of 64. With one pencil per wave the blocks of a tile sweep its pencils in
        // order and are short, so the pencils in flight on an XCD are one window of neighbours
        // (dz = 50 m: 961 pencils -> 256 groups; stencil 640 -> 558 us per launch against 64
        // groups of 4 pencils per wave, whose strided sweeps kept two tiles' windows in each L2;
        // 512 groups leave half the waves idle, 602 us). dz = 25 m: 64,516 pencils -> 16,192
        // groups; grids above kFoldAbove groups sum their partials in two levels
        // (k_fold_partials) so that k_finalize stays one short pass. Depends on the grid only,
        // never on the batch.
        {
            const int np = op->p32.NY * op->p32.NY * op->p32.nseg;
            const int per = EQD_PENCILS_PER_GROUP * 64;
            op->G = std::min(kMaxGroups, std::max(kGroups, (np + per - 1) / per * 64));
            if (op->G > kFoldAbove) {
                op->nfold = 16 * ((op->G + 16 * 256 - 1) / (16 * 256));
                op->fchunk = (op->G + op->nfold - 1) / op->nfold;
            }
        }
        op->rv = RhsView{op->tb.NX, op->tb.NY, g->n23, lo, cnt, op->nf, op->up(op->tb.Jt), op->up(op->tb.Jn),
                         op->v64.Hx, op->v64.Hy, op->tab<double>(op->tb.D1full)};
        op->ev = ExtractView{op->tb.NX, op->tb.NY, op->tb.mface, g->n23, lo, cnt, op->nf, g->mu, op->tb.Jext,
                             op->up(op->tb.Dn), op->tab<double>(op->tb.D1ext)};
        *out = op.release();
    });
}

eqd_status eqd_operator_destroy(eqd_op* op) {
    return guard([&] {
        if (op && --op->refs == 0) delete op;
    });
}

eqd_status eqd_operator_info(const eqd_op* op, eqd_op_info* info) {
    return guard([&] {
        if (!op || !info) throw Error(EQD_ERR_ARG, "null argument");
        info->nf = op->nf;
        info->n2f = info->n3f = op->n2f;
        info->ntot = op->ntot;
        info->nreduced = (int64_t)op->NN * 3;
        info->h1 = op->tb.h1;
        info->h2 = op->tb.h2;
    });
}

eqd_status eqd_fault_grid_axes(const eqd_op* op, double* x2f, double* x3f) {
    return guard([&] {
        if (!op || !x2f || !x3f) throw Error(EQD_ERR_ARG, "null argument");
        for (int i = 0; i < op->n2f; ++i) x2f[i] = x3f[i] = -op->g.L_fault + (op->i2lo + i) * op->tb.h2;
    });
}

eqd_status eqd_apply_A(eqd_op* op, const double* x, double* y) {
    return guard([&] {
        if (!op || !x || !y) throw Error(EQD_ERR_ARG, "null argument");
        const size_t n = (size_t)op->NN * 3;
        std::vector<double> xr(n), yr(n);
        reduce_full(*op, x, xr.data(), false);  // P x
        DevMem dx(n * 8), dy(n * 8), part((size_t)kGroups * 8);
        int one = 1;
        DevMem on(sizeof(int));
        HIPCHK(hipMemcpy(on.p, &one, sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dx.p, xr.data(), n * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL((k_apply_dot<double, 1>), dim3(kGroups, 1), dim3(kBlock), 0, 0, op->v64, op->NN,
                           dx.as<double>(), dy.as<double>(), on.as<int>(), part.as<double>());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(yr.data(), dy.p, n * 8, hipMemcpyDeviceToHost));
        embed_full(*op, yr.data(), y);
    });
}

eqd_status eqd_jacobi_preconditioner(const eqd_grid_desc* g, double* minv, int64_t n) {
    return guard([&] {
        if (!g || !minv) throw Error(EQD_ERR_ARG, "null argument");
        validate_grid(*g);
        const int n1 = g->n1, n23 = g->n23;
        const int64_t N = (int64_t)n1 * n23 * n23;
        if (n != 6 * N) throw Error(EQD_ERR_ARG, "minv must hold the 6 n1 n23^2 full-space entries");
        const HostTables tb = build_tables(n1, n23, g->L_normal, g->L_fault, g->lambda, g->mu);
        const std::vector<double> d = merged_diagonal(tb);
        auto lin = [&](int i1, int i2, int i3) { return i1 + (int64_t)n1 * (i2 + (int64_t)n23 * i3); };
        // far-field rows of A are zero: their diagonal entries are floored to 1
        std::fill(minv, minv + n, 1.0);
        for (int z = 0; z < tb.NY; ++z)
            for (int y = 0; y < tb.NY; ++y)
                for (int m = 0; m < tb.NX; ++m) {
                    const int M = m + 1;  // plane on the two-block x1 line; n1 - 1 is the fault face
                    const size_t node = (size_t)m + (size_t)tb.NX * (y + (size_t)tb.NY * z);
                    for (int c = 0; c < 3; ++c) {
                        const double v = d[node * 3 + c];
                        if (!(v > 0)) throw Error(EQD_ERR_ARG, "diag(A) has non-positive entries; A is not PSD");
                        if (M <= n1 - 1) minv[c * N + lin(M, y + 1, z + 1)] = 1.0 / v;
                        if (M >= n1 - 1) minv[3 * N + c * N + lin(M - (n1 - 1), y + 1, z + 1)] = 1.0 / v;
                    }
                }
    });
}

eqd_status eqd_solver_create(eqd_op* op, const eqd_cg_opts* opts, eqd_solver** out) {
    return guard([&] {
        if (!op || !out) throw Error(EQD_ERR_ARG, "null argument");
        *out = nullptr;
        eqd_cg_opts o;
        eqd_cg_opts_default(&o);
        if (opts) o = *opts;
        if (o.precond != EQD_PRECOND_NONE && o.precond != EQD_PRECOND_JACOBI)
            throw Error(EQD_ERR_ARG, "precond must be :none or :jacobi");
        if (o.precision != EQD_PREC_FP64 && o.precision != EQD_PREC_MIXED) throw Error(EQD_ERR_ARG, "bad precision");
        if (o.engine != EQD_ENGINE_BATCHED && o.engine != EQD_ENGINE_PLAIN) throw Error(EQD_ERR_ARG, "bad engine");
        if (!(o.rtol >= 0) || !(o.atol >= 0)) throw Error(EQD_ERR_ARG, "rtol and atol must be >= 0");
        if (o.precond == EQD_PRECOND_JACOBI) op->ensure_jacobi();
        std::unique_ptr<eqd_solver> s(new eqd_solver());
        s->op = op;
        s->o = o;
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        op->refs++;
        *out = s.release();
    });
}

eqd_status eqd_solver_destroy(eqd_solver* s) {
    return guard([&] {
        if (!s) return;
        s->ws.reset();
        if (s->stream) (void)hipStreamDestroy(s->stream);
        eqd_op* op = s->op;
        delete s;
        if (op && --op->refs == 0) delete op;
    });
}

eqd_status eqd_solver_duplicate(const eqd_solver* s, eqd_solver** out) {
    return guard([&] {
        if (!s || !out) throw Error(EQD_ERR_ARG, "null argument");
        eqd_status rc = eqd_solver_create(s->op, &s->o, out);
        if (rc) throw Error(rc, g_err);
    });
}

eqd_status eqd_solver_merge_stats(eqd_solver* into, const eqd_solver* from) {
    return guard([&] {
        if (!into || !from) throw Error(EQD_ERR_ARG, "null argument");
        into->st.solves += from->st.solves;
        into->st.iterations += from->st.iterations;
        into->st.max_iterations = std::max(into->st.max_iterations, from->st.max_iterations);
        into->st.unconverged += from->st.unconverged;
    });
}

eqd_status eqd_solver_report(const eqd_solver* s, eqd_report* r) {
    return guard([&] {
        if (!s || !r) throw Error(EQD_ERR_ARG, "null argument");
        *r = s->st;
        r->mean_iterations = s->st.solves ? (double)s->st.iterations / (double)s->st.solves : 0.0;
        r->precond = s->o.precond;
        r->tile_k = s->last_k;
    });
}

eqd_status eqd_solver_last_timing(const eqd_solver* s, double* ms, int64_t* launches, int64_t* tile_iters) {
    return guard([&] {
        if (!s) throw Error(EQD_ERR_ARG, "null argument");
        if (ms) *ms = s->last_ms;
        if (launches) *launches = s->last_launches;
        if (tile_iters) *tile_iters = s->last_tile_iters;
    });
}

eqd_status eqd_solver_kernel_timing(const eqd_solver* s, int kernel, double* total_ms, int64_t* launches,
                                    int64_t* active_tiles) {
    return guard([&] {
        if (!s) throw Error(EQD_ERR_ARG, "null argument");
        if (kernel < 0 || kernel >= KTimer::kKinds) throw Error(EQD_ERR_ARG, "unknown kernel id");
        if (total_ms) *total_ms = s->kt->ms[kernel];
        if (launches) *launches = s->kt->launches[kernel];
        if (active_tiles) *active_tiles = s->kt->active[kernel];
    });
}

eqd_status eqd_build_rhs(eqd_op* op, const double* s2, const double* s3, double* rhs) {
    return guard([&] {
        if (!op || !s2 || !s3 || !rhs) throw Error(EQD_ERR_ARG, "null argument");
        const int n23 = op->g.n23;
        const size_t n = (size_t)op->NN * 3;
        std::vector<double> f(2 * (size_t)n23 * n23), b(n);
        slip_fields(*op, s2, s3, f.data());
        DevMem df(f.size() * 8), db(n * 8), dx(n * 8), dr(n * 8), dp(n * 8);
        HIPCHK(hipMemcpy(df.p, f.data(), f.size() * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL((k_rhs<double, 1>), dim3(kGroups, 1), dim3(kBlock), 0, 0, op->rv, op->NN, df.as<double>(), 1,
                           db.as<double>(), dx.as<double>(), dr.as<double>(), dp.as<double>(), (float*)nullptr,
                           (const int*)nullptr, (const double*)nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(b.data(), db.p, n * 8, hipMemcpyDeviceToHost));
        embed_full(*op, b.data(), rhs);
    });
}

eqd_status eqd_split_node_solve(eqd_solver* s, const double* rhs, double* u) {
    return guard([&] {
        if (!s || !rhs || !u) throw Error(EQD_ERR_ARG, "null argument");
        const eqd_op& op = *s->op;
        const size_t n = (size_t)op.NN * 3;
        std::vector<double> br(n);
        reduce_full(op, rhs, br.data(), true);
        const int K = s->o.engine == EQD_ENGINE_PLAIN ? 1 : 1;
        const int prec = s->o.engine == EQD_ENGINE_PLAIN ? EQD_PREC_FP64 : s->o.precision;
        ensure_ws(s, K, 1, prec);
        Workspace& w = *s->ws;
        int64_t cid = 0;
        HIPCHK(hipMemcpy(w.colid.p, &cid, 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(w.b.p, br.data(), n * 8, hipMemcpyHostToDevice));
        // x = 0, r = b, p = z = b (M^{-1} b with the Jacobi preconditioner)
        std::vector<double> zr(br);
        if (s->o.precond == EQD_PRECOND_JACOBI)
            for (size_t i = 0; i < n; ++i) zr[i] = op.dinv_h[i] * br[i];
        HIPCHK(hipMemset(w.x.p, 0, n * 8));
        if (prec == EQD_PREC_MIXED) {
            std::vector<float> bf(br.begin(), br.end()), zf(zr.begin(), zr.end());
            HIPCHK(hipMemcpy(w.r.p, bf.data(), n * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(w.p.p, zf.data(), n * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemset(w.xc.p, 0, n * 4));
        } else {
            HIPCHK(hipMemcpy(w.r.p, br.data(), n * 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(w.p.p, zr.data(), n * 8, hipMemcpyHostToDevice));
        }
        with_k(1, [&](auto d) {
            decltype(d)::init(s);
            decltype(d)::solve(s);
        });
        ColState c;
        HIPCHK(hipMemcpy(&c, w.cs.p, sizeof(c), hipMemcpyDeviceToHost));
        s->st.solves += 1;
        s->st.iterations += c.iters;
        s->st.max_iterations = std::max<int64_t>(s->st.max_iterations, c.iters);
        if (c.status != kSolved) {
            s->st.unconverged += 1;
            if (s->st.unconverged == 1)
                std::fprintf(stderr, "warning: CG did not converge on a split-node solve after %d iterations\n", c.iters);
        }
        std::vector<double> xr(n);
        HIPCHK(hipMemcpy(xr.data(), w.x.p, n * 8, hipMemcpyDeviceToHost));
        embed_full(op, xr.data(), u);
    });
}

eqd_status eqd_shear_traction(eqd_solver* s, const double* s2, const double* s3, double* dtau2, double* dtau3) {
    return guard([&] {
        if (!s || !s2 || !s3 || !dtau2 || !dtau3) throw Error(EQD_ERR_ARG, "null argument");
        const eqd_op& op = *s->op;
        const int n23 = op.g.n23;
        std::vector<double> f(2 * (size_t)n23 * n23);
        slip_fields(op, s2, s3, f.data());
        DevMem out((size_t)2 * op.nf * 8);
        std::vector<int64_t> cid{0}, opos{0};
        eqd_cg_opts saved = s->o;
        if (s->o.engine != EQD_ENGINE_PLAIN) s->o.tile_k = 1;
        try {
            run_columns(s, cid, opos,
                        [&](Workspace& w, int, int) {
                            HIPCHK(hipMemcpy(w.fields.p, f.data(), f.size() * 8, hipMemcpyHostToDevice));
                        },
                        out.as<double>(), 2 * op.nf);
        } catch (...) {
            s->o = saved;
            throw;
        }
        s->o = saved;
        std::vector<double> h((size_t)2 * op.nf);
        HIPCHK(hipMemcpy(h.data(), out.p, h.size() * 8, hipMemcpyDeviceToHost));
        std::memcpy(dtau2, h.data(), op.nf * 8);
        std::memcpy(dtau3, h.data() + op.nf, op.nf * 8);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Allocation-free forward solve. The Krylov workspace of a one-column solve lives in a static
// device array (one copy per device, held until the process ends) instead of the solver's
// hipMalloc'd workspace; the iterations are launched directly (no graph, no events, no pinned
// host memory). Calls that overlap on one device share the array and race.

constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The array's address on each device: the host-side name of a __device__ variable is a
// stand-in, so every kernel argument and copy uses the address looked up here.
void* scratch_ptr() {
    static std::mutex mu;
    static std::map<int, void*> addr;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(EQD_ERR_DEVICE, "hipGetSymbolAddress returned a null address for the scratch array");
    addr[dev] = p;
    return p;
}

__global__ void k_zero_words(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// Carves 256-byte-aligned pieces off the scratch array, in order (dry: only measures).
struct Carver {
    unsigned char* base;
    bool dry;
    size_t off = 0;
    void take(DevMem& m, size_t n) {
        off = (off + 255) & ~(size_t)255;
        if (!dry) m.borrow(base + off, n);
        off += n;
    }
    size_t aligned() const { return (off + 255) & ~(size_t)255; }
};

// Lays a one-column (K = 1) workspace of the solver's precision out in the scratch array (with
// base == nullptr it only measures) and returns the bytes it spans. `out` receives the traction
// column (2nf doubles); `ctl` spans the control words (partial sums, column state, tile flags,
// ids), which are cleared on every call.
size_t carve_noalloc(const eqd_op& op, int prec, unsigned char* base, Workspace& w, DevMem& out, DevMem& ctl) {
    const size_t nvec = (size_t)op.NN * 3;
    const size_t t = prec == EQD_PREC_MIXED ? sizeof(float) : sizeof(double);
    Carver c{base, base == nullptr};
    c.take(w.b, nvec * 8);
    c.take(w.x, nvec * 8);
    c.take(w.r, nvec * t);
    c.take(w.p, nvec * t);
    c.take(w.q, nvec * t);
    if (prec == EQD_PREC_MIXED) c.take(w.xc, nvec * sizeof(float));
    c.take(w.fields, 2 * (size_t)op.g.n23 * op.g.n23 * 8);
    c.take(out, 2 * (size_t)op.nf * 8);
    const size_t lo = c.aligned();
    c.take(w.partial, (size_t)op.G * 8);
    if (op.nfold) c.take(w.folded, (size_t)op.nfold * 8);
    c.take(w.cs, sizeof(ColState));
    c.take(w.tile_on, sizeof(int));
    c.take(w.tile_rel, sizeof(int));
    c.take(w.colid, 8);
    c.take(w.outpos, 8);
    const size_t hi = c.aligned();
    if (!c.dry) ctl.borrow(base + lo, hi - lo);
    return hi;
}

}  // namespace

extern "C" {

eqd_status eqd_shear_traction_noalloc(eqd_solver* s, const double* s2, const double* s3, double* dtau2,
                                      double* dtau3) {
    return guard([&] {
        if (!s || !s2 || !s3 || !dtau2 || !dtau3) throw Error(EQD_ERR_ARG, "null argument");
        if (s->o.engine != EQD_ENGINE_BATCHED)
            throw Error(EQD_ERR_ARG, "the allocation-free forward solve runs the batched engine (engine = BATCHED)");
        const eqd_op& op = *s->op;
        const int prec = s->o.precision;
        std::unique_ptr<Workspace> own(new Workspace());
        Workspace& w = *own;
        DevMem out, ctl;
        const size_t need = carve_noalloc(op, prec, nullptr, w, out, ctl);
        if (need > kScratchBytes)
            throw Error(EQD_ERR_RANGE, "the allocation-free forward solve needs " + std::to_string(need) +
                                           " B of device scratch for this grid; its static scratch holds " +
                                           std::to_string(kScratchBytes) + " B (64 MiB, about 1.6M stored unknowns "
                                           "in FP64); use eqd_shear_traction");
        carve_noalloc(op, prec, static_cast<unsigned char*>(scratch_ptr()), w, out, ctl);
        w.K = 1;
        w.ntiles = 1;
        w.prec = prec;
        w.nvec = (size_t)op.NN * 3;
        hipStream_t st = s->stream;
        const int n23 = op.g.n23;
        std::vector<double> f(2 * (size_t)n23 * n23);
        slip_fields(op, s2, s3, f.data());
        hipLaunchKernelGGL(k_zero_words, dim3(64), dim3(256), 0, st, ctl.as<uint32_t>(), ctl.bytes / 4);
        HIPCHK(hipGetLastError());
        const int64_t cid = 0, opos = 0;
        HIPCHK(hipMemcpyAsync(w.colid.p, &cid, 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(w.outpos.p, &opos, 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(w.fields.p, f.data(), f.size() * 8, hipMemcpyHostToDevice, st));
        // the solver's own workspace is set aside for the call and put back afterwards
        own.swap(s->ws);
        struct Restore {
            eqd_solver* s;
            std::unique_ptr<Workspace>& own;
            ~Restore() { own.swap(s->ws); }
        } restore{s, own};
        with_k(1, [&](auto d) {
            decltype(d)::rhs(s);
            decltype(d)::init(s);
            if (prec == EQD_PREC_MIXED) Engine<float, 1>::solve_direct(s);
            else Engine<double, 1>::solve_direct(s);
            decltype(d)::extract(s, out.as<double>(), 2 * op.nf);
        });
        ColState c;
        HIPCHK(hipMemcpyAsync(&c, s->ws->cs.p, sizeof(c), hipMemcpyDeviceToHost, st));
        std::vector<double> h((size_t)2 * op.nf);
        HIPCHK(hipMemcpyAsync(h.data(), out.p, h.size() * 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        s->st.solves += 1;
        s->st.iterations += c.iters;
        s->st.max_iterations = std::max<int64_t>(s->st.max_iterations, c.iters);
        if (c.status != kSolved) {
            s->st.unconverged += 1;
            if (s->st.unconverged == 1)
                std::fprintf(stderr, "warning: CG did not converge on a split-node solve after %d iterations\n", c.iters);
        }
        std::memcpy(dtau2, h.data(), op.nf * 8);
        std::memcpy(dtau3, h.data() + op.nf, op.nf * 8);
    });
}

eqd_status eqd_fault_stiffness_device(eqd_solver* s, int mode, const int* cols, int ncols, double* K_dev, int64_t ldK) {
    return guard([&] {
        if (!s || !K_dev) throw Error(EQD_ERR_ARG, "null argument");
        stiffness(s, mode, cols, ncols, K_dev, ldK);
    });
}

eqd_status eqd_fault_stiffness(eqd_solver* s, int mode, const int* cols, int ncols, double* K, int64_t ldK) {
    return guard([&] {
        if (!s || !K) throw Error(EQD_ERR_ARG, "null argument");
        const int nf = s->op->nf;
        const int64_t nout = (mode == EQD_STIFFNESS_EXACT && cols) ? ncols : 2 * (int64_t)nf;
        if (ldK < 2 * nf) throw Error(EQD_ERR_ARG, "ldK must be >= 2nf");
        DevMem dK((size_t)std::max<int64_t>(nout, 1) * ldK * 8);
        stiffness(s, mode, cols, ncols, dK.as<double>(), ldK);
        if (nout > 0) HIPCHK(hipMemcpy(K, dK.p, (size_t)nout * ldK * 8, hipMemcpyDeviceToHost));
    });
}

eqd_status eqd_server_create(const eqd_cg_opts* opts, size_t hbm_budget_bytes, eqd_server** out) {
    return guard([&] {
        if (!out) throw Error(EQD_ERR_ARG, "null argument");
        check_device();
        auto sv = std::make_unique<eqd_server>();
        if (opts) sv->o = *opts;
        else eqd_cg_opts_default(&sv->o);
        if (sv->o.engine != EQD_ENGINE_BATCHED) throw Error(EQD_ERR_ARG, "the serving loop runs the batched engine");
        sv->o.tile_k = 64;
        sv->o.timing = 0;
        sv->budget = hbm_budget_bytes;
        if (sv->budget == 0) {
            size_t freeb = 0, totb = 0;
            HIPCHK(hipMemGetInfo(&freeb, &totb));
            sv->budget = freeb / 2;
        }
        *out = sv.release();
    });
}

eqd_status eqd_server_destroy(eqd_server* sv) {
    return guard([&] { delete sv; });
}

eqd_status eqd_server_add_grid(eqd_server* sv, const eqd_grid_desc* g, int max_columns, int* grid_id, int* slots) {
    return guard([&] {
        if (!sv || !g) throw Error(EQD_ERR_ARG, "null argument");
        auto G = std::make_unique<ServeGrid>();
        eqd_status rc = eqd_operator_create(g, &G->op);
        if (rc != EQD_OK) throw Error(rc, g_err);
        rc = eqd_solver_create(G->op, &sv->o, &G->s);
        if (rc != EQD_OK) throw Error(rc, g_err);
        const size_t per = bytes_per_column(*G->op, sv->o.precision) + 2 * (size_t)G->op->nf * 8;
        const size_t room = sv->budget > sv->used ? sv->budget - sv->used : 0;
        int tiles = (int)std::min<size_t>(room / (per * 64), 1 << 20);
        if (max_columns > 0) tiles = std::min(tiles, (max_columns + 63) / 64);
        if (tiles < 1) throw Error(EQD_ERR_NOMEM, "the HBM budget left does not hold one 64-column tile of this grid");
        ensure_ws(G->s, 64, tiles, sv->o.precision);
        G->ntiles = tiles;
        G->slots = tiles * 64;
        sv->used += (size_t)G->slots * per;
        G->slot_req.assign(G->slots, -1);
        G->slot_idx.assign(G->slots, 0);
        G->colid.assign(G->slots, -1);
        G->hcs.resize(G->slots);
        G->arena.alloc((size_t)G->slots * 2 * G->op->nf * 8);
        G->mask_fresh.alloc(G->slots * sizeof(int));
        G->mask_unit.alloc(G->slots * sizeof(int));
        G->mask_retire.alloc(G->slots * sizeof(int));
        HIPCHK(hipEventCreateWithFlags(&G->ev, hipEventDisableTiming));
        Workspace& w = *G->s->ws;
        std::vector<int64_t> opos(G->slots);
        for (int i = 0; i < G->slots; ++i) opos[i] = i;  // each lane extracts into its arena column
        HIPCHK(hipMemcpy(w.outpos.p, opos.data(), G->slots * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(w.colid.p, G->colid.data(), G->slots * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(w.fields.p, 0, w.fields.bytes));
        HIPCHK(hipMemset(w.b.p, 0, w.b.bytes));
        with_k(64, [&](auto d) { decltype(d)::init(G->s); });  // every lane idle, every tile off
        HIPCHK(hipStreamSynchronize(G->s->stream));
        if (grid_id) *grid_id = (int)sv->grids.size();
        if (slots) *slots = G->slots;
        sv->grids.push_back(std::move(G));
    });
}

eqd_status eqd_server_submit_columns(eqd_server* sv, int grid, const int* cols, int ncols, int64_t* req_id) {
    return guard([&] {
        if (!sv || !cols || ncols < 1) throw Error(EQD_ERR_ARG, "a request needs at least one column");
        serve_check_grid(sv, grid);
        ServeGrid& G = *sv->grids[grid];
        if (ncols > G.slots) throw Error(EQD_ERR_ARG, "request larger than the grid's lane pool");
        const int n2 = 2 * G.op->nf;
        ServeRequest r;
        r.grid = grid;
        r.kind = 0;
        r.ncols = r.pending = ncols;
        for (int j = 0; j < ncols; ++j) {
            if (cols[j] < 0 || cols[j] >= n2) throw Error(EQD_ERR_ARG, "column index out of range");
            r.cols.push_back(cols[j]);
        }
        r.result.assign((size_t)n2 * ncols, 0.0);
        r.iters.assign(ncols, 0);
        const int64_t id = sv->next_id++;
        sv->reqs.emplace(id, std::move(r));
        G.queue.push_back(id);
        if (req_id) *req_id = id;
    });
}

eqd_status eqd_server_submit_forward(eqd_server* sv, int grid, const double* s2, const double* s3, int64_t* req_id) {
    return guard([&] {
        if (!sv || !s2 || !s3) throw Error(EQD_ERR_ARG, "null argument");
        serve_check_grid(sv, grid);
        ServeGrid& G = *sv->grids[grid];
        const int n2 = 2 * G.op->nf;
        ServeRequest r;
        r.grid = grid;
        r.kind = 1;
        r.ncols = r.pending = 1;
        r.fields.assign(2 * (size_t)G.op->g.n23 * G.op->g.n23, 0.0);
        slip_fields(*G.op, s2, s3, r.fields.data());
        r.result.assign(n2, 0.0);
        r.iters.assign(1, 0);
        const int64_t id = sv->next_id++;
        sv->reqs.emplace(id, std::move(r));
        G.queue.push_back(id);
        if (req_id) *req_id = id;
    });
}

eqd_status eqd_server_step(eqd_server* sv, int64_t* lanes_busy, int64_t* queued) {
    return guard([&] {
        if (!sv) throw Error(EQD_ERR_ARG, "null server");
        int64_t busy = 0, q = 0;
        for (auto& G : sv->grids) {
            busy += serve_grid_step(sv, *G);
            q += (int64_t)G->queue.size();
        }
        if (lanes_busy) *lanes_busy = busy;
        if (queued) *queued = q;
    });
}

eqd_status eqd_server_completed(eqd_server* sv, int64_t* ids, int max, int* n) {
    return guard([&] {
        if (!sv || !n) throw Error(EQD_ERR_ARG, "null argument");
        int k = 0;
        while (k < max && !sv->completed.empty()) {
            if (ids) ids[k] = sv->completed.front();
            sv->completed.pop_front();
            ++k;
        }
        *n = k;
    });
}

eqd_status eqd_server_result(eqd_server* sv, int64_t id, double* out, int64_t ld, int* iters, int* unconverged) {
    return guard([&] {
        if (!sv) throw Error(EQD_ERR_ARG, "null server");
        auto it = sv->reqs.find(id);
        if (it == sv->reqs.end()) throw Error(EQD_ERR_ARG, "unknown request id");
        ServeRequest& r = it->second;
        if (!r.done) throw Error(EQD_ERR_ARG, "request not completed");
        const int n2 = 2 * sv->grids[r.grid]->op->nf;
        if (out) {
            if (ld < n2) throw Error(EQD_ERR_ARG, "ld must be at least 2nf");
            for (int j = 0; j < r.ncols; ++j)
                std::memcpy(out + (size_t)j * ld, r.result.data() + (size_t)j * n2, n2 * 8);
        }
        if (iters) std::copy(r.iters.begin(), r.iters.end(), iters);
        if (unconverged) *unconverged = r.unconverged;
        sv->reqs.erase(it);
    });
}

eqd_status eqd_server_grid_info(const eqd_server* sv, int grid, eqd_op_info* info, int* slots, double* hbm_bytes) {
    return guard([&] {
        if (!sv) throw Error(EQD_ERR_ARG, "null server");
        serve_check_grid(sv, grid);
        const ServeGrid& G = *sv->grids[grid];
        if (info) {
            const eqd_status rc = eqd_operator_info(G.op, info);
            if (rc != EQD_OK) throw Error(rc, g_err);
        }
        if (slots) *slots = G.slots;
        if (hbm_bytes)
            *hbm_bytes = (double)G.slots * (double)(bytes_per_column(*G.op, sv->o.precision) + 2 * (size_t)G.op->nf * 8);
    });
}

}  // extern "C"
