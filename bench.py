"""Benchmark: exact fault-stiffness K build at dz = 50 m, L_fault = 800, L_normal = 400.

One step = one full :exact K build (2178 unit-slip columns, l_f = L_fault = 800 m,
58,806 DOF) by the mixed-precision batched CG (FP32 stencil/CG inside FP64 reliable
updates, rtol 1e-10), with the operator resident on the device and K left in device
memory. Prints one JSON line (see the keys below).

Besides the timed steps it measures, on the same GPU: the plain one-RHS FP64 CG
(host alpha/beta, no graphs) on a column sample, the batched FP64 build, K's error
against the batched FP64 K and against the CPU oracle on sampled columns, the
stencil kernel's average launch duration (HIP events on the solver's stream) for the
roofline, the CPU oracle's threaded build on a column sample, and the BP8 time loop
(30 days of BP8-QD-GS at dz = 50 m with a device-built Toeplitz K) beside the CPU oracle's
integration of the same K, and the serving loop under a seeded Poisson request stream
(p50/p99 latency, columns/s).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "earthquakediffinitive.jl_amd")]

import eqd  # noqa: E402
from eqd import _lib  # noqa: E402

HBM_PEAK = 8.0e12


class DeviceBuffer:
    """A raw device allocation (hipMalloc) for K kept on the GPU."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        rc = self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed: {rc}")
        self.nbytes = nbytes

    def to_host(self, shape):
        out = np.zeros(shape, order="F")
        rc = self.hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, ctypes.c_size_t(self.nbytes), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None


def build_device(solver, Kbuf, nf2):
    _lib.check(_lib.lib().eqd_fault_stiffness_device(solver.handle, _lib.STIFFNESS_EXACT, None, 0, Kbuf.ptr, nf2))


def copy_calibration(mib):
    """A plain device-to-device copy of known size (mib MiB read and written per launch) for the
    counter calibration: hipMemcpy between two device buffers runs as one copy kernel on the
    device (the runtime's buffer-copy blit), driven here through the HIP runtime directly."""
    n = mib * 2**20
    src, dst = DeviceBuffer(n), DeviceBuffer(n)
    hip = src.hip
    if hip.hipMemset(src.ptr, ctypes.c_int(1), ctypes.c_size_t(n)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(3):
        if hip.hipMemcpy(dst.ptr, src.ptr, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    if hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")
    src.free()
    dst.free()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_read_per_launch": n, "bytes_written_per_launch": n}))


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE, calibrated against a plain copy), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "traffic_dz50_exact_build.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return int(parts[2])
    return None


def bench_bp8(args):
    """BP8-QD-GS for 30 days at dz = 50 m, L = (800, 400), l_f = 400 (run_bp8.jl's defaults:
    Toeplitz K, saveat 300 s) with K built and kept on the device; the CPU oracle integrates the
    same K beside it."""
    from eqd import bp8
    t0 = time.perf_counter()
    m = bp8.build_model(dz=50.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    t_build = time.perf_counter() - t0
    tf, saveat = 30 * 86400.0, 300.0
    t0 = time.perf_counter()
    sol = bp8.run_bp8(m, tspan=(0.0, tf), saveat=saveat)
    t_run = time.perf_counter() - t0
    t0 = time.perf_counter()
    vmax = bp8.saved_vmax(m, len(sol.t))
    t_vmax = time.perf_counter() - t0
    nf = m.nf
    ic = (m.n2f // 2) * (m.n2f + 1)
    j = int(np.argmax(vmax))
    res = {"nf": nf, "state_length": m.state_len, "K_build_ms": round(t_build * 1e3, 1),
           "run_ms": round(t_run * 1e3, 1), "device_loop_ms": round(sol.ms, 1), "accepted_steps": sol.naccept,
           "rejected_steps": sol.nreject, "rhs_evaluations": sol.nrhs, "saved": len(sol.t),
           "us_per_rhs": round(sol.ms * 1e3 / max(1, sol.nrhs), 2), "saved_vmax_ms": round(t_vmax * 1e3, 1),
           "peak_V": float(vmax[j]), "peak_V_day": round(float(sol.t[j]) / 86400, 3),
           "final_slip_centre": float(sol.u[-1, ic]), "peak_p_centre_MPa": float(sol.u[:, 3 * nf + ic].max() / 1e6)}
    print(f"bench: bp8 30 d: K build {t_build * 1e3:.0f} ms, run {t_run * 1e3:.0f} ms "
          f"({sol.naccept} steps, {sol.nrhs} rhs), peak V {vmax[j]:.3e}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        from oracle.oracle import OracleBP8
        o = OracleBP8(m.K, m.n2f, dict(dz=50.0))
        t0 = time.perf_counter()
        ro = o.run(0.0, tf, saveat)
        t_cpu = time.perf_counter() - t0
        res["cpu_oracle_run_ms"] = round(t_cpu * 1e3, 1)
        res["cpu_oracle_accepted_steps"] = ro["naccept"]
        d = np.abs(vmax - ro["vmax"]) / ro["vmax"]
        res["vmax_max_rel_diff_vs_oracle"] = float(d.max())
        res["vmax_median_rel_diff_vs_oracle"] = float(np.median(d))
        res["vmax_max_rel_diff_day"] = round(float(sol.t[int(np.argmax(d))]) / 86400, 2)
        # the trajectory's own conditioning: the oracle against itself with K scaled by a few ulps
        worst, day = 0.0, 0.0
        for scale in (1 + 1e-15, 1 - 1e-15, 1 + 4e-15, 1 - 4e-15):
            rp = OracleBP8(m.K * scale, m.n2f, dict(dz=50.0)).run(0.0, tf, saveat)
            dp = np.abs(rp["vmax"] - ro["vmax"]) / ro["vmax"]
            if dp.max() > worst:
                worst, day = float(dp.max()), round(float(sol.t[int(np.argmax(dp))]) / 86400, 2)
        res["oracle_vs_itself_K_scaled_1pm1e-15_4e-15"] = {"vmax_max_rel_diff": worst, "day": day}
        print(f"bench: bp8 30 d cpu oracle {t_cpu * 1e3:.0f} ms ({ro['naccept']} steps)", file=sys.stderr,
              flush=True)
    return res


def host_info():
    """nproc and the CPU model of the host the CPU baseline ran on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def bench_lf400(args, lam, mu):
    """The headline grid (dz = 50 m, L = (800, 400)) in the reference's own Omega_f reading, l_f = 400:
    289 fault nodes, 578 columns, the configuration of the reference's published CPU build times
    (201.7 s serial = 2.87 K-columns/s, 94.5 s on 16 threads = 6.12 K-columns/s; Julia on a 12-core
    workstation, PROGRESS.md:370-376 of the reference). The device build here (mixed precision, the
    headline's settings) beside the CPU restatement on this host's cores."""
    g = eqd.bp8_grid_pair(args.dz, args.L_fault, args.L_normal)
    fe = eqd.FaultElasticity(g, lam, mu, l_f=400.0, precision="mixed", tile_k=args.tile_k,
                             reliable_interval=args.reliable_interval)
    nf2 = 2 * eqd.frictional_node_count(fe)
    K = eqd.fault_stiffness(fe)  # untimed: autotune and first graph capture
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        K = eqd.fault_stiffness(fe)
        ts.append(time.perf_counter() - t0)
    t = min(ts)
    res = {"columns": nf2, "device_K_columns_per_s": round(nf2 / t, 1), "device_build_ms": round(t * 1e3, 2),
           "published_reference_cpu": {"one_thread": 2.87, "16_threads": 6.12, "unit": "K-columns/s",
                                       "note": "Julia reference, 12-core workstation, PROGRESS.md:370-376"},
           "device_vs_published_16_threads": round(nf2 / t / 6.12, 1),
           "K_self": float(K[(nf2 // 2) // 2, (nf2 // 2) // 2])}
    if not args.no_cpu:
        from oracle.oracle import OracleSystem, bp8_geometry
        o = OracleSystem(bp8_geometry(args.dz, args.L_fault, args.L_normal, 400.0))
        ncpu = min(os.cpu_count() or 1, 16)
        t0 = time.perf_counter()
        Ko = o.fault_stiffness(nthreads=ncpu)
        tc = time.perf_counter() - t0
        res["cpu_port"] = {"value": round(nf2 / tc, 2), "threads": ncpu, "columns": nf2}
        res["K_rel_err_vs_oracle"] = float(np.linalg.norm(K - Ko) / np.linalg.norm(Ko))
    print(f"bench: l_f=400 reading: {nf2} cols on the device in {t * 1e3:.1f} ms "
          f"({nf2 / t:.0f} K-cols/s, {nf2 / t / 6.12:.0f}x the published 16-thread Julia build)",
          file=sys.stderr, flush=True)
    return res


def bench_plain_config(args):
    """The CPU-runnable configuration end to end: BP8-QD-GS at dz = 100 m with l_f = L_fault = 800 m
    (289 fault nodes, 578 columns). As written (L_normal = 400) its n1 = 5 < 9 is rejected by
    build_model (src/BP8.jl:199-202), so the legal L_normal = 800 is used (15,606 DOF). The exact K
    comes from the FP64 batched CG on the device, then 30 days of Tsit5 (saveat 3600 s) on the
    device; the CPU oracle builds the same K (assembled CSR, one column per thread) and integrates
    it beside."""
    from eqd import bp8
    par = bp8.benchmark_parameters(l_f=800.0, dz=100.0)
    t0 = time.perf_counter()
    m = bp8.build_model(par=par, dz=100.0, L_fault=800.0, L_normal=800.0, stiffness="exact", precision="fp64")
    t_build = time.perf_counter() - t0
    tf, saveat = 30 * 86400.0, 3600.0
    t0 = time.perf_counter()
    sol = bp8.run_bp8(m, tspan=(0.0, tf), saveat=saveat)
    t_run = time.perf_counter() - t0
    vmax = bp8.saved_vmax(m, len(sol.t))
    K = m.K
    res = {"dof": m.grid_info["elastic_dofs"], "columns": 2 * m.nf, "K_build_ms": round(t_build * 1e3, 1),
           "run_ms": round(t_run * 1e3, 1), "device_loop_ms": round(sol.ms, 1), "accepted_steps": sol.naccept,
           "peak_V": float(vmax.max())}
    if not args.no_cpu:
        from oracle.oracle import OracleBP8, OracleSystem, bp8_geometry
        ncpu = min(os.cpu_count() or 1, 16)
        o = OracleSystem(bp8_geometry(100.0, 800.0, 800.0, 800.0))
        t0 = time.perf_counter()
        Ko = o.fault_stiffness(nthreads=ncpu)
        res["cpu_K_build_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        res["cpu_threads"] = ncpu
        res["K_rel_err_vs_oracle"] = float(np.linalg.norm(K - Ko) / np.linalg.norm(Ko))
        ob = OracleBP8(Ko, m.n2f, dict(dz=100.0, l_f=800.0))
        t0 = time.perf_counter()
        ro = ob.run(0.0, tf, saveat)
        res["cpu_run_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        res["cpu_accepted_steps"] = ro["naccept"]
        res["vmax_max_rel_diff_vs_oracle"] = float(np.max(np.abs(vmax - ro["vmax"]) / ro["vmax"]))
    print(f"bench: dz=100 plain config: K {t_build * 1e3:.0f} ms, 30 d {t_run * 1e3:.0f} ms, "
          f"V_max vs oracle {res.get('vmax_max_rel_diff_vs_oracle', float('nan')):.2e}", file=sys.stderr, flush=True)
    return res


def bench_production(args):
    """Production-size stiffness build: dz = 25 m, L_fault = 1600, L_normal = 800, l_f = 1600
    (3,294,918 DOF, 16641 fault nodes, 33282 columns). The :toeplitz build (10 solves, K filled
    on the device, 8.9 GB) and a seeded 512-column subset of the :exact build (mixed precision);
    the Toeplitz columns are checked against the exact ones (< 1e-3 relative, the reference's own
    Toeplitz guard scale), and mixed against batched FP64 on 8 of the columns (1e-8)."""
    mu, nu = 32.04e9, 0.25
    lam = 2 * mu * nu / (1 - 2 * nu)
    g = eqd.bp8_grid_pair(25.0, 1600.0, 800.0)
    fe = eqd.FaultElasticity(g, lam, mu, l_f=1600.0, precision="mixed", tile_k=64)
    nf2 = 2 * eqd.frictional_node_count(fe)
    rng = np.random.default_rng(20261203)
    cols = np.sort(rng.choice(nf2, size=args.prod_cols, replace=False)).astype(np.int32)
    res = {"dof": fe.Ntot, "columns_total": nf2, "exact_subset": len(cols)}
    Kt = DeviceBuffer(nf2 * nf2 * 8)
    t0 = time.perf_counter()
    _lib.check(_lib.lib().eqd_fault_stiffness_device(fe.rs.handle, _lib.STIFFNESS_TOEPLITZ, None, 0, Kt.ptr, nf2))
    res["toeplitz_build_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
    # the Toeplitz K's columns of the subset (device -> host, column by column)
    Ktc = np.zeros((nf2, len(cols)), order="F")
    for j, c in enumerate(cols):
        src = ctypes.c_void_p(Kt.ptr.value + int(c) * nf2 * 8)
        rc = Kt.hip.hipMemcpy(ctypes.c_void_p(Ktc[:, j].ctypes.data), src, ctypes.c_size_t(nf2 * 8), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
    Kt.free()
    Ke = DeviceBuffer(nf2 * len(cols) * 8)
    t0 = time.perf_counter()
    _lib.check(_lib.lib().eqd_fault_stiffness_device(fe.rs.handle, _lib.STIFFNESS_EXACT, cols.ctypes.data, len(cols),
                                                     Ke.ptr, nf2))
    te = time.perf_counter() - t0
    K_ex = Ke.to_host((nf2, len(cols)))
    Ke.free()
    rep = eqd.solver_report(fe.rs)
    res["exact_subset_ms"] = round(te * 1e3, 1)
    res["exact_K_columns_per_s"] = round(len(cols) / te, 2)
    res["exact_mean_cg_iterations"] = round(rep["iterations"] / max(1, rep["solves"]), 1)
    res["exact_unconverged"] = rep["unconverged"]
    col_err = np.linalg.norm(Ktc - K_ex, axis=0) / np.linalg.norm(K_ex, axis=0)
    res["toeplitz_vs_exact_max_col_rel_err"] = float(col_err.max())
    res["toeplitz_vs_exact_median_col_rel_err"] = float(np.median(col_err))
    res["toeplitz_cols_within_1e-3"] = round(float(np.mean(col_err < 1e-3)), 4)
    res["toeplitz_vs_exact_frobenius_rel_err"] = float(np.linalg.norm(Ktc - K_ex) / np.linalg.norm(K_ex))
    fe64 = eqd.FaultElasticity(g, lam, mu, l_f=1600.0, precision="fp64", tile_k=8)
    sub = cols[:8]
    K64 = eqd.fault_stiffness(fe64, cols=sub)
    res["mixed_vs_fp64_rel_err_8cols"] = float(np.linalg.norm(K_ex[:, :8] - K64) / np.linalg.norm(K64))
    print(f"bench: dz=25 production: toeplitz {res['toeplitz_build_ms']} ms, exact {len(cols)} cols in "
          f"{res['exact_subset_ms']} ms, toeplitz vs exact max col err {col_err.max():.2e}", file=sys.stderr, flush=True)
    return res


def bench_ensemble(args):
    """Trajectory ensemble: dz = 20 m, L_fault = 800, L_normal = 400, l_f = 800 (6561 fault nodes, K
    13122 x 13122 FP64 = 1.38 GB from the device :toeplitz build), 256 members with seeded
    perturbations of Q0 (+-20%), a - b (+-10%) and sigma0 (+-5%), Gaussian source, integrated over
    args.ens_hours (default 30 days) with saveat 3600 s. K S per stage is one 2D FFT convolution
    batched over the members (transform 256 x 256; its roofline is HBM, by the bytes of the
    transform planes it writes and reads back). The Peaceman-well members run on a model of the same
    K handed over dense, whose K S is one FP64 MFMA GEMM (2 * 13122^2 * 256 = 88.2 GFLOP, roofline
    the dense FP64 matrix peak): both products are timed."""
    from eqd import bp8
    from eqd.ensemble import Ensemble, perturbed_members
    par = bp8.benchmark_parameters(l_f=800.0, dz=20.0)
    t0 = time.perf_counter()
    m = bp8.build_model(par=par, dz=20.0, L_fault=800.0, L_normal=400.0, stiffness="toeplitz", precision="mixed")
    t_build = time.perf_counter() - t0
    # one trajectory over the same horizon: K s from the Toeplitz separation kernel (build_model's
    # model) and from the dense 1.38 GB K (the same K handed to model_from_stiffness)
    single = {}
    Kh = m.K
    ks_us = {}
    for meth in ("direct", "fft"):
        m.set_ks_method(meth)
        ks_us[meth] = round(m.ks_timing(200) * 1e3, 2)
    for name, mm, meth in (("toeplitz_fft", m, "fft"), ("toeplitz_direct", m, "direct"),
                           ("dense_K", bp8.model_from_stiffness(Kh, m.n2f, par=m.par), None)):
        if meth:
            mm.set_ks_method(meth)
        sol = bp8.run_bp8(mm, tspan=(0.0, args.ens_hours * 3600.0), saveat=3600.0)
        single[name] = {"device_loop_ms": round(sol.ms, 1), "rhs_evaluations": sol.nrhs,
                        "us_per_rhs": round(sol.ms * 1e3 / max(1, sol.nrhs), 2), "accepted_steps": sol.naccept,
                        "final_slip_centre": float(sol.u[-1, (m.n2f // 2) * (m.n2f + 1)])}
        del mm
    m.set_ks_method("fft")
    single["ks_us_per_product"] = ks_us
    mem = perturbed_members(m.par, args.ens_members, seed=20261203)
    e = Ensemble(m, args.ens_members, **mem)
    t0 = time.perf_counter()
    st = e.run(tspan=(0.0, args.ens_hours * 3600.0), saveat=3600.0)
    t_run = time.perf_counter() - t0
    print(f"bench: ensemble GS {args.ens_members} members {args.ens_hours} h: loop {e.loop_ms:.0f} ms",
          file=sys.stderr, flush=True)
    cpu_check = None
    if not getattr(args, "no_cpu", True):
        # member-wise against the CPU oracle at full size: members 0 (base) and 1 over the first
        # 2 h (the oracle's dense K product costs ~0.15 s per rhs! at nf = 6561)
        from oracle.oracle import OracleBP8
        ens_v, _ = e.saved_vmax(3)
        worst = 0.0
        t0 = time.perf_counter()
        for j in (0, 1):
            par_j = dict(dz=20.0, l_f=800.0, Q0=float(mem["Q0"][j]), a=float(mem["a"][j]),
                         sigma0=float(mem["sigma0"][j]))
            ro = OracleBP8(Kh, m.n2f, par_j).run(0.0, 7200.0, 3600.0)
            worst = max(worst, float(np.max(np.abs(ens_v[j] - ro["vmax"]) / ro["vmax"])))
        cpu_check = {"members": [0, 1], "hours": 2, "vmax_max_rel_diff_vs_oracle": worst,
                     "cpu_ms": round((time.perf_counter() - t0) * 1e3, 1)}
    ms_ks, _ = e.gemm_timing(20)
    # the transform planes (rows x M complex, written and read back twice) plus the slips read and
    # the tractions written, per member
    Mt = 1 << (2 * m.n2f - 2).bit_length()
    plane_bytes = args.ens_members * (4 * m.n2f * Mt * 16 + 2 * 2 * m.nf * 8)
    fft_bw = plane_bytes / (ms_ks * 1e-3)
    acc = [s.naccept for s in st]
    ens_loop_ms = e.loop_ms
    del e
    # the same members with the Peaceman well (BP8-QD-PW) over a shorter horizon: its steps grow
    # as the well pressure builds (104 in the first 3 h, ~1e6 per 100 h at this spacing by the
    # reference's dz^-4 scaling of its 31,249 steps per 100 h at dz = 50 m)
    pw = None
    gemm = None
    if args.ens_pw_hours > 0:
        mpw = bp8.model_from_stiffness(Kh, m.n2f, par=m.par, injection="peaceman")
        epw = Ensemble(mpw, args.ens_members, **mem)
        print(f"bench: ensemble PW {args.ens_members} members, {args.ens_pw_hours} h ...", file=sys.stderr,
              flush=True)
        t0 = time.perf_counter()
        # a step budget per member: PW steps grow steeply with the horizon, so a long horizon ends
        # with a MaxIters retcode (counted below) instead of an unbounded run
        stp = epw.run(tspan=(0.0, args.ens_pw_hours * 3600.0), saveat=600.0, maxiters=args.ens_pw_maxiters)
        t_pw = time.perf_counter() - t0
        accp = [s.naccept for s in stp]
        ms_gemm, flops = epw.gemm_timing(20)
        tf = flops / (ms_gemm * 1e-3)
        gemm = {"bound": "mfma", "achieved": round(tf / 1e12, 2), "peak": 78.6, "unit": "TFLOP/s",
                "frac": round(tf / 78.6e12, 4), "ms_per_gemm": round(ms_gemm, 3), "flop_per_gemm": flops}
        pw = {"hours": args.ens_pw_hours, "run_ms": round(t_pw * 1e3, 1), "device_loop_ms": round(epw.loop_ms, 1),
              "accepted_steps_min_max": [int(min(accp)), int(max(accp))],
              "member_steps_per_s": round(sum(accp) / (epw.loop_ms * 1e-3), 1),
              "maxiters": args.ens_pw_maxiters, "members_at_maxiters": int(sum(s.retcode != 0 for s in stp)),
              "retcodes_ok": bool(all(s.retcode == 0 for s in stp))}
        print(f"bench: ensemble PW {args.ens_members} members {args.ens_pw_hours} h: loop {epw.loop_ms:.0f} ms, "
              f"steps {min(accp)}-{max(accp)}", file=sys.stderr, flush=True)
        del epw, mpw
    # BP8-QD-PW over the full 30 days with the implicit-local integrator (per-node 3 x 3 Newton
    # blocks inside L-stable SDIRK4 stages): the same seeded members in lockstep Newton iterations
    # (one batched FFT K S and one rhs! over all members per iteration, each member with its own
    # step size and error control) on a Peaceman model with the Toeplitz K; member 0 also alone
    # (eqd_bp8_integrate_imex) for the sequential rate and a trajectory check
    pw_imex = None
    if args.ens_pw_imex_members > 0:
        ne = min(args.ens_pw_imex_members, args.ens_members)
        mpi = bp8.build_model(par=par, dz=20.0, L_fault=800.0, L_normal=400.0, injection="peaceman",
                              stiffness="toeplitz", precision="mixed")
        ei = Ensemble(mpi, ne, **{k: v[:ne] for k, v in mem.items()})
        print(f"bench: ensemble PW imex {ne} members, {args.ens_pw_imex_hours} h ...", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        sti = ei.run(tspan=(0.0, args.ens_pw_imex_hours * 3600.0), saveat=3600.0, alg="imex", maxiters=400000)
        wall = time.perf_counter() - t0
        acci = [s.naccept for s in sti]
        s0 = ei.solution(0)
        p0 = bp8.benchmark_parameters(l_f=800.0, dz=20.0, Q0=float(mem["Q0"][0]), a=float(mem["a"][0]),
                                      sigma0=float(mem["sigma0"][0]))
        m0 = bp8.build_model(par=p0, dz=20.0, L_fault=800.0, L_normal=400.0, injection="peaceman",
                             stiffness="toeplitz", precision="mixed")
        r0 = bp8.run_bp8(m0, tspan=(0.0, args.ens_pw_imex_hours * 3600.0), saveat=3600.0, alg="imex",
                         maxiters=400000)
        nrow = min(len(s0.t), len(r0.t))
        pwell_diff = float(np.max(np.abs(s0.u[:nrow, -1] - r0.u[:nrow, -1]) / np.abs(r0.u[:nrow, -1]).max()))
        pw_imex = {"hours": args.ens_pw_imex_hours, "members": ne, "wall_s": round(wall, 2),
                   "device_loop_ms": round(ei.loop_ms, 1),
                   "accepted_steps_min_max": [int(min(acci)), int(max(acci))],
                   "member_steps_per_s": round(sum(acci) / (ei.loop_ms * 1e-3), 1),
                   "member_hours_per_s": round(ne * args.ens_pw_imex_hours / (ei.loop_ms * 1e-3), 1),
                   "retcodes_ok": bool(all(s.retcode == 0 for s in sti)),
                   "single_member0": {"device_ms": round(r0.ms, 1), "accepted_steps": int(r0.naccept),
                                      "p_well_max_MPa": round(float(r0.u[:, -1].max()) / 1e6, 3),
                                      "ensemble_member0_steps": int(sti[0].naccept),
                                      "p_well_max_rel_diff": pwell_diff},
                   "speedup_vs_sequential_single": round(ne * r0.ms / ei.loop_ms, 1),
                   "note": "eqd_bp8_ensemble_integrate_imex; Tsit5 needs ~1e6 steps per 100 h at dz = 20 m on "
                           "this variant"}
        print(f"bench: ensemble PW imex {ne} members {args.ens_pw_imex_hours} h: loop {ei.loop_ms / 1e3:.2f} s, "
              f"steps {min(acci)}-{max(acci)}; member 0 alone {r0.ms / 1e3:.2f} s", file=sys.stderr, flush=True)
        del ei, mpi, m0
    res = {"nf": m.nf, "members": args.ens_members, "hours": args.ens_hours, "K_build_ms": round(t_build * 1e3, 1),
           "run_ms": round(t_run * 1e3, 1), "device_loop_ms": round(ens_loop_ms, 1),
           "accepted_steps_min_max": [int(min(acc)), int(max(acc))],
           "member_steps_per_s": round(sum(acc) / (ens_loop_ms * 1e-3), 1), "single_trajectory": single,
           "cpu_member_check": cpu_check, "peaceman": pw, "peaceman_imex": pw_imex,
           "ks_fft": {"bound": "hbm", "achieved": round(fft_bw / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                      "frac": round(fft_bw / 8e12, 4), "ms_per_product": round(ms_ks, 4),
                      "transform": [Mt, Mt], "bytes_per_product": plane_bytes,
                      "vs_dense_gemm": round(gemm["ms_per_gemm"] / ms_ks, 2) if gemm else None},
           "gemm_dense_K": gemm}
    print(f"bench: ensemble dz=20 {args.ens_members} members {args.ens_hours} h: K build {t_build * 1e3:.0f} ms, "
          f"loop {ens_loop_ms:.0f} ms, FFT K S {ms_ks:.3f} ms"
          + (f", dense GEMM {gemm['ms_per_gemm']:.3f} ms = {gemm['achieved']:.1f} TFLOP/s" if gemm else ""),
          file=sys.stderr, flush=True)
    return res


def _serve_stream(sv, gids, axes, specs, weights, rate, seconds, seed):
    """One seeded Poisson stream (mean `rate` req/s for `seconds` s) through the server; returns
    latency percentiles, throughput and the backlog the stream leaves when its last request
    arrives."""
    rng = np.random.default_rng(seed)
    n_req = rng.poisson(rate * seconds)
    arrivals = np.sort(rng.uniform(0.0, seconds, n_req))
    reqs = []
    for a in arrivals:
        gi = int(rng.choice(len(specs), p=weights))
        nf = sv.grids[gi]["nf"]
        if rng.random() < 0.5:
            k = int(rng.integers(1, 65))
            reqs.append((a, gi, "columns", rng.choice(2 * nf, size=k, replace=False), None))
        else:
            x = axes[gi]
            X, Y = np.meshgrid(x, x, indexing="ij")
            w = rng.uniform(100.0, 400.0)
            c2, c3 = rng.uniform(x[0], x[-1], 2)
            bump = (1e-3 * np.exp(-((X - c2) ** 2 + (Y - c3) ** 2) / (2 * w * w))).ravel(order="F")
            zero = np.zeros(nf)
            reqs.append((a, gi, "forward", None, (bump, zero) if rng.random() < 0.5 else (zero, bump)))
    pending = {}
    lat = [[] for _ in specs]
    cols_done = 0
    cols_offered = sum(1 if r[3] is None else len(r[3]) for r in reqs)
    backlog_at_last_arrival = None
    t_last_done = 0.0
    i = 0
    start = time.perf_counter()
    t_report = 10.0
    while i < len(reqs) or pending:
        if time.perf_counter() - start > t_report:
            print(f"bench: serving {rate} req/s: {i} submitted, {i - len(pending)} done at {t_report:.0f} s",
                  file=sys.stderr, flush=True)
            t_report += 10.0
        now = time.perf_counter() - start
        while i < len(reqs) and reqs[i][0] <= now:
            a, gi, kind, cols, slip = reqs[i]
            rid = sv.submit_columns(gids[gi], cols) if kind == "columns" else sv.submit_forward(gids[gi], *slip)
            pending[rid] = (a, gi, 1 if cols is None else len(cols))
            i += 1
            if i == len(reqs):
                backlog_at_last_arrival = dict(requests=len(pending), columns=int(sum(v[2] for v in pending.values())))
        busy, queued = sv.step()
        for rid in sv.completed():
            a, gi, nc = pending.pop(rid)
            sv.result(rid, gids[gi], nc)
            t_last_done = time.perf_counter() - start
            lat[gi].append(t_last_done - a)
            cols_done += nc
        if busy == 0 and queued == 0 and i < len(reqs):
            time.sleep(max(0.0, min(reqs[i][0] - (time.perf_counter() - start), 0.002)))
        if time.perf_counter() - start > seconds * 40 + 180:
            raise RuntimeError("serving run did not drain")
    wall = time.perf_counter() - start
    allat = np.concatenate([np.array(l) for l in lat if l]) * 1e3
    info = dict(offered_req_per_s=rate, duration_s=seconds, requests=len(reqs), offered_columns=cols_offered,
                wall_s=round(wall, 2), drain_s=round(wall - (arrivals[-1] if len(arrivals) else 0.0), 2),
                completed_req_per_s=round(len(reqs) / wall, 2), columns=cols_done,
                columns_per_s=round(cols_done / wall, 1), backlog_at_last_arrival=backlog_at_last_arrival,
                p50_ms=round(float(np.percentile(allat, 50)), 1), p99_ms=round(float(np.percentile(allat, 99)), 1),
                mean_ms=round(float(allat.mean()), 1))
    info["by_grid"] = {f"dz{int(d)}_Lf{int(f)}_Ln{int(n)}": (
        dict(requests=len(l), p50_ms=round(float(np.percentile(np.array(l) * 1e3, 50)), 1),
             p99_ms=round(float(np.percentile(np.array(l) * 1e3, 99)), 1)) if l else dict(requests=0))
        for (d, f, n), l in zip(specs, lat)}
    print(f"bench: serving {rate} req/s: {len(reqs)} requests in {wall:.1f} s (drain {info['drain_s']} s): "
          f"p50 {info['p50_ms']} ms, p99 {info['p99_ms']} ms, {info['columns_per_s']} columns/s, "
          f"{info['completed_req_per_s']} req/s completed", file=sys.stderr, flush=True)
    return info


def bench_serving(args):
    """Serving loop under seeded Poisson streams over three grids with l_f = L_fault: (dz,
    L_fault, L_normal) = (100, 800, 800) -- the legal form of (100, 800, 400), whose n1 = 5 < 9 --
    (50, 800, 400) and (25, 1600, 800), weights 0.5 / 0.35 / 0.15. A request is, with equal
    probability, k ~ U{1..64} unit-slip columns at random nodes and directions, or a forward solve
    with a Gaussian slip bump (amplitude 1e-3 m, width U[100, 400] m, centre uniform on Omega_f,
    random direction). Two streams on the same server: args.serve_rate req/s (default 10) for
    args.serve_seconds s, and args.serve_rate_high req/s (default 50, the offered load of the serving
    configuration) for args.serve_seconds_high s. The second exceeds what one device serves on this
    mix; its completed req/s and columns/s are the saturation throughput, and its backlog when the
    last request arrives and its drain time say by how much."""
    from eqd.serve import Server
    mu, nu = 32.04e9, 0.25
    lam = 2 * mu * nu / (1 - 2 * nu)
    specs = [(100.0, 800.0, 800.0), (50.0, 800.0, 400.0), (25.0, 1600.0, 800.0)]
    weights = np.array([0.5, 0.35, 0.15])
    sv = Server(precision="mixed")
    gids, axes = [], []
    t0 = time.perf_counter()
    for dz, Lf, Ln in specs:
        lanes = args.serve_lanes_dz25 if dz == 25.0 else args.serve_lanes
        gids.append(sv.add_grid(eqd.bp8_grid_pair(dz, Lf, Ln), lam, mu, l_f=Lf, max_columns=lanes))
        n2f = sv.grids[-1]["n2f"]
        axes.append(np.linspace(-Lf, Lf, n2f))
    t_setup = time.perf_counter() - t0
    info = _serve_stream(sv, gids, axes, specs, weights, args.serve_rate, args.serve_seconds, 20261203)
    info.update(setup_s=round(t_setup, 2), hbm_pool_GB=round(sum(g["hbm_bytes"] for g in sv.grids) / 1e9, 2),
                lanes=[g["slots"] for g in sv.grids])
    if args.serve_rate_high > 0 and args.serve_seconds_high > 0:
        info["high_rate"] = _serve_stream(sv, gids, axes, specs, weights, args.serve_rate_high,
                                          args.serve_seconds_high, 20261204)
    return info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dz", type=float, default=50.0)
    ap.add_argument("--L_fault", type=float, default=800.0)
    ap.add_argument("--L_normal", type=float, default=400.0)
    ap.add_argument("--l_f", type=float, default=800.0)
    ap.add_argument("--tile_k", type=int, default=0, help="columns per tile (0 = autotuned over 8/16/32/64)")
    ap.add_argument("--reliable_interval", type=int, default=0,
                    help="mixed precision: iterations between FP64 residual refreshes (0 = library default)")
    ap.add_argument("--plain_cols", type=int, default=16)
    ap.add_argument("--cpu_cols", type=int, default=1024,
                    help="columns of the CPU oracle sample (about 10 s on 16 host threads)")
    ap.add_argument("--no_cpu", action="store_true")
    ap.add_argument("--cpu_cols_1core", type=int, default=48,
                    help="columns of the one-thread CPU oracle sample (about 7 s)")
    ap.add_argument("--no_extras", action="store_true", help="only the timed mixed-precision steps")
    ap.add_argument("--no_bp8", action="store_true", help="skip the BP8 time-loop measurement")
    ap.add_argument("--no_plain", action="store_true", help="skip the dz = 100 m end-to-end configuration")
    ap.add_argument("--no_lf400", action="store_true", help="skip the l_f = 400 reading of the headline grid")
    ap.add_argument("--no_serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no_production", action="store_true", help="skip the dz = 25 m production-size build")
    ap.add_argument("--prod_cols", type=int, default=512, help="exact columns of the dz = 25 m subset")
    ap.add_argument("--no_ensemble", action="store_true", help="skip the dz = 20 m trajectory ensemble")
    ap.add_argument("--ens_members", type=int, default=256)
    ap.add_argument("--ens_hours", type=float, default=720.0, help="ensemble horizon (30 days)")
    ap.add_argument("--ens_pw_hours", type=float, default=3.0,
                    help="horizon of the Peaceman-well ensemble (0: skip; its steps grow fast with the horizon)")
    ap.add_argument("--ens_pw_imex_members", type=int, default=256,
                    help="members of the Peaceman-well ensemble run over --ens_pw_imex_hours with the implicit-local "
                         "integrator (lockstep over the members; at most --ens_members)")
    ap.add_argument("--ens_pw_imex_hours", type=float, default=720.0)
    ap.add_argument("--ens_pw_maxiters", type=int, default=3000,
                    help="step budget per Peaceman-well member (a longer run returns MaxIters)")
    ap.add_argument("--serve_seconds", type=float, default=10.0)
    ap.add_argument("--serve_rate", type=float, default=10.0, help="mean request rate of the first stream")
    ap.add_argument("--serve_rate_high", type=float, default=50.0,
                    help="mean request rate of the second stream (0: skip); 50 req/s over this grid mix exceeds "
                         "one GPU, so the stream measures the saturation throughput and the backlog")
    ap.add_argument("--serve_seconds_high", type=float, default=5.0)
    ap.add_argument("--serve_lanes", type=int, default=512, help="lanes per grid in the serving pools")
    ap.add_argument("--serve_lanes_dz25", type=int, default=64,
                    help="lanes of the dz = 25 m pool: one tile already fills the device there, so a larger "
                         "pool adds latency and no throughput")
    ap.add_argument("--copy_calib_mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (3 times) and exit: the "
                         "reference kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calib_mib:
        copy_calibration(args.copy_calib_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the K build would shard by "
              f"columns across GPUs, one process per GPU; only the one-GPU path is built)", file=sys.stderr)
        sys.exit(2)

    mu, nu = 32.04e9, 0.25
    lam = 2 * mu * nu / (1 - 2 * nu)
    g = eqd.bp8_grid_pair(args.dz, args.L_fault, args.L_normal)
    fe = eqd.FaultElasticity(g, lam, mu, l_f=args.l_f, precision="mixed", tile_k=args.tile_k,
                             reliable_interval=args.reliable_interval)
    nf = eqd.frictional_node_count(fe)
    nf2 = 2 * nf
    ncols = nf2
    dof = fe.Ntot
    nred = fe.op.nreduced
    print(f"bench: dz={args.dz} L=({args.L_fault},{args.L_normal}) l_f={args.l_f}: DOF {dof}, "
          f"device unknowns {nred}, nf {nf}, {ncols} columns", file=sys.stderr, flush=True)
    Kbuf = DeviceBuffer(nf2 * nf2 * 8)

    for i in range(args.warmup):
        build_device(fe.rs, Kbuf, nf2)
        print(f"bench: warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    times = []
    for i in range(args.steps):
        t0 = time.perf_counter()
        build_device(fe.rs, Kbuf, nf2)
        times.append(time.perf_counter() - t0)
        print(f"bench: step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    ms_per_step = 1e3 * sum(times) / len(times)
    value = ncols / (ms_per_step / 1e3)
    K_mixed = Kbuf.to_host((nf2, nf2))
    rep = eqd.solver_report(fe.rs)
    iters_per_col = rep["iterations"] / rep["solves"]

    out = {
        "metric": "K-columns/s at dz=50 to 1e-8 rel (mixed block CG vs FP64); % HBM peak; p99 lat",
        "value": round(value, 2), "unit": "K-columns/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 CG in fp64 reliable updates",
        "data": "synthetic: unit-slip right-hand sides of the exact K build",
        "config": {"workload": f"exact K build, dz={args.dz:g} m, L_fault={args.L_fault:g}, "
                               f"L_normal={args.L_normal:g}, l_f={args.l_f:g}",
                   "columns": ncols, "dof": dof, "device_unknowns": nred, "tile_k": rep["tile_k"],
                   "tile_k_autotuned": args.tile_k == 0,
                   "rtol": 1e-10, "mean_cg_iterations": round(iters_per_col, 2)},
    }

    if not args.no_extras:
        # batched FP64 (device alpha/beta) on the same GPU
        fe64 = eqd.FaultElasticity(g, lam, mu, l_f=args.l_f, precision="fp64", tile_k=args.tile_k)
        build_device(fe64.rs, Kbuf, nf2)  # untimed: the first build also autotunes the tile width
        t0 = time.perf_counter()
        build_device(fe64.rs, Kbuf, nf2)
        t64 = time.perf_counter() - t0
        K64 = Kbuf.to_host((nf2, nf2))
        err_vs_fp64 = float(np.linalg.norm(K_mixed - K64) / np.linalg.norm(K64))
        print(f"bench: batched fp64 build {t64 * 1e3:.1f} ms, mixed vs fp64 K rel err {err_vs_fp64:.2e}",
              file=sys.stderr, flush=True)
        # plain one-RHS FP64 CG, host alpha/beta, on a column sample
        rng = np.random.default_rng(20261203)
        sample = np.sort(rng.choice(ncols, size=min(args.plain_cols, ncols), replace=False))
        fep = eqd.FaultElasticity(g, lam, mu, l_f=args.l_f, engine="plain")
        t0 = time.perf_counter()
        Kp = eqd.fault_stiffness(fep, cols=sample)
        tp = time.perf_counter() - t0
        plain_value = len(sample) / tp
        err_plain = float(np.linalg.norm(Kp - K64[:, sample]) / np.linalg.norm(K64[:, sample]))
        print(f"bench: plain fp64 {len(sample)} cols in {tp:.2f} s ({plain_value:.1f} cols/s), "
              f"vs batched fp64 {err_plain:.2e}", file=sys.stderr, flush=True)
        # stencil kernel timing (events around every launch) for the roofline
        fet = eqd.FaultElasticity(g, lam, mu, l_f=args.l_f, precision="mixed", tile_k=args.tile_k, timing=1,
                                  reliable_interval=args.reliable_interval)
        build_device(fet.rs, Kbuf, nf2)
        names = ["stencil", "update", "pupdate", "finalize", "xaccum", "residual64"]
        kt = {n: fet.rs.kernel_timing(i) for i, n in enumerate(names)}
        st_ms, st_n, st_tiles = kt["stencil"]
        avg_us = 1e3 * st_ms / st_n
        # algorithmic bytes of q = A p per launch: read p, write q (FP32) for every active column
        bytes_per_launch = (st_tiles / st_n) * eqd.solver_report(fet.rs)["tile_k"] * nred * 4 * 2
        achieved = bytes_per_launch / (avg_us * 1e-6)
        tot_ms = sum(v[0] for v in kt.values())
        # whole-iteration metric of the survey: 10 vector passes x DOF x 4 B per column-iteration
        col_iters = rep["iterations"] / max(1, rep["solves"]) * ncols
        survey_gbs = 10 * dof * 4 * col_iters / (ms_per_step / 1e3)
        # x1 lines of 9 to 16 nodes: the x1-inner layout and its x3-marching stencil; 17 to 64
        # nodes: the x1-inner layout with one wave per 16-node segment; longer lines: segmented
        # node-major pencils
        nx = 2 * g.n1 - 3
        tk = eqd.solver_report(fet.rs)["tile_k"]
        if tk != 64:
            stencil_name = f"k_apply_dot<float,{tk}>"
        elif 9 <= nx <= 16:
            stencil_name = "k_xi_pencil<float,float,0,3>"
        elif nx <= 64:
            stencil_name = "k_xl_pencil<float,float,0,3>"
        else:
            stencil_name = "k_pencil<float,float,0,false>"
        out["roofline"] = {"bound": "hbm", "kernel": stencil_name,
                           "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK, 4), "traffic": measured_traffic(stencil_name),
                           "avg_launch_us": round(avg_us, 2), "launches": st_n,
                           "bytes_per_launch": int(bytes_per_launch),
                           # the same two passes (read p, write q) counted over every DOF of the
                           # split-node system, as the per-iteration 10 x DOF x 4 B figure counts them
                           "achieved_full_dof": round(achieved * dof / nred / 1e9, 1),
                           "frac_full_dof": round(achieved * dof / nred / HBM_PEAK, 4)}
        out["kernel_share"] = {n: round(v[0] / tot_ms, 4) for n, v in kt.items()}
        out["survey_algorithmic_GBps"] = round(survey_gbs / 1e9, 1)
        # the survey's full formula: + 6 DOF 8 B per refinement sweep of a column (here every
        # column of a tile that the FP64 residual pass covered, from the timing-mode build) and
        # + 3 DOF 8 B per column (RHS, U, extraction)
        sweeps = kt["residual64"][2] * 64
        full_bytes = 10 * dof * 4 * col_iters + 6 * dof * 8 * sweeps + 3 * dof * 8 * ncols
        out["survey_algorithmic_GBps_full"] = round(full_bytes / (ms_per_step / 1e3) / 1e9, 1)
        out["refinement_sweeps_per_column"] = round(sweeps / ncols, 2)
        out["plain_fp64_one_rhs"] = {"value": round(plain_value, 2), "unit": "K-columns/s",
                                     "sample_columns": len(sample)}
        out["batched_fp64"] = {"value": round(ncols / t64, 2), "unit": "K-columns/s"}
        out["speedup_vs_plain_fp64"] = round(value / plain_value, 1)
        # the ratio's two factors: batching (64-column tiles, device-side alpha/beta, graphs) in
        # FP64, then FP32 CG inside FP64 reliable updates
        out["speedup_decomposition"] = {"batched_fp64_vs_plain_fp64": round((ncols / t64) / plain_value, 1),
                                        "mixed_vs_batched_fp64": round(value / (ncols / t64), 2)}
        out["K_rel_err_vs_fp64"] = err_vs_fp64

        if not args.no_cpu:
            from oracle.oracle import OracleSystem, bp8_geometry
            o = OracleSystem(bp8_geometry(args.dz, args.L_fault, args.L_normal, args.l_f))
            ncpu = min(os.cpu_count() or 1, 16)
            csample = np.sort(rng.choice(ncols, size=min(args.cpu_cols, ncols), replace=False))
            t0 = time.perf_counter()
            Kc = o.fault_stiffness_cols(csample, nthreads=ncpu)
            tc = time.perf_counter() - t0
            err_oracle = float(np.linalg.norm(K_mixed[:, csample] - Kc) / np.linalg.norm(Kc))
            # the same restatement on one thread, on the first columns of the sample
            c1 = csample[:max(1, min(args.cpu_cols_1core, len(csample)))]
            t0 = time.perf_counter()
            o.fault_stiffness_cols(c1, nthreads=1)
            t1 = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(len(csample) / tc, 3), "unit": "K-columns/s", "cores": ncpu,
                                   "kind": "port",
                                   "sample": f"{len(csample)} seeded random columns of the same K build, "
                                             f"assembled-CSR FP64 CG, one column per thread",
                                   "one_core": {"value": round(len(c1) / t1, 3), "columns": len(c1)},
                                   "host": host_info()}
            print(f"bench: cpu oracle {len(c1)} cols on 1 thread in {t1:.1f} s", file=sys.stderr, flush=True)
            out["K_rel_err_vs_oracle_sample"] = err_oracle
            print(f"bench: cpu oracle {len(csample)} cols in {tc:.1f} s on {ncpu} threads; "
                  f"mixed K vs oracle {err_oracle:.2e}", file=sys.stderr, flush=True)
    Kbuf.free()
    if not args.no_extras and not args.no_lf400:
        out["reference_reading_lf400"] = bench_lf400(args, lam, mu)
    if not args.no_extras and not args.no_bp8:
        out["bp8_30day_gs"] = bench_bp8(args)
    if not args.no_extras and not args.no_bp8 and not args.no_plain:
        out["plain_config_dz100"] = bench_plain_config(args)
    if not args.no_extras and not args.no_production:
        out["production_dz25"] = bench_production(args)
    if not args.no_extras and not args.no_ensemble:
        out["ensemble_dz20"] = bench_ensemble(args)
    if not args.no_extras and not args.no_serve:
        out["serving"] = bench_serving(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
