"""Runs SEAS BP8-QD on the device and writes the benchmark's section 4 output files: the
command-line driver of scripts/run_bp8.jl:1-79 over libeqd.

    python -m eqd.run_bp8 [gs|pw] [dz] [L_fault] [L_normal] [--outdir DIR] [--alg tsit5|imex]
                          [--tf_days DAYS]

The positional arguments and their defaults are the reference script's (scripts/run_bp8.jl:14-17):
injection "pw" selects the Peaceman well, anything else the Gaussian source; dz = 50 m,
L_fault = 800 m, L_normal = 400 m. The model is build_model with the reference's defaults
(Toeplitz stiffness, built on the device), the run is run_bp8 with saveat = 300 s over t_f (30
days), then the summary (peak slip rate, final slip, peak pressure at the origin, the effective
stress and resolution reports) and write_outputs into output/BP8-QD-<GS|PW>_dz<dz>_Lf<L>_Ln<L>
(under the current directory unless --outdir is given). The modeler written into the file headers
is $BP8_MODELER. --alg imex selects the implicit-local integrator for the stiff Peaceman variant;
--tf_days shortens the horizon.
"""
import argparse
import os
import sys
import time

import numpy as np

from . import bp8


def parse_args(argv):
    ap = argparse.ArgumentParser(prog="python -m eqd.run_bp8", description=__doc__.split("\n\n")[0])
    ap.add_argument("injection", nargs="?", default="gs", help="gs (Gaussian source) or pw (Peaceman well)")
    ap.add_argument("dz", nargs="?", type=float, default=50.0)
    ap.add_argument("L_fault", nargs="?", type=float, default=800.0)
    ap.add_argument("L_normal", nargs="?", type=float, default=400.0)
    ap.add_argument("--outdir", default=None, help="output directory (default output/BP8-QD-...)")
    ap.add_argument("--alg", default="tsit5", choices=("tsit5", "imex"))
    ap.add_argument("--tf_days", type=float, default=None, help="horizon in days (default t_f = 30 days)")
    a = ap.parse_args(argv)
    a.injection = "peaceman" if a.injection.lower() == "pw" else "gaussian"
    return a


def default_outdir(injection: str, dz: float, L_fault: float, L_normal: float) -> str:
    """The reference's output directory name (scripts/run_bp8.jl:19-20)."""
    tag = "GS" if injection == "gaussian" else "PW"
    return os.path.join("output", f"BP8-QD-{tag}_dz{int(dz)}_Lf{int(L_fault)}_Ln{int(L_normal)}")


def summarize(m, sol, out=sys.stdout):
    """The reference script's summary (scripts/run_bp8.jl:37-75): peak max |V| over the saves (each
    evaluated with evaluate! in save order), final slip at the node nearest the origin and the
    largest final slip, the peak pressure there, then the effective stress and resolution reports.
    Returns the figures."""
    nf = m.nf
    ic = int(np.argmin(np.abs(m.x2))) + int(np.argmin(np.abs(m.x3))) * len(m.x2)
    vmax = bp8.saved_vmax(m, len(sol.t))
    jp = int(np.argmax(vmax))
    vpeak, tpeak = float(vmax[jp]), float(sol.t[jp])
    ppeak = max(0.0, float(sol.u[:, 3 * nf + ic].max()))
    uend = sol.u[-1]
    w = out.write
    w(f"\npeak slip rate      {vpeak:.4E} m/s at t = {tpeak / 86400:.3f} days\n")
    w(f"final slip (0,0)    {uend[ic]:.4E} m\n")
    w(f"final max slip      {np.abs(uend[:nf]).max():.4E} m\n")
    w(f"peak pressure (0,0) {ppeak / 1e6:.4f} MPa   (min σ̄ = {(m.par['sigma0'] - ppeak) / 1e6:.4f} MPa)\n")
    es = bp8.effective_stress_report(m)
    if es["bound"]:
        w(f"\nWARNING: effective normal stress went to {es['sigma_lowest'] / 1e6:.4f} MPa (below the "
          f"{es['floor']:.0f} Pa floor)\n")
        w(f"         in {es['floor_hits']} evaluations. Fluid pressure has fully unclamped\n")
        w("         the fault there, so BP8's no-opening condition (eq. 3) no longer\n")
        w("         holds and the model is outside its range of validity at those nodes.\n")
    rr = bp8.resolution_report(m)
    w(f"\nresolution: Δz = {rr['dz']:.0f} m, process zone L_b = {rr['L_b']:.1f} m -> {rr['cells_per_Lb']:.2f} "
      f"cells per L_b\n")
    w(f"            Gaussian source width L_gauss = {m.par['L_gauss']:.0f} m -> {rr['cells_per_L_gauss']:.2f} cells\n")
    if not rr["converged"]:
        w("            NOT resolution-converged: slip rate depends exponentially on σ̄,\n")
        w("            so peak V here is indicative, not quantitative.\n")
    out.flush()
    return dict(vpeak=vpeak, tpeak=tpeak, final_slip_centre=float(uend[ic]),
                final_max_slip=float(np.abs(uend[:nf]).max()), ppeak=ppeak, stress=es, resolution=rr)


def main(argv=None, out=sys.stdout):
    a = parse_args(sys.argv[1:] if argv is None else argv)
    modeler = os.environ.get("BP8_MODELER", "unknown")
    tag = "GS" if a.injection == "gaussian" else "PW"
    outdir = a.outdir or default_outdir(a.injection, a.dz, a.L_fault, a.L_normal)
    out.write(f"[ Info: BP8-QD-{tag} dz = {a.dz} L_fault = {a.L_fault} L_normal = {a.L_normal} outdir = {outdir}\n")
    out.flush()
    t0 = time.perf_counter()
    m = bp8.build_model(dz=a.dz, L_fault=a.L_fault, L_normal=a.L_normal, injection=a.injection, verbose=True)
    out.write(f"[ Info: model built seconds = {time.perf_counter() - t0:.1f} frictional_nodes = {m.nf}\n")
    out.flush()
    tf = m.par["t_f"] if a.tf_days is None else a.tf_days * 86400.0
    t0 = time.perf_counter()
    sol = bp8.run_bp8(m, tspan=(0.0, tf), saveat=300.0, verbose=True, alg=a.alg)
    out.write(f"[ Info: integrated seconds = {time.perf_counter() - t0:.1f} steps = {len(sol.t)} "
              f"retcode = {sol.retcode}\n")
    out.flush()
    res = summarize(m, sol, out)
    bp8.write_outputs(m, sol, outdir, modeler=modeler)
    out.write(f"[ Info: outputs written outdir = {outdir} files = {len(os.listdir(outdir))}\n")
    out.flush()
    res.update(outdir=outdir, naccept=sol.naccept, retcode=sol.retcode, nsaved=len(sol.t))
    return res


if __name__ == "__main__":
    main()
