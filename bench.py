#!/usr/bin/env python3
"""bench.py -- time-to-converge of the 256^3 / 128 A / 846-atom KH 3D-RISM solve
on one MI355X, mixed precision (headline `value`) against the FP64 path.

A step is one complete job on inputs already resident in HBM: grid potentials,
the fixed-point solve from the initial estimate to RMS residual <= 1e-6, and
the thermodynamic reductions (Delta-mu, PMV, SE).  Warm-up steps run first,
then `--steps` timed steps bracketed by device synchronisation.

The JSON line also carries:
  roofline      the dominant iteration kernel of the mixed path: algorithmic
                bytes per launch / average launch time (HIP events around every
                launch on the solver's stream, collected during one extra
                profiled step), against the 8 TB/s HBM peak;
  cpu_baseline  the FP64 host oracle (OpenMP, the host's cores) on the same
                deck, the whole job measured (--cpu-sample: a bounded sample,
                scaled); Delta-mu, PMV and g of both GPU paths are reported
                against this oracle run.
Solvent: 3-site water, chi_vv(k) from the library's own 1D DRISM-KH solve (eps 78.4)
(16384 points, dr 0.01 A, 298.15 K; rism3d.configs.solvent_file), computed
before any timing and cached.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "3d-rism-cuda_amd"), REPO]

import numpy as np  # noqa: E402
import torch  # noqa: E402,F401  (first: its HIP runtime then serves the solver library too)

METRIC = "s-to-converge 256^3 KH, mixed vs FP64 at Δμ≤0.01 kcal/mol; % HBM roofline"
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30  # calibration copy: reads and writes this many bytes
TRAFFIC = os.path.join(REPO, "profiles", "iteration_kernels_256cubed_mixed.json")


def kernel_bytes(V, P, H, mixed):
    """Algorithmic HBM bytes per grid point of each iteration kernel: every
    operand read once and every result written once at its stored width, for
    a steady-state Anderson step with H filled history slots (V sites, P
    complex packs of two sites).

    FP64 path: the state t_s, the closure potential u, the residual R and
    the H-slot histories of t_s and R in FP64; the spectrum in complex FP64.
    Mixed path, FP32 chain (every step but the few FP64 refreshes): the
    closure pass reads the reference A = T - u (FP32), the change dl (FP32, also
    written), the current R and H history residuals from the ring, H
    increments and R_ref (FP32), writes one increment and B (FP32) and the
    spectrum (complex FP32); the residual pass reads the spectrum, B and the
    H history residuals and writes R into the ring; no long-range term.
    """
    if mixed:
        return {
            "closure_fx": V * (4 + 4 + 4 + 4 + 8 * H + 4 + 4 + 4) + 8 * P,
            "fft_y_fwd": 16 * P,
            "fft_z_oz": 16 * P,
            "fft_y_inv": 16 * P,
            "resid": 8 * P + V * (4 + 4 * H + 4),
        }
    return {
        # t_s, u, R, H x (t_i, R_i) in; new slot (t, R) and t_s out; spectrum out
        "closure_fx": V * (8 + 8 + 8 + 16 * H + 16 + 8) + 16 * P,
        "fft_y_fwd": 32 * P,
        # spectrum in/out + the half-grid long-range term
        "fft_z_oz": 32 * P + 8,
        "fft_y_inv": 32 * P,
        # spectrum, t_s and u, H history R_i in; R out
        "resid": 16 * P + V * (8 + 8 + 8 * H + 8),
    }


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by its copy-kernel
    calibration); None when no summary covers it."""
    try:
        with open(TRAFFIC) as f:
            return json.load(f)["traffic_bytes_per_launch"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None


def per_kernel_table(kt, kb, N):
    out = {}
    for k, (kms, kn) in kt.items():
        if kn:
            e = {"avg_us": round(kms / kn * 1e3, 2), "launches": kn}
            if k in kb:
                e["bytes_per_pt"] = kb[k]
                e["GBps"] = round(kb[k] * N / (kms / kn / 1e3) / 1e9, 1)
            out[k] = e
    return out


def run_job(s):
    s.build_potentials()
    st, steps, rms = s.solve()
    x = s.thermo()
    return st, steps, rms, x


def time_path(rism3d, w, solvent, prec, steps, warmup):
    import torch
    opts = rism3d.Options(precision=prec, nhist=w.nhist)
    s = rism3d.Solver(w.n, w.box, solvent, w.solute, opts)
    for _ in range(warmup):
        run_job(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = run_job(s)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    st, nit, rms, x = res
    g = s.field(rism3d.FIELD_G).reshape(-1)
    stats = s.stats()
    s.profile(True)
    run_job(s)
    kt = s.kernel_times()
    s.profile(False)
    geo = s.geometry()
    s.close()
    return {"s": dt, "status": st, "iterations": nit, "rms": rms, "xmu": x, "stats": stats,
            "kernels": kt, "geometry": geo, "g": g}


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, fp64_iterations, full=True, iters=8, sdiv=4):
    """The FP64 host oracle (C/OpenMP restatement of the reference path, all
    host cores) on the same deck.  full: the whole job measured (k-grid and
    spline, potentials, the solve to RMS 1e-6, thermodynamics), its g written
    for the GPU paths' error; otherwise a bounded sample (potentials and SE on
    1/sdiv of the points, `iters` iterations) scaled to the job.  Per-step
    lines of the oracle go to stderr every 50 steps."""
    sys.path.insert(0, REPO)
    from rism3d import configs
    from oracle import oracle as orc
    if not os.path.exists(orc.EXE):
        return None, None
    d = tempfile.mkdtemp()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    g = None
    try:
        shutil.copy(configs.solvent_file("water"), os.path.join(d, "h2o_vv.xvk"))
        configs.write_inp(os.path.join(d, "c1.inp"), w, "h2o_vv.xvk")
        cmd = [orc.EXE, "-j", "-f", "-T", str(threads)]
        cmd += ["-G", "g.bin"] if full else ["-q", "-n", str(iters), "-x", str(sdiv)]
        p = subprocess.Popen(cmd + ["c1.inp"], cwd=d, stdout=subprocess.PIPE, text=True)
        last = ""
        for ln in p.stdout:
            if ln.startswith(" Step = ") and int(ln.split()[2]) % 50 == 0:
                print(f"cpu_baseline oracle:{ln.rstrip()}", file=sys.stderr, flush=True)
            if ln.strip():
                last = ln
        if p.wait(timeout=900) != 0:
            raise RuntimeError(f"oracle exited with {p.returncode}: {last[:200]}")
        rec = json.loads(last)
        if full:
            g = np.fromfile(os.path.join(d, "g.bin"), dtype=np.float64)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    host = f"{cpu_model()}, nproc {os.cpu_count()}, {rec['threads']} OpenMP threads"
    if full:
        total = rec["t_setup"] + rec["t_iter"] + rec["t_thermo"]
        return {"value": round(total, 3), "unit": "s", "cores": rec["threads"], "kind": "port",
                "sample": (f"the whole job, measured: FP64 oracle (C/OpenMP restatement of the reference path) on "
                           f"the same 256^3 deck, k-grid+spline {rec['t_setup'] - rec['t_pot']:.1f} s, potentials "
                           f"{rec['t_pot']:.1f} s, {rec['steps']} iterations {rec['t_iter']:.1f} s, thermo with SE "
                           f"{rec['t_thermo']:.1f} s; {host}"),
                "steps": rec["steps"], "sfe_sc_J_mol": rec["sfe_sc"], "sfe_gf_J_mol": rec["sfe_gf"],
                "pmv_L_mol": rec["pmv"], "host": host}, g
    k = rec["sample_div"]
    per_it = rec["t_iter"] / rec["steps"]
    setup = rec["t_setup"] - rec["t_pot"] + rec["t_pot"] * k
    total = setup + per_it * fp64_iterations + rec["t_thermo"] * k
    return {"value": round(total, 3), "unit": "s", "cores": rec["threads"], "kind": "port",
            "sample": (f"FP64 oracle (C/OpenMP restatement of the reference path) on the same 256^3 deck: "
                       f"k-grid+spline {rec['t_setup'] - rec['t_pot']:.1f} s, potentials on 1/{k} of the "
                       f"points {rec['t_pot']:.1f} s (x{k}), {rec['steps']} iterations at {per_it:.2f} s "
                       f"(x{fp64_iterations}, the FP64 solve's count), thermo with SE on 1/{k} of the "
                       f"points {rec['t_thermo']:.1f} s (x{k}); {host}"),
            "measured_s": round(rec["t_setup"] + rec["t_iter"] + rec["t_thermo"], 2), "host": host}, None


def side_workload(which, rism3d, configs):
    """The other two measured configurations (not the headline line)."""
    import numpy as np
    if which == "batch":
        # tile shapes tuned once per grid shape (rism_autotune through rism_solve)
        os.environ.setdefault("RISM_AUTOTUNE", "3")
        sv = configs.solvent("water")
        systems = [(w.n, w.box, w.solute) for w in configs.batch_config(256)]
        rism3d.batch_solve(systems[:8], sv, rism3d.Options(), workers=4)  # warm-up
        r64, s64 = rism3d.batch_solve(systems, sv, rism3d.Options(), workers=8)
        rmx, smx = rism3d.batch_solve(systems, sv, rism3d.Options(precision=rism3d.MIXED), workers=8)
        err = max(abs(a["SFE_SC"] - b["SFE_SC"]) for a, b in zip(rmx, r64))
        return {"metric": "solutes/s, 256 seeded 8-40-atom solutes on 60-72^3 grids (KH, RMS 1e-6)",
                "value": round(len(systems) / smx, 2), "unit": "solutes/s", "higher_is_better": True,
                "fp64_solutes_per_s": round(len(systems) / s64, 2), "workers": 8,
                "autotune_reps": int(os.environ["RISM_AUTOTUNE"]),
                "converged": sum(r["status"] == 0 for r in rmx), "max_dmu_err_vs_fp64_J_mol": round(err, 4)}
    sv = configs.solvent("nacl")
    w = configs.config(4)
    out = {"metric": "s-to-converge 512^3 / 256 A / 5000-atom KH job (potentials + solve + thermo)",
           "unit": "s", "higher_is_better": False,
           "solvent": "water + 0.15 M NaCl (4 reduced sites), chi from the library's 1D DRISM-KH solve (eps 78.4)"}
    for name, prec in (("mixed", rism3d.MIXED), ("fp64", rism3d.FP64)):
        # buffers a destroyed handle left for reuse would not show as new
        # device memory: return them first, so device_GB is this path's own
        rism3d.release_cache()
        free0, _ = torch.cuda.mem_get_info()
        s = rism3d.Solver(w.n, w.box, sv, w.solute, rism3d.Options(precision=prec, nhist=w.nhist))
        free1, _ = torch.cuda.mem_get_info()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.build_potentials()
        st, steps, rms = s.solve()
        x = s.thermo()
        dt = time.perf_counter() - t0
        r = {"status": st, "iterations": steps, "s": round(dt, 3), "device_GB": round((free0 - free1) / 1e9, 1),
             "SFE_SC": x["SFE_SC"], "PMV": x["PMV"]}
        if name == "mixed":
            t1 = time.perf_counter()
            g = s.adjoint_grad(1e-8)
            r["adjoint_s"] = round(time.perf_counter() - t1, 3)
            r["adjoint_iterations"] = g["steps"]
            cg = s.grad()
            ref = cg[:, :3] + cg[:, 3:]
            r["adjoint_vs_cal_grad_max_rel"] = float(np.abs(g["xyz"] - ref).max() / np.abs(ref).max())
        s.close()
        out[name] = r
    out["value"] = out["mixed"]["s"]
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["s"] / out["mixed"]["s"], 3)
    out["dmu_mixed_minus_fp64_J_mol"] = out["mixed"]["SFE_SC"] - out["fp64"]["SFE_SC"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-sample", action="store_true",
                    help="CPU baseline on a bounded sample of the job (scaled) instead of the whole job")
    ap.add_argument("--mixed-only", action="store_true", help="time only the mixed path")
    ap.add_argument("--fp64-only", action="store_true",
                    help="time only the FP64 path (profiling its kernels on their own; no headline line)")
    ap.add_argument("--workload", choices=("256", "512", "batch"), default="256",
                    help="256 (default): the headline 256^3 job; 512: the 512^3 / 5000-atom electrolyte "
                         "job (memory, times, adjoint forces); batch: 256 small solutes (solutes/s)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run the 1 GiB streaming copy 3 times instead of 20 (fewer dispatches under a "
                         "counter-collecting profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        sys.exit(2)
    import rism3d
    from rism3d import configs
    lib = rism3d.lib()
    if lib.rism_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload != "256":
        print(json.dumps(side_workload(args.workload, rism3d, configs)))
        return
    # the measured HBM ceiling: the library's plain streaming copy of 1 GiB
    # (its dispatches, rism::k_stream_copy, are also the known-size reference
    # for scaling the FETCH_SIZE / WRITE_SIZE counters)
    copy_gbps, copy_us = rism3d.stream_copy(COPY_BYTES, 3 if args.calibrate_copy else 20)
    w = configs.config(1)
    solvent = configs.solvent("water")
    V = solvent.nsite
    P = (V + 1) // 2
    N = int(np.prod(w.n))
    if args.fp64_only:
        fp64 = time_path(rism3d, w, solvent, rism3d.FP64, args.steps, args.warmup)
        kb = kernel_bytes(V, P, w.nhist, False)
        print(json.dumps({"workload": w.name, "path": "fp64", "s": round(fp64["s"], 4),
                          "iterations": fp64["iterations"], "copy_GBps": round(copy_gbps, 1),
                          "kernels_fp64": per_kernel_table(fp64["kernels"], kb, N)}))
        return
    mixed = time_path(rism3d, w, solvent, rism3d.MIXED, args.steps, args.warmup)
    fp64 = None if args.mixed_only else time_path(rism3d, w, solvent, rism3d.FP64, args.steps, args.warmup)
    # roofline: dominant iteration kernel of the mixed path
    kb = kernel_bytes(V, P, w.nhist, True)
    it_kernels = {k: v for k, v in mixed["kernels"].items() if k in kb and v[1] > 0}
    dom = max(it_kernels, key=lambda k: it_kernels[k][0])
    ms, n = it_kernels[dom]
    avg_s = ms / n / 1e3
    bytes_launch = kb[dom] * N
    achieved = bytes_launch / avg_s / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
            "bytes_per_launch": bytes_launch, "avg_launch_us": round(avg_s * 1e6, 2),
            "copy_GBps": round(copy_gbps, 1), "frac_of_copy": round(achieved / copy_gbps, 4)}
    per_kernel = per_kernel_table(mixed["kernels"], kb, N)
    cpu, g_orc = None, None
    if not args.no_cpu and fp64 is not None:
        try:
            cpu, g_orc = cpu_baseline(w, fp64["iterations"], full=not args.cpu_sample)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "error": str(e)[:200]}
    out = {
        "metric": METRIC,
        "value": round(mixed["s"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["s"] * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields and changes / f64 reference",
        "data": "synthetic",
        "config": {"workload": w.name, "grid": list(w.n), "box_A": list(w.box), "atoms": len(w.solute.q),
                   "solvent": "3-site water, chi from the library's 1D DRISM-KH solve (eps 78.4, 16384 pts, dr 0.01 A)", "closure": "KH",
                   "tol": 1e-6, "anderson": {"nhist": w.nhist, "count": 2, "m": 0.6, "mp": 0.2},
                   "parallelism": "1 GPU"},
        "iterations_mixed": mixed["iterations"],
        "rms_mixed": mixed["rms"],
        "roofline": roof,
        "kernels_mixed": per_kernel,
        "cpu_baseline": cpu,
    }
    if fp64 is not None:
        out.update({
            "fp64_s": round(fp64["s"], 4),
            "iterations_fp64": fp64["iterations"],
            "speedup_mixed_vs_fp64": round(fp64["s"] / mixed["s"], 3),
            "dmu_mixed_minus_fp64_J_mol": round(mixed["xmu"]["SFE_SC"] - fp64["xmu"]["SFE_SC"], 4),
            "sfe_sc_fp64_J_mol": fp64["xmu"]["SFE_SC"],
            "pmv_fp64_L_mol": fp64["xmu"]["PMV"],
            "kernels_fp64": per_kernel_table(fp64["kernels"], kernel_bytes(V, P, w.nhist, False), N),
        })
    if cpu and cpu.get("sfe_sc_J_mol") is not None:
        # both GPU paths against the oracle's own solve of the same deck on this host
        out["dmu_mixed_minus_oracle_J_mol"] = round(mixed["xmu"]["SFE_SC"] - cpu["sfe_sc_J_mol"], 4)
        if fp64 is not None:
            out["dmu_fp64_minus_oracle_J_mol"] = round(fp64["xmu"]["SFE_SC"] - cpu["sfe_sc_J_mol"], 6)
        out["pmv_mixed_minus_oracle_L_mol"] = mixed["xmu"]["PMV"] - cpu["pmv_L_mol"]
        if g_orc is not None and g_orc.size == mixed["g"].size:
            out["g_maxabs_mixed_vs_oracle"] = float(np.abs(mixed["g"] - g_orc).max())
            if fp64 is not None:
                out["g_maxabs_fp64_vs_oracle"] = float(np.abs(fp64["g"] - g_orc).max())
    print(json.dumps(out))


if __name__ == "__main__":
    main()
