"""D2Q9 throughput on one MI355X: GLUPS at 4096 x 1024 (plain FP32 vs chosen form), % HBM roofline, L2 error.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--k 8] [--cpu-seconds 10] [--l2-steps 20]

Workload: channel flow past a cylinder, 4096 x 1024, R = 64 at (1024, 512), Re = 1000, Cs = 0.15,
parabolic inlet u_max = 0.05 (synthetic initial state: feq at rho = 1, u = inlet profile, seeded noise).
A "step" is one collide-stream update of all W*H cells with the state resident in HBM.  Timed region: K
steps bracketed by torch.cuda.synchronize(); HIP events on the launch stream give the per-launch kernel
time used for the roofline (72 B/cell/step algorithmic traffic for FP32 single steps).
The CPU baseline is the oracle's FP32 mirror (C, OpenMP over rows) on a bounded number of steps of the
same lattice, timed on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "lbm-cuda_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "D2Q9 GLUPS at 4096x1024 (plain fp32 vs chosen form), % HBM roofline, L2 err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_FP32 = 72  # 9 populations read + 9 written, 4 B each


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--k", type=int, default=8, help="steps per launch of the fused form")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="bound on the CPU-baseline sample")
    ap.add_argument("--l2-steps", type=int, default=20, help="steps for the FP32-vs-FP64 L2 error (0 = skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=("all", "aa", "fused"), default="all",
                    help="run only the plain single step or only the fused form (for profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy a device buffer of this many MiB (a known-size kernel for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine runs one lattice per GPU and has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from lbm import _lib, configs
    from lbm.sim import Simulation, _ptr, _stream

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        del src, dst
    case = configs.cylinder_large()
    W, H = case.W, case.H
    cells = W * H
    log(f"bench: {case.name}, {args.warmup} warmup + {args.steps} timed AA steps")
    sim = Simulation(W, H, case.tau_inv, **case.sim_kwargs(), device=dev)
    L = sim.lib

    def run_aa(n):
        rc = L.lbm_step_aa_n(ctypes.byref(sim.params), _ptr(sim.mask), _ptr(sim.f), _ptr(sim.inlet_u), sim.parity,
                             n, _stream())
        _lib.check(rc, "lbm_step_aa_n")
        sim.parity = (sim.parity + n) & 1
        sim.steps += n

    if args.only == "fused":
        args.warmup, args.steps = 2, 2
    # ---------------- plain FP32 single AA steps
    run_aa(args.warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    run_aa(args.steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    glups = cells / (ms_per_step * 1e-3) / 1e9
    achieved = BYTES_PER_CELL_FP32 * cells / (kern_ms * 1e-3) / 1e9
    log(f"plain fp32 AA: {ms_per_step * 1e3:.1f} us/step wall, {kern_ms * 1e3:.1f} us/launch (events), "
        f"{glups:.2f} GLUPS, {achieved:.0f} GB/s")
    forms = {"aa_fp32": {"glups": glups, "us_per_step": kern_ms * 1e3, "achieved_gbs": achieved}}

    # ---------------- fused form: autotuned (k, strip, chunk), then timed like the plain form
    if L.lbm_fused_max_k() >= 2 and args.k >= 2 and args.only != "aa":
        from lbm import autotune
        ks = tuple(k for k in autotune.KS if k <= args.k)
        best, table = autotune.tune_fused(sim, ks=ks, use_cache=False)
        for k, strip, chunk, us in table:
            log(f"tuned k={k}: strip={strip} chunk={chunk} {us:.2f} us/step")
        k, strip, chunk, _ = best
        nlaunch = max(1, args.steps // k)
        sim.step(2 * k, k=k, cfg=(strip, chunk))
        torch.cuda.synchronize()
        e0.record()
        sim.step(nlaunch * k, k=k, cfg=(strip, chunk))
        e1.record()
        torch.cuda.synchronize()
        fk_ms = e0.elapsed_time(e1) / (nlaunch * k)
        fg = cells / (fk_ms * 1e-3) / 1e9
        forms[f"fused_fp32_k{k}"] = {"glups": fg, "us_per_step": fk_ms * 1e3, "x_plain": fg / glups,
                                     "strip": strip, "chunk": chunk,
                                     "tuned": {f"k{r[0]}": {"strip": r[1], "chunk": r[2], "us_per_step": r[3]}
                                               for r in table}}
        log(f"fused fp32 k={k} strip={strip} chunk={chunk}: {fk_ms * 1e3:.1f} us/step, {fg:.2f} GLUPS "
            f"({fg / glups:.2f}x plain)")

    # ---------------- L2 error of the FP32 GPU run against the FP64 oracle (same initial state)
    l2 = None
    if args.l2_steps > 0:
        import oracle as O
        s2 = Simulation(W, H, case.tau_inv, **case.sim_kwargs(), device=dev)
        s2.step(args.l2_steps)
        rho_g, ux_g, uy_g = (t.cpu().numpy().astype(np.float64) for t in s2.macro())
        p = s2.params
        op = O.params(W, H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp)
        f64 = np.ascontiguousarray(case.f0, dtype=np.float64)
        t1 = time.perf_counter()
        par = O.aa_steps(op, s2.mask_host, f64, case.inlet_u.astype(np.float64), 0, args.l2_steps)
        log(f"FP64 oracle {args.l2_steps} steps: {time.perf_counter() - t1:.1f} s")
        rho_o, ux_o, uy_o = O.macro(op, s2.mask_host, f64, case.inlet_u.astype(np.float64), par)
        upd = O.observable(op, s2.mask_host, 1)[0]
        u_err = np.sqrt(((ux_g - ux_o) ** 2 + (uy_g - uy_o) ** 2)[upd].sum() / (ux_o ** 2 + uy_o ** 2)[upd].sum())
        r_err = np.sqrt(((rho_g - rho_o) ** 2)[upd].sum() / (rho_o ** 2)[upd].sum())
        l2 = {"steps": args.l2_steps, "rho": float(r_err), "u": float(u_err), "vs": "FP64 oracle, same IC"}
        log(f"L2 err after {args.l2_steps} steps: rho {r_err:.2e}, u {u_err:.2e}")
        del s2

    # ---------------- CPU baseline: the oracle's FP32 mirror on this host
    cpu = None
    if not args.no_cpu:
        import oracle as O
        p = sim.params
        op = O.params(W, H, p.tau_inv, cs=p.cs, model=p.model, bc=p.bc, clamp=p.clamp)
        fc = np.ascontiguousarray(case.f0, dtype=np.float32)
        prof = case.inlet_u.astype(np.float32)
        cores = len(os.sched_getaffinity(0))
        n, par = 0, 0
        t1 = time.perf_counter()
        while True:
            O.aa_step(op, sim.mask_host, fc, prof, par)
            par ^= 1
            n += 1
            el = time.perf_counter() - t1
            if el >= args.cpu_seconds or n >= 400:
                break
        cpu = {"value": cells * n / el / 1e9, "unit": "GLUPS", "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
               "kind": "port",
               "sample": f"oracle FP32 mirror (C, OpenMP), {n} AA steps of the full 4096x1024 lattice in {el:.1f} s"}
        log(f"CPU baseline: {cpu['value'] * 1e3:.1f} MLUPS on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": case.name, "form": "aa_single_step_fp32", "W": W, "H": H, "re": 1000, "cs": 0.15,
                   "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_aa (even/odd average)", "bytes_per_cell": BYTES_PER_CELL_FP32},
        "cpu_baseline": cpu,
        "forms": forms,
        "l2_err": l2,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
