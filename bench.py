"""Throughput of the DGSEM rhs! on one MI355X, in GDOF.RHS/s.

Workload (default): dry Euler (energy form) with gravity on the cubed-sphere
shell, 6 x 16^2 trees x 8 layers = 12,288 elements, polydeg 4 (125 nodes per
element, 1,536,000 nodes), flux_kennedy_gruber + Waruszewski volume,
flux_lax_friedrichs + Waruszewski surface, slip walls, isothermal 250 K
background with 10 seeded 1 K bumps.  A step is one CarpenterKennedy2N54 step
(5 rhs! evaluations, each one fused launch) on the device-resident state.
GDOF counts nodes (Trixi's convention), not nodes x variables.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "trixiatmo.jl_amd"))

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(mesh, eq, solver, u0, seconds=15.0):
    """the oracle (FP64 C++ OpenMP restatement of the reference rhs!) on the
    same workload, timed on this host's cores"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, params_for
    o = Oracle()
    p = params_for((mesh, eq, solver, 0.0))
    o.rhs(p, mesh, u0)  # warm-up
    n_calls, t0 = 0, time.perf_counter()
    while True:
        o.rhs(p, mesh, u0)
        n_calls += 1
        el = time.perf_counter() - t0
        if el > seconds or n_calls >= 20:
            break
    nodes = mesh.n_elements * mesh.nnodes ** 3
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": nodes * n_calls / el / 1e9, "unit": "GDOF*RHS/s", "cores": cores,
            "kind": "port",
            "sample": f"{n_calls} full rhs! calls of the same workload ({el:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--trees", type=int, default=16)
    ap.add_argument("--layers", type=int, default=8)
    ap.add_argument("--polydeg", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (to scale HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {a.gpus}); the multi-GPU path "
              "(element partition with a face-halo exchange per stage) is not built.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import trixiatmo_hip as T

    if a.copy_calibration_mib:
        import torch
        nbytes = a.copy_calibration_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: 3 device copies of {nbytes} bytes each")

    log("building workload")
    runs = {}
    for prec in ("fp64", "fp32_kahan"):
        mesh, eq, solver, bcs, u0 = T.workloads.shell(a.trees, a.layers, polydeg=a.polydeg, seed=0,
                                                      precision=prec)
        semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
        nodes = mesh.n_elements * mesh.nnodes ** 3
        du_one = semi.rhs(None, u0)
        semi.set_state(u0)
        dt = semi.max_dt(0.5)
        log(f"[{prec}] warmup {a.warmup} steps, dt = {dt:.4g} s")
        semi.step(dt, a.warmup)
        semi.synchronize()
        t0 = time.perf_counter()
        semi.step(dt, a.steps)
        semi.synchronize()
        el = time.perf_counter() - t0
        log(f"[{prec}] timed {a.steps} steps in {el:.4f} s")
        # per-launch duration of the fused stage kernel from HIP events on its stream
        semi.enable_timing(True)
        semi.step(dt, max(2, a.steps // 2))
        launches, total_ms = semi.kernel_times(0)
        semi.enable_timing(False)
        ok = bool(np.all(np.isfinite(semi.get_state())))
        runs[prec] = dict(el=el, gdof=nodes * 5 * a.steps / el / 1e9, avg_s=total_ms / launches / 1e3,
                          du=du_one, ok=ok)
        del semi
    f64, mix = runs["fp64"], runs["fp32_kahan"]
    gdof, el, avg_s, ok = f64["gdof"], f64["el"], f64["avg_s"], f64["ok"] and mix["ok"]
    # algorithmic HBM bytes of one fused stage per node: read u, write u_out,
    # read + write the RK register k (6 FP64 each), read Ja (9) and 1/J (1);
    # the mixed path reads Ja in FP32
    bytes_per_node = 8 * (6 + 6 + 6 + 6 + 9 + 1)
    bytes_mixed = 8 * (6 + 6 + 6 + 6 + 1) + 4 * 9
    achieved = bytes_per_node * nodes / avg_s / 1e9
    d64 = f64["du"].reshape(6, -1, order="F")
    dmx = mix["du"].reshape(6, -1, order="F")
    rel_err = (np.abs(dmx - d64).max(axis=1) / np.abs(d64).max(axis=1))[:5]

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(mesh, eq, solver, u0)

    out = {
        "metric": "GDOF*RHS/s (FP64 rhs! on 1 GPU)",
        "value": gdof,
        "unit": "GDOF*RHS/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded isothermal shell with 10 temperature bumps)",
        "config": {"workload": f"dry Euler energy+gravity, cubed-sphere shell 6x{a.trees}^2x{a.layers}, "
                               f"polydeg {a.polydeg}, KG+Waruszewski volume, LLF+Waruszewski surface, "
                               "CarpenterKennedy2N54",
                   "elements": mesh.n_elements, "nodes": nodes, "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "e3_stage<5,1,1,double,false>", "avg_launch_us": avg_s * 1e6,
                     "bytes_per_node": bytes_per_node},
        "cpu_baseline": cpu,
        "mixed": {"precision": "fp32 fluxes + Kahan sums, fp64 state/pressure/RK",
                  "value": mix["gdof"], "speedup_vs_fp64": mix["gdof"] / gdof,
                  "ms_per_step": mix["el"] / a.steps * 1e3,
                  "avg_launch_us": mix["avg_s"] * 1e6,
                  "hbm_achieved_gbs": bytes_mixed * nodes / mix["avg_s"] / 1e9,
                  "rhs_rel_err_vs_fp64_per_var": [float(x) for x in rel_err]},
        "state_finite": ok,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
