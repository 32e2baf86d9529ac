"""Throughput of the DGSEM rhs! on one MI355X, in GDOF.RHS/s.

Workload (default): dry Euler (energy form) with gravity on the cubed-sphere
shell, 6 x 16^2 trees x 8 layers = 12,288 elements, polydeg 4 (125 nodes per
element, 1,536,000 nodes), flux_kennedy_gruber + Waruszewski volume,
local Lax-Friedrichs (FluxPlusDissipation(flux_kennedy_gruber,
DissipationLocalLaxFriedrichs())) + Waruszewski surface, slip walls,
isothermal 250 K background with 10 seeded 1 K bumps, 100 steps.  A step is
one CarpenterKennedy2N54 step (5 rhs! evaluations, each one fused launch) on
the device-resident state.  GDOF counts nodes (Trixi's convention), not
nodes x variables.  The FP64 form is `value`; the mixed form (FP32 fluxes with
Kahan sums, FP64 state, pressure, RK and hydrostatic background) is reported
beside it with its error against the FP64 form.
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
FP32_VECTOR_PEAK_TFLOPS = 157.3
METRIC = "GDOF\u00b7RHS/s, mixed vs FP64 on 1 GPU at stated rel. err; % of VALU roofline"

# HBM traffic and executed FP64 flops of the FP64 stage kernel are measured by
# this script itself (measure_counters below: rocprofv3 --pmc passes over a
# short child run of the same workload).  Where rocprofv3 is missing or a pass
# fails, these figures of the committed profile
# (profiles/e3_stage_shell16x8_p4_pmc.txt) stand in, and the JSON line says so:
# FETCH_SIZE and WRITE_SIZE in KiB per dispatch, the FETCH_SIZE scale (the
# stage issues 128 B read requests, counted as 64 B, as for a plain device
# copy of known size) and the executed flops per node
# (64 x (ADD + MUL + TRANS) + 128 x FMA wave instructions / nodes).
STAGE_FETCH_KIB, STAGE_WRITE_KIB, FETCH_FACTOR = 2.020e5, 1.258e5, 2.0
EXEC_FP64_FLOPS_PER_NODE = 1383.0
COUNTER_PASSES = (("FETCH_SIZE",), ("WRITE_SIZE",),
                  ("SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
                   "SQ_INSTS_VALU_TRANS_F64"))
CALIBRATION_COPY_BYTES = 256 << 20
# Algorithmic FP64 work of one rhs! per node of the shell 6x16^2x8, polydeg 4
# (KG + Waruszewski volume, KG + LLF + Waruszewski surface, slip walls): the
# oracle's own code run with an operation-counting scalar (oracle/count_flops.py;
# + - * / sqrt log exp pow one each), every node pair of a line evaluated once
# as the reference does
ALG_FP64_FLOPS_PER_NODE = 1680.4
EXEC_MIXED_FLOPS_PER_NODE = {"fp32": 1703.0, "fp64": 298.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def counter_probe(a):
    """the child run the counter passes profile: one device copy of
    CALIBRATION_COPY_BYTES (the largest copy of the run, which scales the
    byte counters), then the FP64 shell stage in the given workgroup shape
    (one warm-up step and two more, fifteen stage dispatches)"""
    import torch
    import trixiatmo_hip as T
    src = torch.ones(CALIBRATION_COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    mesh, eq, solver, bcs, u0 = T.workloads.shell(a.trees, a.layers, polydeg=a.polydeg, seed=0,
                                                  precision="fp64")
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    semi.set_kernel_variant(a.counter_probe)
    semi.set_state(u0)
    dt = semi.max_dt(0.5)
    semi.step(dt, 3)
    semi.synchronize()
    print("counter probe done", flush=True)


def measure_counters(a, variant, nodes, time_limit=240):
    """rocprofv3 --pmc passes (one per group of COUNTER_PASSES) over
    counter_probe in a child process; returns per-dispatch means of the stage
    kernel's counters and the copy's, or None and the reason"""
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not on PATH"
    if "rocprofiler" in os.environ.get("LD_PRELOAD", "") or any(
            k.startswith("ROCPROF") for k in os.environ):
        # this run is itself under a profiler, whose preloaded library has
        # already initialised the GPU: a nested profiler could not start the
        # probe, and the outer profiler's own counters are the record
        return None, "bench.py runs under a profiler"
    stage = f"e3_stage_direct<{a.polydeg + 1}, 1, 1, double, false, {variant + 1}, double>"
    values = {}
    with tempfile.TemporaryDirectory(prefix="atmo_pmc_") as tmp:
        for i, counters in enumerate(COUNTER_PASSES):
            out_dir = os.path.join(tmp, f"pass{i}")
            cmd = [prof, "--pmc", *counters, "-f", "csv",
                   "--kernel-include-regex", "e3_stage_direct|copyBuffer", "-d", out_dir, "-o", "run",
                   "--", sys.executable, os.path.abspath(__file__), "--counter-probe", str(variant),
                   "--trees", str(a.trees), "--layers", str(a.layers), "--polydeg", str(a.polydeg)]
            log(f"counters: pass {i + 1} of {len(COUNTER_PASSES)} ({' '.join(counters)})")
            p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, cwd=tmp,
                                 start_new_session=True, env=dict(os.environ, TMPDIR=tmp))
            try:
                text = p.communicate(timeout=time_limit)[0].decode(errors="replace")
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                return None, f"pass {i + 1} exceeded {time_limit} s"
            files = glob.glob(os.path.join(out_dir, "**", "*counter_collection.csv"), recursive=True)
            # rocprofv3 can end with a fault of its own after a complete run:
            # the probe's last line and the result file decide
            if "counter probe done" not in text or not files:
                return None, f"pass {i + 1} failed (exit {p.returncode}): {text.strip()[-300:]}"
            got, why = counter_means(files, stage, counters)
            if got is None:
                return None, f"pass {i + 1}: {why}"
            values.update(got)
    return counter_summary(values, stage, nodes), None


def counter_means(files, stage, counters):
    """per counter: the mean over the stage kernel's dispatches (rows of one
    dispatch summed) and, for the byte counters, the largest copy dispatch's
    value, from rocprofv3's counter_collection.csv files"""
    import csv
    per = {}
    for f in files:
        with open(f) as fh:
            for n, r in enumerate(csv.DictReader(fh)):
                name, c = r["Kernel_Name"], r["Counter_Name"]
                key = "stage" if stage in name else ("copy" if "copyBuffer" in name else None)
                if key is None:
                    continue
                d = r.get("Dispatch_Id") or r.get("Correlation_Id") or f"{f}:{n}"
                per.setdefault((key, c), {}).setdefault(d, 0.0)
                per[(key, c)][d] += float(r["Counter_Value"])
    values = {}
    for c in counters:
        st = list(per.get(("stage", c), {}).values())
        if not st:
            return None, f"no {c} value for {stage}"
        values[c] = sum(st) / len(st)
        values[c + "_dispatches"] = len(st)
        if c in ("FETCH_SIZE", "WRITE_SIZE"):
            cp = list(per.get(("copy", c), {}).values())
            if not cp:
                return None, f"no {c} value for the calibration copy"
            values["copy_" + c] = max(cp)
    return values, None


def counter_summary(values, stage, nodes):
    """HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE, each scaled by
    what the counter reads for the calibration copy of known size, and the
    executed FP64 flops per node"""
    read_factor = CALIBRATION_COPY_BYTES / (values["copy_FETCH_SIZE"] * 1024.0)
    write_factor = CALIBRATION_COPY_BYTES / (values["copy_WRITE_SIZE"] * 1024.0)
    flops = (64.0 * (values["SQ_INSTS_VALU_ADD_F64"] + values["SQ_INSTS_VALU_MUL_F64"] +
                     values["SQ_INSTS_VALU_TRANS_F64"]) + 128.0 * values["SQ_INSTS_VALU_FMA_F64"])
    return {"kernel": stage, "dispatches": values["FETCH_SIZE_dispatches"],
            "fetch_kib": values["FETCH_SIZE"], "write_kib": values["WRITE_SIZE"],
            "copy_fetch_kib": values["copy_FETCH_SIZE"], "copy_write_kib": values["copy_WRITE_SIZE"],
            "read_factor": read_factor, "write_factor": write_factor,
            "traffic": (values["FETCH_SIZE"] * read_factor + values["WRITE_SIZE"] * write_factor) * 1024.0,
            "exec_fp64_flops_per_node": flops / nodes}


def cpu_baseline(mesh, eq, solver, u0, du_gpu=None, max_seconds=30.0, min_calls=20,
                 others=None):
    """the oracle (FP64 C++ OpenMP restatement of the reference rhs!) on the
    same workload, timed on this host's cores: 5 warm-up calls, then the
    median of at least `min_calls` calls (more while under `max_seconds`).
    With the GPU's rhs of the same state (du_gpu) it also reports, as the
    checker, the GPU's error against the oracle per conserved variable, on
    the scale of the summed terms and on the scale of max |du|."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, params_for
    o = Oracle()
    p = params_for((mesh, eq, solver, 0.0))
    for _ in range(5):
        ref = o.rhs(p, mesh, u0)
    times = []
    t_start = time.perf_counter()
    while len(times) < min_calls or (time.perf_counter() - t_start < max_seconds and
                                     len(times) < 200):
        t0 = time.perf_counter()
        o.rhs(p, mesh, u0)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    nodes = mesh.n_elements * mesh.nnodes ** 3
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"value": nodes / med / 1e9, "unit": "GDOF*RHS/s", "cores": threads,
           "nproc": os.cpu_count(), "kind": "port",
           "sample": f"median of {len(times)} full FP64 rhs! calls of the same workload (the FP64 "
                     f"state after the timed steps) after 5 warm-up calls ({med * 1e3:.1f} ms per "
                     f"call, {threads} OpenMP threads)"}
    if du_gpu is not None:
        mag = o.rhs_magnitude(p, mesh, u0)
        nv = 6
        d = np.abs((du_gpu - ref).reshape(nv, -1, order="F")).max(axis=1)
        term = d / np.maximum(np.abs(mag.reshape(nv, -1, order="F")).max(axis=1), 1e-300)
        plain = d / np.maximum(np.abs(ref.reshape(nv, -1, order="F")).max(axis=1), 1e-300)
        out["gpu_fp64_rhs_rel_err_vs_oracle"] = {"term_scale": [float(x) for x in term[:5]],
                                                  "du_scale": [float(x) for x in plain[:5]]}
        for name, du_other in (others or {}).items():
            d = np.abs((du_other - ref).reshape(nv, -1, order="F")).max(axis=1)
            out[f"gpu_{name}_rhs_rel_err_vs_oracle"] = {
                "term_scale": [float(x) for x in (d / np.maximum(np.abs(
                    mag.reshape(nv, -1, order="F")).max(axis=1), 1e-300))[:5]],
                "du_scale": [float(x) for x in (d / np.maximum(np.abs(
                    ref.reshape(nv, -1, order="F")).max(axis=1), 1e-300))[:5]]}
    return out


def bench_covariant(steps, warmup, trees=32, polydeg=3, n_members=32):
    """covariant shallow water on the cubed sphere (radius 6.371e6 m, g = 9.81,
    polydeg 3, flux_ec volume, EC + LLF surface): one member of 6 x 32^2
    elements, and an ensemble of 32 members with 6 k^2 elements each, k drawn
    from {8, 16, 24, 32}, advanced by one launch per RK stage; each with FP64
    geometry and with FP32 geometry storage (FP64 arithmetic)"""
    import numpy as np
    import trixiatmo_hip as T
    W = T.workloads

    def timed(semi, run):
        run(warmup)
        semi.synchronize()
        t0 = time.perf_counter()
        run(steps)
        semi.synchronize()
        el = time.perf_counter() - t0
        semi.enable_timing(True)
        run(max(2, steps // 2))
        launches, total_ms = semi.kernel_times(0)
        semi.enable_timing(False)
        return el, total_ms / launches / 1e3

    out = {}
    ref_du = {}
    for prec, tag in (("fp64", ""), ("fp32_geometry", "_fp32_geometry")):
        # algorithmic bytes per node and stage: u, k read, k and u_out written
        # (3 FP64 each), the geometry values the stage reads (G, h_s, the
        # Christoffel symbols, z / r, and with FP32 storage J; G^-1, and in
        # FP64 J, are recomputed from G), and per face node (4 n of them
        # per n^2 nodes) the 2 x 2 basis map and the neighbour link (int32)
        gb = 8 if prec == "fp64" else 4
        # (the zonal jets have no bottom topography: the h_s row is not read)
        bpn = 8 * 4 * 3 + (8 * 10 if gb == 8 else 4 * 11) + (gb * 4 + 4) * 4 / (polydeg + 1)
        case = W.zonal_jet(trees, polydeg, seed=0, precision=prec)
        semi = case.semi()
        u0 = case.u0()
        # rhs on a perturbed state, against the FP64 form (which matches the
        # oracle to 1e-12 on the scale of the terms)
        rng = np.random.default_rng(5)
        up = u0 * (1.0 + 0.01 * rng.standard_normal(u0.shape))
        du = semi.rhs(None, up).reshape(3, -1, order="F")
        if prec == "fp64":
            ref_du["single"] = du
        semi.set_state(u0)
        shape, shape_ms = semi.autotune()
        dt = semi.max_dt(0.4)
        el, avg_s = timed(semi, lambda k: semi.step(dt, k))
        nodes = semi.ndofs
        rec = {"workload": f"6x{trees}^2 elements, polydeg {polydeg}", "precision": prec,
               "nodes": nodes, "value": nodes * 5 * steps / el / 1e9, "unit": "GDOF*RHS/s",
               "ms_per_step": el / steps * 1e3, "avg_launch_us": avg_s * 1e6,
               "hbm_achieved_gbs": bpn * nodes / avg_s / 1e9, "bytes_per_node": bpn,
               "shape": shape, "shape_us": [round(x * 1e3, 2) for x in shape_ms],
               "finite": bool(np.all(np.isfinite(semi.get_state())))}
        if prec != "fp64":
            d = np.abs(du - ref_du["single"]).max(axis=1)
            rec["rhs_rel_err_vs_fp64"] = (d / np.abs(ref_du["single"]).max(axis=1)).tolist()
            rec["speedup_vs_fp64"] = out["single"]["avg_launch_us"] / rec["avg_launch_us"]
        out["single" + tag] = rec
        del semi
        members, eq, solver, src = W.zonal_jet_ensemble(n_members=n_members, polydeg=polydeg,
                                                        seed=0, precision=prec)
        ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
        u0 = ens.compute_coefficients()
        ens.set_state(u0)
        shape, shape_ms = ens.autotune()
        dts = ens.max_dt_members(0.4)
        el, avg_s = timed(ens, lambda k: ens.step_members(dts, k))
        nodes = ens.ndofs
        rec = {"workload": f"{n_members} members, 6k^2 elements, k in {{8,16,24,32}}",
               "precision": prec, "sizes": W.ensemble_sizes(n_members, 0), "nodes": nodes,
               "value": nodes * 5 * steps / el / 1e9, "unit": "GDOF*RHS/s",
               "ms_per_step": el / steps * 1e3, "avg_launch_us": avg_s * 1e6,
               "hbm_achieved_gbs": bpn * nodes / avg_s / 1e9,
               "shape": shape, "shape_us": [round(x * 1e3, 2) for x in shape_ms],
               "finite": bool(np.all(np.isfinite(ens.get_state())))}
        if prec != "fp64":
            rec["speedup_vs_fp64"] = out["ensemble"]["avg_launch_us"] / rec["avg_launch_us"]
        out["ensemble" + tag] = rec
        del ens
    out["precision_note"] = ("fp32_geometry stores the stage's metric terms, Christoffel "
                             "symbols and face maps in FP32 and computes in FP64; FP32 flux "
                             "arithmetic is refused for the covariant form (no faster than FP64 "
                             "on this GPU's non-packed vector units, and the stage is bound by "
                             "bytes)")
    return out


def _time_loops(semi, u0, cfl, steps, check_every=256):
    """the StepsizeCallback loop over the span that `steps` steps of the
    initial step size cover, run as a host loop (atmo_max_dt + atmo_step per
    step) and resident on the device (atmo_solve); returns ms per step of each"""
    import numpy as np
    semi.set_state(u0)
    t_end = steps * semi.max_dt(cfl)
    out = {}
    for mode in ("host", "resident"):
        for rep in range(2):  # the first pass warms up (graph capture, autotuning)
            semi.set_state(u0)
            semi.set_time(0.0)
            semi.synchronize()
            t0 = time.perf_counter()
            if mode == "host":
                t, n = 0.0, 0
                while t < t_end:
                    dt = min(semi.max_dt(cfl), t_end - t)
                    semi.step(dt, 1)
                    t += dt
                    n += 1
                n = np.array([n])
            else:
                n, _ = semi.solve_resident(t_end, cfl, check_every)
            semi.synchronize()
            el = time.perf_counter() - t0
        out[mode] = {"steps": int(n.max()), "ms_per_step": el / max(1, int(n.max())) * 1e3,
                     "seconds": el}
    out["speedup"] = out["host"]["ms_per_step"] / out["resident"]["ms_per_step"]
    return out


def bench_time_loop(steps, trees, layers, polydeg, cov_trees=32, n_members=32):
    """time to solution of the StepsizeCallback loop, host-driven against
    resident, on the 3D shell, the covariant single member, (resident only,
    members on their own step sizes) the covariant ensemble and the rainy
    bubble (a quarter of the steps)"""
    import trixiatmo_hip as T
    W = T.workloads
    out = {}
    mesh, eq, solver, bcs, u0 = W.shell(trees, layers, polydeg=polydeg, seed=0)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    out["euler3d_shell"] = _time_loops(semi, u0, 0.5, steps)
    del semi
    case = W.zonal_jet(cov_trees, 3, seed=0)
    out["covariant_single"] = _time_loops(case.semi(), case.u0(), 0.4, steps)
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=n_members, polydeg=3, seed=0)
    ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
    u0 = ens.compute_coefficients()
    ens.set_state(u0)
    t_end = steps * ens.max_dt_members(0.4).min()
    for rep in range(2):
        ens.set_state(u0)
        ens.set_time(0.0)
        ens.synchronize()
        t0 = time.perf_counter()
        n, _ = ens.solve_resident(t_end, 0.4, 256)
        ens.synchronize()
        el = time.perf_counter() - t0
    out["covariant_ensemble_resident"] = {
        "steps_per_member_min_max": [int(n.min()), int(n.max())], "seconds": el,
        "ms_per_step": el / int(n.max()) * 1e3}
    del ens
    mesh, eq, solver, src, lim, u0, _ = W.rainy_bubble(seed=0)
    rainy = T.SemidiscretizationRainy(mesh, eq, solver, source_terms=src, stage_limiter=lim)
    out["rainy_bubble"] = _time_loops(rainy, u0, 0.5, max(2, steps // 4))
    for k, v in out.items():
        log(f"[time loop {k}] {v}")
    return out


def bench_setup(trees, layers, polydeg):
    """set-up time of the 3D shell: mesh (node coordinates and connectivity on
    the host), then the handle with the metric terms computed on the device
    from the node coordinates, or on the host and uploaded"""
    import trixiatmo_hip as T
    W = T.workloads
    t0 = time.perf_counter()
    mesh, eq, solver, bcs, _ = W.shell(trees, layers, polydeg=polydeg, seed=0)
    t_mesh = time.perf_counter() - t0
    out = {"workload": f"shell 6x{trees}^2x{layers}, polydeg {polydeg}", "mesh_seconds": t_mesh}
    for on_device in (True, False):
        mesh._metric = None
        t0 = time.perf_counter()
        semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs,
                                              metric_terms_on_device=on_device)
        semi.synchronize()
        out["handle_seconds_" + ("device_metric" if on_device else "host_metric")] = \
            time.perf_counter() - t0
        del semi
    log(f"[setup] {out}")
    return out


def bench_precision_stress(steps=1000):
    """the entropy-conservative shell (6 x 8^2 x 4, polydeg 7) in three forms:
    FP64, FP32 fluxes with Kahan sums, FP32 fluxes with plain sums; each runs
    `steps` CarpenterKennedy2N54 steps from the same state and reports the
    relative RHS error against FP64 on the FP64 end state, and the mass and
    entropy drift over the run"""
    import numpy as np
    import trixiatmo_hip as T
    forms = {}
    u_end64 = None
    semis = {}
    dt = None
    for prec in ("fp64", "fp32_kahan", "fp32"):
        mesh, eq, solver, bcs, u0 = T.workloads.precision_stress(prec)
        semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
        semi.set_state(u0)
        if dt is None:
            dt = semi.max_dt(0.5)
        i0 = semi.integrals(fresh_rhs=True)
        t0 = time.perf_counter()
        semi.step(dt, steps)
        semi.synchronize()
        el = time.perf_counter() - t0
        u_end = semi.get_state()
        i1 = semi.integrals(fresh_rhs=True)
        forms[prec] = {"mass_drift": (i1[0] - i0[0]) / i0[0],
                       "entropy_drift": (i1[1] - i0[1]) / abs(i0[1]),
                       "seconds": el, "ms_per_step": el / steps * 1e3,
                       "finite": bool(np.all(np.isfinite(u_end)))}
        log(f"[stress {prec}] {steps} steps in {el:.2f} s, entropy drift "
            f"{forms[prec]['entropy_drift']:.3e}")
        if prec == "fp64":
            u_end64 = u_end
        semis[prec] = semi
    d64 = semis["fp64"].rhs(None, u_end64).reshape(6, -1, order="F")
    for prec in ("fp32_kahan", "fp32"):
        d = semis[prec].rhs(None, u_end64).reshape(6, -1, order="F")
        forms[prec]["rhs_rel_err_vs_fp64_per_var"] = [
            float(x) for x in np.abs(d - d64)[:5].max(axis=1) / np.abs(d64)[:5].max(axis=1)]
    semis.clear()
    nodes = mesh.n_elements * mesh.nnodes ** 3
    return {"workload": "shell 6x8^2x4, polydeg 7, flux_ranocha + Waruszewski volume and surface",
            "nodes": nodes, "steps": steps, "dt": dt, "forms": forms}


def bench_rainy(steps, warmup, train_seeds=4, mlp_epochs=20, saturated_training=False):
    """rainy Euler 2D, 256 x 128 cells, polydeg 3 (524,288 nodes): stage-limited
    CarpenterKennedy2N54 with the FP64 saturation Newton solve (tol 1e-9, the
    reference's NonlinearSolveDG default) fused into every stage, unseeded
    and seeded by the MLP trained on 4 x 524,288 = 2,097,152 states of this
    configuration (4 further seeds, each after 2 steps)"""
    import numpy as np
    import torch
    import trixiatmo_hip as T
    from trixiatmo_hip.rainy import harvest_features, train_saturation_mlp
    W = T.workloads
    mesh, eq, solver, src, lim, u0, info = W.rainy_bubble(seed=0)
    semi = T.SemidiscretizationRainy(mesh, eq, solver, source_terms=src, stage_limiter=lim)
    nodes = semi.ndofs
    out = {"workload": f"moist bubble {mesh.nx}x{mesh.ny} cells, polydeg {mesh.polydeg}, "
                       "flux_ec_rain + LLF, source_terms_rainy, Newton tol 1e-9",
           "nodes": nodes, "relative_humidity": info["relative_humidity"]}

    def run(tag):
        semi.set_state(u0)
        dt = semi.max_dt(0.5)
        semi.step(dt, warmup)
        semi.synchronize()
        t0 = time.perf_counter()
        semi.step(dt, steps)
        semi.synchronize()
        el = time.perf_counter() - t0
        semi.enable_timing(True)
        semi.step(dt, 2)
        launches, total_ms = semi.kernel_times(0)
        semi.enable_timing(False)
        semi.enable_newton_stats(True)
        semi.step(dt, 2)
        st = semi.newton_stats()
        semi.enable_newton_stats(False)
        finite = bool(np.all(np.isfinite(semi.get_state())))
        log(f"[rainy {tag}] {steps} steps in {el:.4f} s, Newton {st['updates_per_node'] / 10:.3f} "
            "updates per node and stage")
        return {"value": nodes * 5 * steps / el / 1e9, "unit": "GDOF*RHS/s",
                "ms_per_step": el / steps * 1e3, "avg_stage_launch_us": total_ms / launches * 1e3,
                "newton_updates_per_node_per_stage": st["updates"] / (10.0 * nodes),
                "newton_max_updates": st["max_updates"], "newton_failures": st["failures"],
                "fallbacks": st["fallbacks"], "fallback_rate": st["fallbacks"] / (10.0 * nodes),
                "finite": finite}, dt

    out["unseeded"], dt = run("unseeded")
    u_unseeded = semi.get_state()
    # algorithmic HBM bytes of one fused stage per node: read u and the RK
    # register k, write k and the new state (9 FP64 each); the uniform
    # Cartesian cells need no per-node geometry
    stage_bytes = 8 * (9 + 9 + 9 + 9)
    avg_s = out["unseeded"]["avg_stage_launch_us"] * 1e-6
    out["roofline"] = {"bound": "hbm", "kernel": "rainy_stage<4>", "bytes_per_node": stage_bytes,
                       "achieved": stage_bytes * nodes / avg_s / 1e9, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": stage_bytes * nodes / avg_s / 1e9 / HBM_PEAK_GBS,
                       "note": "the fused Newton solve of the stage limiter makes this stage "
                               "VALU-heavy (exp and log per update); unseeded launch"}
    semi.seed_stage_limiter("closed_form")
    out["closed_form_seeded"], _ = run("closed_form_seeded")
    semi.seed_stage_limiter(False)
    # training states: further seeds of this configuration (relative humidity
    # below 1); saturated_training adds a supersaturated variant (relative
    # humidity 1.05), which measured worse on this configuration's solves
    states = [W.rainy_bubble(seed=s)[5] for s in range(1, train_seeds + 1)]
    if saturated_training:
        states.append(W.rainy_bubble(seed=train_seeds + 1, relative_humidity=1.05)[5])
    t0 = time.perf_counter()
    x, y = harvest_features(semi, states, steps=2, dt=dt)
    mlp, hist = train_saturation_mlp(x, y, epochs=mlp_epochs)
    torch.cuda.synchronize()
    out["mlp"] = {"samples": int(x.shape[0]), "epochs": mlp_epochs,
                  "final_batch_loss": hist[-1], "first_epoch_loss": hist[0],
                  "train_seconds": time.perf_counter() - t0}
    semi.attach_mlp(mlp)
    semi.seed_stage_limiter(True)
    out["seeded"], _ = run("seeded")
    u_seeded = semi.get_state()
    semi.seed_stage_limiter(False)
    a, b = u_seeded.reshape(-1, 9), u_unseeded.reshape(-1, 9)
    out["seeded_vs_unseeded_state_rel_diff"] = float(np.abs(a - b).max() / np.abs(b).max())
    out["seeded_vs_unseeded_T_rel_diff"] = float(np.abs(a[:, 8] - b[:, 8]).max() / np.abs(b[:, 8]).max())
    # cold solve: the saturation variables reset to a fixed guess (T = 250 K,
    # no vapour, no cloud), then the default and the seeded solve
    cold = u_unseeded.reshape(-1, 9).copy()
    cold[:, 6:9] = np.array([0.0, 0.0, 250.0])
    semi.set_state(cold.reshape(-1))
    st0 = semi.newton(seeded=False)
    semi.set_state(cold.reshape(-1))
    st1 = semi.newton(seeded=True)
    semi.set_state(cold.reshape(-1))
    st2 = semi.newton(seeded="closed_form")
    out["cold_solve"] = {"guess": "T = 250 K, rho_vapour = rho_cloud = 0",
                         "default_updates_per_node": st0["updates_per_node"],
                         "seeded_updates_per_node": st1["updates_per_node"],
                         "closed_form_updates_per_node": st2["updates_per_node"],
                         "seeded_fallbacks": st1["fallbacks"],
                         "failures": [st0["failures"], st1["failures"], st2["failures"]]}
    # a saturated case: a bubble at relative humidity 1.1 (a seed not used in
    # training), solved from its own guess (all moisture as vapour) and cold
    semi.set_state(W.rainy_bubble(seed=train_seeds + 2, relative_humidity=1.1)[5])
    sat = semi.get_state().reshape(-1, 9)
    res = {}
    for tag, guess in (("own_guess", None), ("cold", np.array([0.0, 0.0, 250.0]))):
        v = sat.copy()
        if guess is not None:
            v[:, 6:9] = guess
        semi.set_state(v.reshape(-1))
        s0 = semi.newton(seeded=False)
        u_def = semi.get_state()
        semi.set_state(v.reshape(-1))
        s1 = semi.newton(seeded=True)
        u_seed = semi.get_state()
        semi.set_state(v.reshape(-1))
        s2 = semi.newton(seeded="closed_form")
        res[tag] = {"default_updates_per_node": s0["updates_per_node"],
                    "seeded_updates_per_node": s1["updates_per_node"],
                    "closed_form_updates_per_node": s2["updates_per_node"],
                    "seeded_fallbacks": s1["fallbacks"],
                    "failures": [s0["failures"], s1["failures"], s2["failures"]],
                    "state_rel_diff": float(np.abs(u_seed - u_def).max() / np.abs(u_def).max())}
    cloud = semi.get_state().reshape(-1, 9)[:, 7]
    res["cloudy_node_fraction"] = float(np.mean(cloud > 0))
    out["saturated_solve"] = res
    del semi
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--timing-steps", type=int, default=4,
                    help="steps run with per-launch HIP-event timing after the timed region")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--trees", type=int, default=16)
    ap.add_argument("--layers", type=int, default=8)
    ap.add_argument("--polydeg", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-covariant", action="store_true",
                    help="skip the covariant shallow-water lines")
    ap.add_argument("--no-rainy", action="store_true", help="skip the rainy-Euler line")
    ap.add_argument("--stress-steps", type=int, default=1000,
                    help="steps of the precision-stress forms (0 skips them)")
    ap.add_argument("--rainy-steps", type=int, default=20)
    ap.add_argument("--loop-steps", type=int, default=200,
                    help="steps of the host-driven and resident time loops (0 skips them)")
    ap.add_argument("--traffic-kib", type=float, nargs=3, default=None,
                    metavar=("FETCH", "WRITE", "FACTOR"),
                    help="rocprofv3 FETCH_SIZE / WRITE_SIZE of the stage kernel and the FETCH scale")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (to scale HBM counters)")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (traffic and executed flops then "
                         "come from the committed profile)")
    ap.add_argument("--counter-probe", type=int, default=None, help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.counter_probe is not None:
        counter_probe(a)
        return
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
    semis, results = {}, {}
    dt = None
    fp64_objects = None
    for prec in ("fp64", "fp32_kahan", "fp32_geometry"):
        mesh, eq, solver, bcs, u0 = T.workloads.shell(a.trees, a.layers, polydeg=a.polydeg, seed=0,
                                                      precision=prec)
        semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
        nodes = mesh.n_elements * mesh.nnodes ** 3
        if prec == "fp64":
            # the CPU baseline and its parity check run on these (with the
            # evolved FP64 state, below)
            fp64_objects = (mesh, eq, solver)
        semi.set_state(u0)
        if dt is None:
            dt = semi.max_dt(0.5)
        shape, shape_ms = semi.autotune()
        i0 = semi.integrals(fresh_rhs=True)
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
        semi.step(dt, a.timing_steps)
        launches, total_ms = semi.kernel_times(0)
        semi.enable_timing(False)
        u_end = semi.get_state()
        i1 = semi.integrals(fresh_rhs=True)
        results[prec] = dict(el=el, gdof=nodes * 5 * a.steps / el / 1e9, shape=shape,
                             shape_ms=[float(x) for x in shape_ms[:2]],
                             avg_s=total_ms / launches / 1e3, u_end=u_end,
                             ok=bool(np.all(np.isfinite(u_end))),
                             mass_drift=(i1[0] - i0[0]) / i0[0],
                             entropy_drift=(i1[1] - i0[1]) / abs(i0[1]))
        semis[prec] = semi
    f64, mix = results["fp64"], results["fp32_kahan"]
    gdof, el, avg_s, ok = f64["gdof"], f64["el"], f64["avg_s"], f64["ok"] and mix["ok"]
    n_total = a.warmup + a.steps + a.timing_steps
    # RHS error of the mixed form against the FP64 form on the same evolved
    # state (the FP64 state after all steps), and the state error after the run
    du_end = semis["fp64"].rhs(None, f64["u_end"])
    d64 = du_end.reshape(6, -1, order="F")
    dmx = semis["fp32_kahan"].rhs(None, f64["u_end"]).reshape(6, -1, order="F")
    du_geo = semis["fp32_geometry"].rhs(None, f64["u_end"])
    geo = results["fp32_geometry"]
    geo_err = (np.abs(du_geo.reshape(6, -1, order="F") - d64).max(axis=1) /
               np.abs(d64).max(axis=1))[:5]
    ugeo = geo["u_end"].reshape(6, -1, order="F")
    rhs_err = (np.abs(dmx - d64).max(axis=1) / np.abs(d64).max(axis=1))[:5]
    u64 = f64["u_end"].reshape(6, -1, order="F")
    umx = mix["u_end"].reshape(6, -1, order="F")
    u_start = u0.reshape(6, -1, order="F")
    state_err = (np.abs(umx - u64).max(axis=1) / np.abs(u64).max(axis=1))[:5]
    change_err = (np.abs(umx - u64).max(axis=1) /
                  np.maximum(np.abs(u64 - u_start).max(axis=1), 1e-300))[:5]
    semis.clear()
    # algorithmic HBM bytes of one fused stage per node: read u (6 FP64),
    # write u_out, read + write the RK register k (5 FP64 each: phi has no
    # tendency and is the same in both state buffers), read Ja (9) and 1/J (1);
    # the mixed path reads Ja in FP32
    bytes_per_node = 8 * (6 + 5 + 5 + 5 + 9 + 1)
    bytes_mixed = 8 * (6 + 5 + 5 + 5 + 1) + 4 * 9
    bytes_geo = bytes_mixed
    achieved = bytes_per_node * nodes / avg_s / 1e9
    counters, counter_note = None, "skipped (--no-counters)"
    if not a.no_counters and not a.traffic_kib:
        try:
            counters, counter_note = measure_counters(a, f64["shape"], nodes)
        except Exception as exc:  # a profiler or file format this script does not know
            counters, counter_note = None, f"counter passes failed: {exc!r}"[:300]
    if counters is not None:
        traffic, exec_flops = counters["traffic"], counters["exec_fp64_flops_per_node"]
        counter_source = "rocprofv3 --pmc passes in this run"
    else:
        fetch_kib, write_kib, fetch_factor = a.traffic_kib or (STAGE_FETCH_KIB, STAGE_WRITE_KIB,
                                                               FETCH_FACTOR)
        traffic = (fetch_kib * fetch_factor + write_kib) * 1024.0
        exec_flops = EXEC_FP64_FLOPS_PER_NODE
        counter_source = ("--traffic-kib" if a.traffic_kib else
                          f"profiles/e3_stage_shell16x8_p4_pmc.txt ({counter_note})")
    valu64 = exec_flops * nodes / avg_s / 1e12
    valu_mix = (EXEC_MIXED_FLOPS_PER_NODE["fp32"] / FP32_VECTOR_PEAK_TFLOPS +
                EXEC_MIXED_FLOPS_PER_NODE["fp64"] / FP64_VECTOR_PEAK_TFLOPS) * nodes / mix["avg_s"] / 1e12

    cov = None
    if not a.no_covariant:
        log("covariant shallow water")
        cov = bench_covariant(a.steps, a.warmup)

    rainy = None
    if not a.no_rainy:
        log("rainy Euler")
        rainy = bench_rainy(a.rainy_steps, min(a.warmup, 3))

    setup = bench_setup(a.trees, a.layers, a.polydeg)

    loops = None
    if a.loop_steps > 0:
        log("time loops")
        loops = bench_time_loop(a.loop_steps, a.trees, a.layers, a.polydeg)

    stress = None
    if a.stress_steps > 0:
        log("precision stress")
        stress = bench_precision_stress(a.stress_steps)

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(*fp64_objects, f64["u_end"], du_end,
                           others={"fp32_geometry": du_geo})

    out = {
        "metric": METRIC,
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
                               f"polydeg {a.polydeg}, flux_kennedy_gruber + Waruszewski volume, "
                               "FluxPlusDissipation(flux_kennedy_gruber, "
                               "DissipationLocalLaxFriedrichs()) + Waruszewski surface (the "
                               "reference's flux_lax_friedrichs reads phi as rho_e and diverges "
                               "on this shell), slip walls, CarpenterKennedy2N54",
                   "elements": mesh.n_elements, "nodes": nodes, "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"e3_stage_direct<{a.polydeg + 1},1,1,double,false,{f64['shape'] + 1}>",
                     "elements_per_workgroup": f64["shape"] + 1,
                     "autotuner_ms_per_shape": f64["shape_ms"], "avg_launch_us": avg_s * 1e6,
                     "bytes_per_node": bytes_per_node, "traffic_bytes_per_node": traffic / nodes,
                     "traffic_source": counter_source, "counters": counters},
        "valu": {"fp64_executed_tflops": valu64, "executed_flops_per_node": exec_flops,
                 "frac_of_fp64_vector_peak": valu64 / FP64_VECTOR_PEAK_TFLOPS,
                 "algorithmic_flops_per_node": ALG_FP64_FLOPS_PER_NODE,
                 "fp64_algorithmic_tflops": ALG_FP64_FLOPS_PER_NODE * nodes / avg_s / 1e12,
                 "algorithmic_frac_of_fp64_vector_peak":
                     ALG_FP64_FLOPS_PER_NODE * nodes / avg_s / 1e12 / FP64_VECTOR_PEAK_TFLOPS,
                 "mixed_frac_of_vector_peak": valu_mix,
                 "note": "executed VALU flops from rocprofv3 SQ_INSTS_VALU_*_F64 counters "
                         f"({counter_source})"},
        "cpu_baseline": cpu,
        "mixed": {"precision": "fp32 fluxes + Kahan sums, fp64 state/pressure/RK",
                  "value": mix["gdof"], "speedup_vs_fp64": mix["gdof"] / gdof,
                  "ms_per_step": mix["el"] / a.steps * 1e3,
                  "avg_launch_us": mix["avg_s"] * 1e6,
                  "hbm_achieved_gbs": bytes_mixed * nodes / mix["avg_s"] / 1e9,
                  "rhs_rel_err_vs_fp64_per_var": [float(x) for x in rhs_err],
                  "state_rel_err_vs_fp64_after_steps": [float(x) for x in state_err],
                  "state_err_rel_to_change": [float(x) for x in change_err],
                  "steps_compared": n_total,
                  "mass_drift": mix["mass_drift"], "entropy_drift": mix["entropy_drift"]},
        "fp32_geometry": {"precision": "fp64 arithmetic, contravariant vectors stored in fp32",
                          "value": geo["gdof"], "speedup_vs_fp64": geo["gdof"] / gdof,
                          "ms_per_step": geo["el"] / a.steps * 1e3,
                          "avg_launch_us": geo["avg_s"] * 1e6,
                          "bytes_per_node": bytes_geo,
                          "hbm_achieved_gbs": bytes_geo * nodes / geo["avg_s"] / 1e9,
                          "rhs_rel_err_vs_fp64_per_var": [float(x) for x in geo_err],
                          "state_rel_err_vs_fp64_after_steps":
                              [float(x) for x in (np.abs(ugeo - u64).max(axis=1) /
                                                  np.abs(u64).max(axis=1))[:5]],
                          "state_err_rel_to_change":
                              [float(x) for x in (np.abs(ugeo - u64).max(axis=1) /
                                                  np.maximum(np.abs(u64 - u_start).max(axis=1),
                                                             1e-300))[:5]],
                          "mass_drift": geo["mass_drift"], "entropy_drift": geo["entropy_drift"]},
        "fp64_drift": {"mass": f64["mass_drift"], "entropy": f64["entropy_drift"],
                       "steps": n_total},
        "covariant_swe": cov,
        "rainy_euler": rainy,
        "precision_stress": stress,
        "time_loop": loops,
        "setup": setup,
        "state_finite": ok,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
