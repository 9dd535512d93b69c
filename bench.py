"""Throughput of the explicit RK4 time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--windows 3] [--workload sphere512] [--precision fp32|fp64]

A "step" is one full RK4 time step (4 stages, edge BCs, immersed filter and ghost update, dt and norm
reductions) of the named workload, with the state resident in HBM. After W untimed steps, K timed steps run
as consecutive windows, each bracketed by a device synchronisation; the median window is reported. Prints
ONE JSON line with the metric (M node-updates/s = NI*NJ*NK * steps / wall time of the step loop), a roofline
object for the dominant kernel (the RK4 stage kernel; algorithmic bytes = 20 field values per node per stage
launch, i.e. 80 per RK4 step, timed live with HIP events on the library's stream), by default the other
forms the metric names (`forms`: the FP64 step on the same workload with its own roofline, the FP32/FP64
ratio, and the 64-case batched sweep against the same cases stepped one after another) and a CPU baseline
(the serial FP64 oracle, 1 core, on the 48^3 sphere for its 200 steps and the 256^3 cylinder for 5 steps).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-ns-fsi_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per stage launch from rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy), one file
# per profiled workload form (workload, precision, combine) at one tile; other forms report traffic as null.
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f"k_stage_fused_{p}_sphere512_traffic.json") for p in ("fp32", "fp64")]
VALUES_PER_NODE_STEP = 80  # 3-register RK4: stage 1 reads 5 writes 10, stages 2-3 read 15 write 10, stage 4 reads 10 writes 5


def cpu_baseline(samples=(("sphere48", 48, 200), ("cylinder256", 256, 5))):
    """The serial FP64 oracle (1 core) on the reference's CPU-runnable configuration (the 48^3 sphere, its
    full 200 steps) and on the 256^3 z-cylinder precision study for 5 steps; `value` is the first."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cases
    import oracle

    out = []
    for workload, N, steps in samples:
        cfg = cases.WORKLOADS[workload](N=N, steps=steps, precision="fp64")
        o = oracle.Oracle(cfg)
        o.initialize()
        t0 = time.perf_counter()
        o.march(steps)
        el = time.perf_counter() - t0
        o.close()
        out.append({"workload": workload, "grid": N, "steps": steps, "seconds": round(el, 2),
                    "value": round(N ** 3 * steps / el / 1e6, 4)})
    first = out[0]
    return {
        "value": first["value"],
        "unit": "M node-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": "; ".join(f"{x['workload']} {x['grid']}^3, {x['steps']} RK4 steps: {x['value']} M/s ({x['seconds']} s)"
                            for x in out) + f"; serial FP64 oracle (g++ -O3, no FMA), one thread; host has "
                                            f"{os.cpu_count()} cores",
        "samples": out,
    }


def measured_traffic(workload, precision, kahan, kernel):
    """(bytes per stage launch, tile it was measured at) for this workload form, or (None, None)."""
    for path in TRAFFIC_FILES:
        try:
            with open(path) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if (kernel == "fused" and t.get("workload") == workload and t.get("precision") == precision
                and bool(t.get("compensated_combine")) == bool(kahan)):
            return t.get("traffic_bytes_per_stage_launch"), list(t.get("tile", []))
    return None, None


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes (a kernel of known traffic for scaling FETCH/WRITE counters)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)


def hip_sync():
    """Device-wide synchronisation (the timed windows are bracketed by it)."""
    import ctypes as C

    global _HIP
    if "_HIP" not in globals():
        _HIP = C.CDLL("libamdhip64.so")
    assert _HIP.hipDeviceSynchronize() == 0


def timed_windows(step, warmup, steps, windows):
    """warmup untimed steps, then `steps` timed steps split into `windows` consecutive windows, each bracketed
    by a device synchronisation; returns the per-window seconds per step (median first) and all of them."""
    step(warmup)
    windows = max(1, min(windows, steps))
    sizes = [steps // windows + (1 if q < steps % windows else 0) for q in range(windows)]
    per = []
    for n in sizes:
        hip_sync()
        t0 = time.perf_counter()
        step(n)  # returns after synchronising the library's stream
        hip_sync()
        per.append((time.perf_counter() - t0) / n)
    return float(np.median(per)), per


def sweep_numbers(a, steps, warmup):
    """The batched sweep: 64 independent FP32 cases (grid edges 32..128, own body radius, Mach, Re and dt)
    advanced together by ns_batch_step, against the same cases stepped one after another (each its own
    solver replaying its own step graph, with its own autotuned tile)."""
    import cases
    import nsfsi

    cfgs = cases.sweep64(steps=steps + warmup + 2)
    for c in cfgs:
        c.conv_history = 2
        c.kahan = 0 if a.no_kahan else 1
    nodes = sum(c.NI * c.NJ * c.NK for c in cfgs)
    b = nsfsi.Batch(cfgs)
    b.initialize()
    if a.tile:
        ty, kc = (int(x) for x in a.tile.split(","))
        b.setOption("tile_ty", ty)
        b.setOption("tile_kc", kc)
        tile = b.tile()
    else:
        tile = b.autotune(2)
    b.setOption("graph", 0 if a.no_graph else 1)
    sec, per = timed_windows(b.step, warmup, steps, a.windows)
    b.close()
    value = nodes / sec / 1e6
    out = {"value": round(value, 3), "ms_per_step": round(sec * 1e3, 4),
           "windows_ms_per_step": [round(x * 1e3, 4) for x in per], "cases": len(cfgs), "nodes": nodes,
           "grid_edges": sorted({c.NI for c in cfgs}), "tile_rows_x_zplanes": list(tile),
           "hip_graph": not a.no_graph, "compensated_combine": not a.no_kahan,
           "roofline": {"bound": "hbm", "achieved": round(nodes * VALUES_PER_NODE_STEP * 4 / sec / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(nodes * VALUES_PER_NODE_STEP * 4 / sec / 1e9 / HBM_PEAK_GBS, 4),
                        "kernel": "whole batched step (algorithmic bytes / step time)"}}
    if not a.batch_only:
        # the same cases one after another, each with its own autotuned tile and step graph
        el_seq = 0.0
        for cfg in cfgs:
            s = nsfsi.Solver(cfg)
            s.initialize()
            s.autotune(2)
            s.setOption("graph", 0 if a.no_graph else 1)
            s.marchTimeStep(2)
            hip_sync()
            t1 = time.perf_counter()
            s.marchTimeStep(steps)
            hip_sync()
            el_seq += time.perf_counter() - t1
            s.close()
        seq_value = nodes * steps / el_seq / 1e6
        out["sequential_per_case"] = {"value": round(seq_value, 3), "ms_per_step": round(el_seq / steps * 1e3, 4),
                                      "tile": "each case autotuned", "hip_graph": not a.no_graph}
        out["speedup_of_batch"] = round(value / seq_value, 3)
    return out


def bench_sweep(a):
    """--workload sweep64: the batched sweep as the headline line."""
    sw = sweep_numbers(a, a.steps, a.warmup)
    line = {
        "metric": "M node-updates/s per RK4 step",
        "value": sw["value"],
        "unit": "M node-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sw["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (64 seeded cases: freestream + Gaussian perturbation)",
        "config": {"workload": "sweep64", **{k: sw[k] for k in ("cases", "nodes", "grid_edges", "compensated_combine",
                                                                  "tile_rows_x_zplanes", "hip_graph")}},
        "roofline": {**sw["roofline"], "traffic": None},
        "windows_ms_per_step": sw["windows_ms_per_step"],
        "sequential_per_case": sw.get("sequential_per_case"),
        "speedup_of_batch": sw.get("speedup_of_batch"),
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def single_case(a, workload, prec, kernel, steps, warmup, kahan):
    """One case of `workload` at precision `prec`: the step rate (median of the timed windows) and the
    stage kernel's mean launch time from HIP events on the library's stream."""
    import cases
    import nsfsi

    kw = {"N": a.N} if a.N else {}
    cfg = cases.WORKLOADS[workload](steps=steps + warmup + 8, precision=prec, **kw)
    cfg.conv_history = 2
    cfg.kahan = 1 if kahan else 0
    nodes = cfg.NI * cfg.NJ * cfg.NK
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", 1 if kernel == "fused" else 0)
    if kernel != "fused":
        tile = None
    elif a.tile:
        ty, kc = (int(x) for x in a.tile.split(","))
        s.setOption("tile_ty", ty)
        s.setOption("tile_kc", kc)
        tile = s.tile()
    else:
        tile = s.autotune(2)
    s.setOption("graph", 0 if a.no_graph else 1)
    if a.fused_scans is not None:
        s.setOption("fused_scans", a.fused_scans)
    sec, per = timed_windows(s.marchTimeStep, warmup, steps, a.windows)
    ms, launches = s.profile(max(2, min(5, steps)))
    g, f, _ = s.mesh.counts()
    s.close()
    eb = 4 if prec == "fp32" else 8
    alg_bytes = VALUES_PER_NODE_STEP / 4 * eb * nodes  # per stage launch
    achieved = alg_bytes / (ms[0] * 1e-3) / 1e9 if ms[0] > 0 else None
    return {
        "cfg": cfg, "nodes": nodes, "tile": tile, "ghosts": g, "filter_nodes": f,
        "value": nodes / sec / 1e6, "ms_per_step": sec * 1e3, "windows": [x * 1e3 for x in per],
        "kernel_ms": ms[0], "achieved": achieved,
        "breakdown": {"stage_kernel": round(ms[0] * launches[0], 3), "primitive_pass": round(ms[1] * launches[1], 3),
                      "boundaries": round(ms[2] * launches[2], 3), "reductions": round(ms[3] * launches[3], 3)},
        "step_frac": nodes * VALUES_PER_NODE_STEP * eb / sec / 1e9 / HBM_PEAK_GBS,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed steps, split into --windows windows")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--windows", type=int, default=3, help="timed windows; the median window is reported")
    ap.add_argument("--workload", default="sphere512")
    ap.add_argument("--precision", default=None)
    ap.add_argument("--N", type=int, default=None, help="override the grid edge")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default=None,
                    help="comma list of extra forms in the line: fp64, sweep64 (default both for the 512^3 sphere)")
    ap.add_argument("--kernel", default="fused", choices=["fused", "unfused"])
    ap.add_argument("--fused-scans", type=int, default=None, choices=[0, 1],
                    help="end-of-step scans inside the last stage kernel (default: the library's choice)")
    ap.add_argument("--no-graph", action="store_true", help="launch kernel by kernel instead of replaying a HIP graph")
    ap.add_argument("--no-kahan", action="store_true", help="FP32: plain stage combine instead of the compensated one")
    ap.add_argument("--tile", default=None, help="fixed fused tile 'rows,zplanes' instead of autotuning")
    ap.add_argument("--batch-only", action="store_true",
                    help="sweep64: time the batched step only, not the cases one after another")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, one device-to-device copy of MIB MiB (counter calibration under rocprofv3)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build's path runs on a single MI355X (no multi-GPU sharding); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.workload == "sweep64":
        bench_sweep(a)
        return

    prec = a.precision or ("fp32" if a.workload == "sphere512" else "fp64")
    kahan = prec == "fp32" and not a.no_kahan
    r = single_case(a, a.workload, prec, a.kernel, a.steps, a.warmup, kahan)
    cfg, nodes, tile = r["cfg"], r["nodes"], r["tile"]
    achieved = r["achieved"]
    traffic, traffic_tile = measured_traffic(a.workload, prec, kahan, a.kernel)
    traffic = round(traffic) if traffic else None
    line = {
        "metric": "M node-updates/s per RK4 step",
        "value": round(r["value"], 3),
        "unit": "M node-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64",
        "data": "synthetic (seeded freestream + perturbation initial field)",
        "config": {"workload": a.workload, "grid": [cfg.NI, cfg.NJ, cfg.NK], "precision": prec,
                   "compensated_combine": kahan,
                   "ghosts": r["ghosts"], "filter_nodes": r["filter_nodes"], "stage_kernel": a.kernel,
                   "tile_rows_x_zplanes": list(tile) if tile else None,
                   "hip_graph": not a.no_graph,
                   "fused_scans": bool(a.kernel == "fused" and (a.fused_scans is None or a.fused_scans))},
        "windows_ms_per_step": [round(x, 4) for x in r["windows"]],
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            # the tile the traffic was measured at, when the autotuner chose another this run
            "traffic_tile": traffic_tile if traffic and tile and list(tile) != traffic_tile else None,
            "kernel": "k_stage_fused (RK4 stage)" if a.kernel == "fused" else "k_rhs_plain (RK4 stage)",
            "kernel_ms": round(r["kernel_ms"], 4),
            "step_breakdown_ms": r["breakdown"],
            "step_frac": round(r["step_frac"], 4),
        },
    }
    forms = a.forms.split(",") if a.forms is not None else (["fp64", "sweep64"] if a.workload == "sphere512" and
                                                             prec == "fp32" and a.kernel == "fused" else [])
    forms = [f for f in forms if f]
    extra = {}
    if "fp64" in forms:
        r64 = single_case(a, a.workload, "fp64", a.kernel, a.steps, a.warmup, False)
        extra["fp64_fused"] = {
            "value": round(r64["value"], 3), "ms_per_step": round(r64["ms_per_step"], 4), "dtype": "f64",
            "windows_ms_per_step": [round(x, 4) for x in r64["windows"]],
            "tile_rows_x_zplanes": list(r64["tile"]) if r64["tile"] else None,
            "roofline": {"bound": "hbm", "achieved": round(r64["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(r64["achieved"] / HBM_PEAK_GBS, 4), "kernel_ms": round(r64["kernel_ms"], 4),
                         "step_frac": round(r64["step_frac"], 4),
                         "algorithmic_bytes_per_node_step": VALUES_PER_NODE_STEP * 8}}
        extra["fp32_over_fp64"] = round(r["value"] / r64["value"], 3)
    if "sweep64" in forms:
        sw_steps = max(6, min(a.steps, 30))
        extra["sweep64"] = sweep_numbers(a, sw_steps, a.warmup)
    if extra:
        line["forms"] = extra
    if a.copy_calibration:
        copy_calibration(a.copy_calibration << 20)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
