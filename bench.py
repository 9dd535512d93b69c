"""Throughput of the explicit RK4 time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sphere512] [--precision fp32|fp64]

A "step" is one full RK4 time step (4 stages, edge BCs, immersed filter and ghost update, dt and norm
reductions) of the named workload, with the state resident in HBM. Prints ONE JSON line with the metric
(M node-updates/s = NI*NJ*NK * steps / wall time of the step loop), a roofline object for the dominant
kernel (the RK4 stage kernel; algorithmic bytes = 20 field values per node per stage launch, i.e. 80
per RK4 step, timed live with HIP events on the library's stream) and a CPU baseline (the serial FP64
oracle on a bounded sample of the same workload, 1 core).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-ns-fsi_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes per stage launch from rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy), one file
# per profiled workload form (workload, precision, combine) at one tile; other forms report traffic as null.
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f"k_stage_fused_{p}_sphere512_traffic.json") for p in ("fp32", "fp64")]
VALUES_PER_NODE_STEP = 80  # 3-register RK4: stage 1 reads 5 writes 10, stages 2-3 read 15 write 10, stage 4 reads 10 writes 5


def cpu_baseline(workload, precision_unused=None):
    """Serial FP64 oracle (1 core) on the same workload at 128^3 for a fixed number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cases
    import oracle

    N, steps = 128, 8
    cfg = cases.WORKLOADS[workload](N=N, steps=steps, precision="fp64")
    o = oracle.Oracle(cfg)
    o.initialize()
    t0 = time.perf_counter()
    o.march(steps)
    el = time.perf_counter() - t0
    o.close()
    return {
        "value": N ** 3 * steps / el / 1e6,
        "unit": "M node-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{workload} geometry and physics at {N}^3, {steps} RK4 steps, serial FP64 oracle "
                  f"(g++ -O3, no FMA), {el:.1f} s; host has {os.cpu_count()} cores",
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


def bench_sweep(a):
    """The batched sweep: 64 independent FP32 cases (grid edges 32..128, own body radius, Mach, Re and dt)
    advanced together by ns_batch_step, against the same cases stepped one after another (each its own
    solver replaying its own step graph, with its own autotuned tile; also reported with the batch's
    tile)."""
    import cases
    import nsfsi

    cfgs = cases.sweep64(steps=a.steps + a.warmup)
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
    b.step(a.warmup)
    t0 = time.perf_counter()
    b.step(a.steps)  # synchronises
    el = time.perf_counter() - t0
    b.close()
    value = nodes * a.steps / el / 1e6
    # the same cases one at a time: with the batch's tile, and with each case's own autotuned tile (the
    # stronger baseline, the one the speed-up is quoted against)
    def sequential(own_tile):
        el_seq = 0.0
        for cfg in cfgs:
            s = nsfsi.Solver(cfg)
            s.initialize()
            if own_tile:
                s.autotune(2)
            else:
                s.setOption("tile_ty", tile[0])
                s.setOption("tile_kc", tile[1])
            s.setOption("graph", 0 if a.no_graph else 1)
            s.marchTimeStep(2)
            t1 = time.perf_counter()
            s.marchTimeStep(a.steps)
            el_seq += time.perf_counter() - t1
            s.close()
        return el_seq

    ms_step = el / a.steps * 1e3
    if a.batch_only:
        seq_info = None
    else:
        seq_same = sequential(False)
        seq = sequential(True)
        seq_value = nodes * a.steps / seq / 1e6
        seq_info = {"value": round(seq_value, 3), "ms_per_step": round(seq / a.steps * 1e3, 4),
                    "speedup_of_batch": round(value / seq_value, 3), "tile": "each case autotuned",
                    "same_tile_as_batch": {"value": round(nodes * a.steps / seq_same / 1e6, 3),
                                           "ms_per_step": round(seq_same / a.steps * 1e3, 4)}}
    line = {
        "metric": "M node-updates/s per RK4 step",
        "value": round(value, 3),
        "unit": "M node-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (64 seeded cases: freestream + Gaussian perturbation)",
        "config": {"workload": "sweep64", "cases": len(cfgs), "nodes": nodes,
                   "grid_edges": sorted({c.NI for c in cfgs}), "compensated_combine": not a.no_kahan,
                   "tile_rows_x_zplanes": list(tile), "hip_graph": not a.no_graph},
        "roofline": {"bound": "hbm", "achieved": round(nodes * VALUES_PER_NODE_STEP * 4 / (ms_step * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(nodes * VALUES_PER_NODE_STEP * 4 / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "whole batched step (algorithmic bytes / step time)"},
        "sequential_per_case": seq_info,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="sphere512")
    ap.add_argument("--precision", default=None)
    ap.add_argument("--N", type=int, default=None, help="override the grid edge")
    ap.add_argument("--no-cpu", action="store_true")
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

    import cases
    import nsfsi

    mk = cases.WORKLOADS[a.workload]
    kw = {}
    if a.N:
        kw["N"] = a.N
    prec = a.precision or ("fp32" if a.workload == "sphere512" else "fp64")
    cfg = mk(steps=a.steps + a.warmup, precision=prec, **kw)
    cfg.conv_history = 2
    cfg.kahan = 0 if a.no_kahan else 1
    nodes = cfg.NI * cfg.NJ * cfg.NK
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("stage_kernel", 1 if a.kernel == "fused" else 0)
    if a.kernel != "fused":
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
    s.marchTimeStep(a.warmup)
    t0 = time.perf_counter()
    s.marchTimeStep(a.steps)  # returns after a device synchronisation
    el = time.perf_counter() - t0
    ms_step = el / a.steps * 1e3
    value = nodes * a.steps / el / 1e6

    # dominant kernel timed live with HIP events on the library's stream
    ms, launches = s.profile(max(2, min(5, a.steps)))
    eb = 4 if prec == "fp32" else 8
    alg_bytes = VALUES_PER_NODE_STEP / 4 * eb * nodes  # per stage launch
    achieved = alg_bytes / (ms[0] * 1e-3) / 1e9 if ms[0] > 0 else None
    g, f, _ = s.mesh.counts()
    kahan = bool(prec == "fp32" and cfg.kahan)
    traffic, traffic_tile = measured_traffic(a.workload, prec, kahan, a.kernel)
    traffic = round(traffic) if traffic else None
    line = {
        "metric": "M node-updates/s per RK4 step",
        "value": round(value, 3),
        "unit": "M node-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64",
        "data": "synthetic (seeded freestream + perturbation initial field)",
        "config": {"workload": a.workload, "grid": [cfg.NI, cfg.NJ, cfg.NK], "precision": prec,
                   "compensated_combine": kahan,
                   "ghosts": g, "filter_nodes": f, "stage_kernel": a.kernel,
                   "tile_rows_x_zplanes": list(tile) if tile else None,
                   "hip_graph": not a.no_graph,
                   "fused_scans": bool(a.kernel == "fused" and (a.fused_scans is None or a.fused_scans))},
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
            "kernel_ms": round(ms[0], 4),
            "step_breakdown_ms": {"stage_kernel": round(ms[0] * launches[0], 3),
                                  "primitive_pass": round(ms[1] * launches[1], 3),
                                  "boundaries": round(ms[2] * launches[2], 3),
                                  "reductions": round(ms[3] * launches[3], 3)},
            "step_frac": round(nodes * VALUES_PER_NODE_STEP * eb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
    }
    s.close()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration << 20)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(a.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
