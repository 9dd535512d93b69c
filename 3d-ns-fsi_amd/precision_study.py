"""Precision study: FP32 forms of the time step against the FP64 form on one case, or on the batched sweep.

    python 3d-ns-fsi_amd/precision_study.py [--workload cylinder256] [--N 256] [--steps 1000] [--out FILE]
    python 3d-ns-fsi_amd/precision_study.py --workload sweep64 [--steps 500] [--out FILE]

Runs the case in FP64 (the form that matches the serial FP64 oracle to <= 1e-12 relative L2 per field,
see tests/test_gpu_parity.py), in FP32 with the compensated combine and in FP32 with the plain combine,
each with the body-force record on, and reports for each FP32 form the relative L2 and Linf error of every
conserved field against FP64 after the last step, and the drag coefficient (last step and mean over the
last tenth of the run) of every form. The FP64 device run is the reference at these sizes because the
serial oracle would need hours for 256^3 x 1000 steps.

With --workload sweep64 the 64 cases of the batched sweep run together in FP32 with the compensated combine
(the batched form), each case also alone in FP64, and every case's relative L2 and Linf errors per conserved
field are reported after the last step; cases whose FP64 run also goes non-finite within the steps (the
reference's scheme loses them too) are listed as such.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import cases  # noqa: E402
import nsfsi  # noqa: E402


def run(cfg, steps, label):
    t0 = time.perf_counter()
    s = nsfsi.Solver(cfg)
    s.initialize()
    s.setOption("body_force", 1)
    s.setOption("graph", 1)
    done = 0
    while done < steps:  # progress lines for long runs
        n = min(250, steps - done)
        s.marchTimeStep(n)
        done += n
        print(f"  {label}: {done}/{steps} steps", flush=True)
    U = s.mesh.conserved()
    cd = s.dragCoefficient()
    t, lv, dt = s.time()
    s.close()
    return U, cd, t, time.perf_counter() - t0


def errors(U, V):
    out = []
    for q in range(5):
        d = U[q] - V[q]
        n2 = np.linalg.norm(V[q])
        ni = np.abs(V[q]).max()
        out.append((float(np.linalg.norm(d) / n2) if n2 > 0 else float(np.linalg.norm(d)),
                    float(np.abs(d).max() / ni) if ni > 0 else float(np.abs(d).max())))
    return out


def sweep_study(a):
    cfg32 = cases.sweep64(steps=a.steps, precision="fp32")
    cfg64 = cases.sweep64(steps=a.steps, precision="fp64")
    print(f"sweep64: batched FP32 (compensated combine), {a.steps} steps", flush=True)
    t0 = time.perf_counter()
    b = nsfsi.Batch(cfg32)
    b.initialize()
    b.setOption("graph", 1)
    done = 0
    while done < a.steps:
        n = min(100, a.steps - done)
        try:
            b.step(n)
        except nsfsi.NSError as e:  # a case went non-finite: the batch still advanced every case
            if e.status != 3:
                raise
        done += n
        print(f"  batch: {done}/{a.steps} steps", flush=True)
    U32 = [b.conserved(c) for c in range(len(cfg32))]
    wall32 = time.perf_counter() - t0
    b.close()
    rows, worst = [], (0.0, 0.0)
    t0 = time.perf_counter()
    for c, cfg in enumerate(cfg64):
        s = nsfsi.Solver(cfg)
        s.initialize()
        s.setOption("graph", 1)
        try:
            s.marchTimeStep(a.steps)
        except nsfsi.NSError as e:
            if e.status != 3:
                raise
        U64 = s.mesh.conserved()
        s.close()
        fin32, fin64 = bool(np.all(np.isfinite(U32[c]))), bool(np.all(np.isfinite(U64)))
        r = {"case": c, "N": cfg.NI, "radius": cfg.body_radius, "M": cfg.M_0, "finite_fp32": fin32,
             "finite_fp64": fin64}
        if fin32 and fin64:
            errs = errors(U32[c], U64)
            r["rel_l2"] = [e[0] for e in errs]
            r["rel_linf"] = [e[1] for e in errs]
            worst = (max(worst[0], max(r["rel_l2"])), max(worst[1], max(r["rel_linf"])))
        rows.append(r)
        if c % 8 == 7:
            print(f"  FP64 cases: {c + 1}/64", flush=True)
    wall64 = time.perf_counter() - t0
    lines = [f"# Precision study: sweep64, 64 cases (grid edges 32-128), {a.steps} RK4 steps", "",
             "Batched FP32 with the compensated combine (every case equal bit for bit to its single-case FP32 run)",
             "against each case alone in FP64 (the form that matches the serial oracle within its one-ulp",
             "sensitivity). Relative L2 = ||U - U64|| / ||U64||, relative Linf = max|U - U64| / max|U64|, per conserved",
             "field after the last step; each case at its own step sizes in both forms.", "",
             f"Wall time: batched FP32 {wall32:.1f} s, the 64 FP64 runs {wall64:.1f} s. Worst over the finite cases and",
             f"fields: relative L2 {worst[0]:.3e}, relative Linf {worst[1]:.3e}.", "",
             "| case | N | sphere r | Mach | " + " | ".join(f"{f} L2 / Linf" for f in nsfsi.FIELDS[:5]) + " |",
             "|---|---|---|---|" + "---|" * 5]
    for r in rows:
        if "rel_l2" in r:
            cells = " | ".join(f"{l2:.2e} / {li:.2e}" for l2, li in zip(r["rel_l2"], r["rel_linf"]))
        else:
            cells = " | ".join(["non-finite in " + ("FP32 and FP64" if not r["finite_fp64"] else "FP32 only")] * 5)
        lines.append(f"| {r['case']} | {r['N']} | {r['radius']:.3f} | {r['M']:.3f} | {cells} |")
    text = "\n".join(lines) + "\n"
    print(text)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"workload": "sweep64", "steps": a.steps, "cases": rows}, f, indent=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cylinder256", choices=sorted(cases.WORKLOADS) + ["sweep64"])
    ap.add_argument("--N", type=int, default=None)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--out", default=None, help="also write the report (markdown) here")
    ap.add_argument("--json", default=None, help="also write the figures as JSON here")
    a = ap.parse_args()
    if a.workload == "sweep64":
        sweep_study(a)
        return
    kw = {"N": a.N} if a.N else {}
    forms = [("fp64", 0), ("fp32", 1), ("fp32", 0)]
    res = {}
    for prec, kahan in forms:
        cfg = cases.WORKLOADS[a.workload](steps=a.steps, precision=prec, **kw)
        cfg.kahan = kahan
        print(f"run {a.workload} {prec} kahan={kahan} grid {cfg.NI}x{cfg.NJ}x{cfg.NK}", flush=True)
        res[(prec, kahan)] = run(cfg, a.steps, f"{prec} kahan={kahan}")
    ref = res[("fp64", 0)]
    last = max(1, a.steps // 10)
    lines = [f"# Precision study: {a.workload}, grid {cfg.NI}x{cfg.NJ}x{cfg.NK}, {a.steps} RK4 steps", "",
             "Errors are relative to the FP64 device run (relative L2 = ||U - U64|| / ||U64||, relative Linf =",
             "max|U - U64| / max|U64|, per conserved field, after the last step). C_D = Fx / (0.5 M_0^2 A_ref).", "",
             "| form | t_end | C_D last | C_D mean (last tenth) | wall s |", "|---|---|---|---|---|"]
    names = {("fp64", 0): "FP64", ("fp32", 1): "FP32 + compensated combine", ("fp32", 0): "FP32 plain combine"}
    for k in forms:
        U, cd, t, wall = res[k]
        lines.append(f"| {names[k]} | {t:.6f} | {cd[-1]:.8f} | {cd[-last:].mean():.8f} | {wall:.1f} |")
    # momentum components are also measured against the norm of the whole momentum field: a component
    # that is zero up to the seeded perturbation has a tiny norm of its own
    mnorm = float(np.sqrt(sum(np.linalg.norm(ref[0][q]) ** 2 for q in (1, 2, 3))))
    lines += ["", "Momentum components: `L2 / ||m||` is ||U_q - U64_q|| / ||(rho_u, rho_v, rho_w)_64||; "
              "`||U64_q|| / ||m||` is the component's share of that norm.", "",
              "| form | field | rel L2 | rel Linf | L2 / \\|m\\| | \\|U64_q\\| / \\|m\\| |", "|---|---|---|---|---|---|"]
    report = {"workload": a.workload, "grid": [cfg.NI, cfg.NJ, cfg.NK], "steps": a.steps, "forms": {}}
    for k in forms[1:]:
        errs = errors(res[k][0], ref[0])
        for q, (l2, li) in enumerate(errs):
            if q in (1, 2, 3):
                dm = float(np.linalg.norm(res[k][0][q] - ref[0][q])) / mnorm
                share = float(np.linalg.norm(ref[0][q])) / mnorm
                extra = f"{dm:.3e} | {share:.3e}"
            else:
                extra = " | "
            lines.append(f"| {names[k]} | {nsfsi.FIELDS[q]} | {l2:.3e} | {li:.3e} | {extra} |")
        cdr = abs(res[k][1][-1] - ref[1][-1]) / abs(ref[1][-1])
        lines.append(f"| {names[k]} | C_D (last step) | {cdr:.3e} | |")
        report["forms"][names[k]] = {"rel_l2": [e[0] for e in errs], "rel_linf": [e[1] for e in errs],
                                     "cd_rel_err": cdr}
    for k in forms:
        report["forms"].setdefault(names[k], {})["cd_last"] = float(res[k][1][-1])
    text = "\n".join(lines) + "\n"
    print(text)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
