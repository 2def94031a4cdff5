"""Benchmark: one MG cycle of the 2.1M-node RANS-SA bump case on one MI355X.

Workload (the metric's configuration): unit box, 128^3 hexahedra -> 12.58M
Kuhn tetrahedra, 2,146,689 nodes, 14,827,904 sides; wall z=0 with Gaussian
bump h=0.05, sigma=0.1; z graded from 1e-6 by 1.15 (69 cells) then uniform;
Re=6e6, M=0.7, alpha=1 deg, Spalart-Allmaras, HLLC-2 with Green-Gauss
gradients and LED clipping, 4 agglomerated MG levels, growing-V cycle
(multigridScheme=3, the reference default), CFL 1.  Synthetic freestream
start (no restart file exists).

A "step" is one oneCycle (MG:907-1196) plus the per-cycle monitor
(writeResiduals reductions + calculateLiftAndDrag).  value = side-loop
evaluations per second summed over levels (Gedges/s): sum_l E_l * n_l per
cycle / cycle time, where n_l counts residual evaluations on level l.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# algorithmic bytes per side of one evaluation, FP64 (counting
# rule: every distinct input read once, every output written once)
ALGO_BYTES_PER_EDGE = {"flux": 76.0, "grad": 67.9, "stage": 19.7}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_case(n, n_levels):
    from flite3d_amd import mesh
    t = time.time()
    lv = mesh.bump_levels(n, n_levels, first_cell=1e-6, ratio=1.15)
    log(f"[bench] mesh {n}^3: {[l.n_nodes for l in lv]} nodes, "
        f"{[l.n_sides for l in lv]} sides, built in {time.time() - t:.1f}s")
    return lv


def params(fp32):
    from flite3d_amd.params import InputVariables
    iv = InputVariables(reynoldsnumber=6.0e6, machnumber=0.7, alpha=1.0, fluxtype=4,
                        hllcfluxorder=2, automaticlipping=1, hllcclipping=1.0,
                        gradientmethod=1, turbulencemodel=1, wallsareisentropic=True,
                        cflnumber=1.0, turbulentcflnumber=1.0, multigridscheme=3,
                        restartnumber=0)
    return iv.to_struct(coarse_fp32=fp32)


def run_gpu(levels, fp32, steps, warmup, timed_kernel="flux_l1"):
    from flite3d_amd import device
    s = device.Solver(params(fp32), levels)
    s.init_freestream()
    c = 1
    r0 = None
    for _ in range(warmup):
        s.cycle(c)
        m = s.monitor()
        r0 = r0 or m["res"][0]
        c += 1
    evals = device.cycle_evaluations(s)
    device.timing_enable(s, timed_kernel)
    s.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.cycle(c)
        m = s.monitor()
        r0 = r0 or m["res"][0]
        c += 1
    s.sync()
    t1 = time.perf_counter()
    kms, kcnt = device.timing_get(s)
    device.timing_enable(s, None)
    res = m["res"]
    s.close()
    return dict(time=t1 - t0, evals=evals, kernel_ms=kms, kernel_launches=kcnt,
                res0=r0, res=res[0], forces=m["forces"][:4].tolist())


def run_converge(levels, fp32, drop, cap_s, every=250):
    """Cycles the solver from freestream until the monitor residual (L2 of
    the density change over the last fine smoothing step, relative to cycle
    1) falls by `drop` decades or `cap_s` seconds pass; returns the history."""
    from flite3d_amd import device
    s = device.Solver(params(fp32), levels)
    s.init_freestream()
    hist = []
    r0 = None
    c = 1
    s.sync()
    t0 = time.perf_counter()
    while True:
        s.cycle(c)
        m = s.monitor()
        r0 = r0 or m["res"][0]
        d = float(np.log10(m["res"][0] / r0)) if m["res"][0] > 0 else -99.0
        el = time.perf_counter() - t0
        if c % every == 0 or d <= -drop or el > cap_s or not np.isfinite(d):
            hist.append((c, round(el, 3), round(d, 4), float(m["forces"][1]), float(m["forces"][0])))
            log(f"[converge] {'fp32-coarse' if fp32 else 'fp64'} cycle {c} t={el:.1f}s "
                f"drop={d:.3f} drag={m['forces'][1]:.12e}")
        if d <= -drop or el > cap_s or not np.isfinite(d):
            break
        c += 1
    s.close()
    return dict(cycles=c, time=el, drop=d, reached=d <= -drop, drag=hist[-1][3],
                lift=hist[-1][4], history=hist)


def cpu_baseline(levels, budget_s=30.0):
    """The FP64 CPU restatement (oracle, one core) on one cycle of the same
    workload, or fewer fine-level smoothing steps if a cycle exceeds the
    budget; reported in the same unit (Gedges/s)."""
    import oracle
    o = oracle.Oracle(params(False), levels)
    o.init_freestream()
    E1 = levels[0].n_sides
    t0 = time.perf_counter()
    o.do_time_iterations(1, 1, True)   # one fine smoothing step = 3 residual evaluations
    t1 = time.perf_counter()
    o.close()
    return dict(value=3 * E1 / (t1 - t0) / 1e9, unit="Gedges/s", cores=1, kind="port",
                sample="one fine-level smoothing step (3 RK stages, calcDiss) of the same "
                       f"128^3 RANS-SA case, {t1 - t0:.1f} s on one host core")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--converge", type=float, default=0.0, metavar="SECONDS",
                    help="instead of the timed steps: run each precision mode to a 1e-8 "
                         "monitor drop or this many seconds, and report the histories")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU design is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)
    levels = build_case(a.n, a.levels)
    sides = [l.n_sides for l in levels]
    if a.converge > 0:
        out = {}
        for tag, fp32 in (("mixed", True), ("fp64", False)):
            out[tag] = run_converge(levels, fp32, 8.0, a.converge)
        out["drag_mixed_minus_fp64"] = out["mixed"]["drag"] - out["fp64"]["drag"]
        out["workload"] = f"RANS-SA bump box {a.n}^3 hexes, {a.levels} MG levels, growing-V"
        print(json.dumps(out), flush=True)
        return
    results = {}
    for tag, fp32 in ((("fp64", False),) if a.fp64_only else (("mixed", True), ("fp64", False))):
        log(f"[bench] GPU run {tag}: warmup {a.warmup}, steps {a.steps}")
        r = run_gpu(levels, fp32, a.steps, a.warmup)
        edges = sum(e * n for e, n in zip(sides, r["evals"]))
        r["gedges"] = edges * a.steps / r["time"] / 1e9
        r["ms_per_cycle"] = 1e3 * r["time"] / a.steps
        log(f"[bench] {tag}: {r['ms_per_cycle']:.2f} ms/cycle, {r['gedges']:.2f} Gedges/s, "
            f"evals/level {r['evals']}, res drop {np.log10(r['res'] / r['res0']):.2f}")
        results[tag] = r
    head = results["fp64" if a.fp64_only else "mixed"]
    k_avg_ms = head["kernel_ms"] / max(1, head["kernel_launches"])
    algo = ALGO_BYTES_PER_EDGE["flux"] * sides[0]
    achieved = algo / (k_avg_ms * 1e-3) / 1e9
    cpu = None if a.no_cpu else cpu_baseline(levels)
    out = {
        "metric": "Gedges/s + time to 1e-8 resid drop, FP32-MG vs FP64, 2.1M-node RANS",
        "value": round(head["gedges"], 4),
        "unit": "Gedges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_cycle"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 fine / f32 coarse" if not a.fp64_only else "f64",
        "data": "synthetic bump-box mesh, freestream start",
        "config": {"workload": f"RANS-SA bump box {a.n}^3 hexes ({levels[0].n_nodes} nodes, "
                               f"{sides[0]} sides), {a.levels} MG levels, growing-V",
                   "evaluations_per_level": head["evals"],
                   "fp64_gedges": round(results["fp64"]["gedges"], 4),
                   "fp64_ms_per_step": round(results["fp64"]["ms_per_cycle"], 3),
                   "residual_drop_log10": round(float(np.log10(head["res"] / head["res0"])), 3),
                   "time_to_1e-8_drop": None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "kernel": "k_flux<double,2,SA> (level 1)",
                     "avg_launch_ms": round(k_avg_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
