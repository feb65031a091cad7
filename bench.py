"""Benchmark: triangle-RHS evaluations per second of the explicit SUPG step.

Workload (default): viscous flow past a circle on an O-grid with Nr=1024,
Ntheta=2048 (2,097,152 nodes, 4,190,208 triangles), M=0.8, Re=1000, fixed
mesh -- the FP64 GPU path (--precision fp64: element residual, node sums and
RK stage fused per element patch; fp32: the same with binary32 element
arithmetic; fp64ref: element kernel and node gather in the reference's
operation order).  The other two paths are timed briefly beside it
("variants").  A step is one full time step of the
reference's main loop (deltat, RK with 4 calcRHS evaluations, ESTAB, the
two mesh-Laplace PCG solves of fluidStructure, U = U1).  value = 4 * nelem *
steps / timed seconds (every calcRHS evaluation complete, none hoisted).

The timed region starts with all inputs resident in HBM and is bracketed by
a stream synchronisation on both sides.  Per-kernel times come from HIP
events recorded on the session's own stream inside the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32|fp64ref]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "triangle-RHS evals/s, %HBM roofline; ALE step time learned vs plain at equal tol"


PRECISION = {"fp64": 0, "fp32": 1, "fp64ref": 2}
KERNEL = {0: "k_patch_stage<double> + k_halo_stage (FP64 residual, node sums, RK stage)",
          1: "k_patch_stage<float> + k_halo_stage (FP32 residual, node sums, RK stage)",
          2: "k_element_rhs (calcRHS element residual, reference order)"}
TRAFFIC = {0: "k_patch_stage_fp64", 1: "k_patch_stage_fp32", 2: "k_element_rhs_fp64ref"}


def element_bytes(nelem: int, npoin: int, precision: int = 0) -> int:
    """Algorithmic bytes of one launch of the timed element kernel, each array
    counted once at the width the variant reads or writes.

    Fused paths (k_patch_stage + k_halo_stage: residual, node sums and the RK
    stage): per element the connectivity 12 B + dNx,dNy + fac,nu,tau1..3 at
    the element width (FP64: 48 + 40 = 88 B; FP32: 24 + 20 = 44 B); per node
    U 32 + T 8 + M 8 + GAMM 8 = 56 B read, and written on average over the 4
    stages (3 x T 8 + final U1 32, RHS 32, T, VX, VY, RHO, E, P, RMACH 56) / 4
    = 36 B.

    Reference order (k_element_rhs, calcRHS element part): per element inpoel
    12 + dNx,dNy 48 + area,dtl,shoc,tau1..3 48 = 108 B; per node U 32 + T 8 =
    40 B (theta == 0 inside RK is not read)."""
    if precision == 0:
        return 100 * nelem + 92 * npoin
    if precision == 1:
        return 56 * nelem + 92 * npoin
    return 108 * nelem + 40 * npoin


def cpu_baseline(cs, seconds_budget: float = 20.0) -> dict:
    """The oracle (C restatement, OpenMP element loops with atomic scatters,
    -O3) on the host cores, on whole steps of the same workload."""
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    orc.set_threads(threads)
    o = orc.OracleSolver(cs)
    o.step()                                  # first touch / warm caches
    t0 = time.time()
    n = 0
    while True:
        o.step()
        n += 1
        if time.time() - t0 > seconds_budget / 2 or n >= 20:
            break
    dt = time.time() - t0
    return {"value": 4.0 * cs.nelem * n / dt, "unit": "triangle-RHS evals/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} full steps (after 1 untimed) of the same {cs.nelem}-triangle case, "
                      f"oracle/cfd_oracle.c -O3 -fopenmp, {threads} threads"}


def ale_compare(steps: int = 40, warmup: int = 5) -> dict:
    """ALE step time, learned vs plain, at equal PCG tolerance: the moving
    ellipse O-grid 512 x 1024 (524,288 nodes, 1,046,528 triangles, M=0.5,
    Re=5000, pitching about x = -0.25).  plain = fused FP64 path, each mesh
    solve started from the previous step's solution (the reference);
    learned = fused FP32-element path, mesh solves started from the learned
    prediction (the rates' linear extrapolation plus the MLP's correction,
    matrix cores, with fallback); extrap = the same without the network (the
    extrapolation alone, the network's ablation).  Under the relative PCG
    rules 1e-10, 1e-8 and 1e-6 (||r|| <= tol ||r0|| of a zero guess) and
    under the reference's absolute rule (r.z < 1e-10).

    Lift/drag: pressure forces after the same steps, as the difference
    relative to max(|Fx|, |Fy|) of the plain run.  The flow is chaotic at
    round-off level, so the plain run is repeated from a state with a
    one-ulp change of one node's density: that difference ("chaos floor") is
    the least any differently-rounded run can show."""
    import numpy as np
    from cfd_amd.mesh import ellipse_case
    from cfd_amd.solver import Session
    cs = ellipse_case()
    res = {}
    for rule, rel, tol in (("relative_1e-10", 1, 1e-10), ("relative_1e-8", 1, 1e-8),
                           ("relative_1e-6", 1, 1e-6), ("absolute_1e-10", 0, 0.0)):
        runs = {}
        for name, prec, ws, ulp in (("plain_fp64", 0, 0, False), ("extrap_fp32", 1, 2, False),
                                    ("learned_fp32", 1, 1, False), ("plain_fp64_ulp", 0, 0, True)):
            S = Session(cs, precision=prec, pcg_rel=rel, pcg_tol=tol, warm_start=ws)
            if ulp:
                U = S.get("U")
                U[5, 0] = np.nextafter(U[5, 0], 2.0)
                S.set("U", U)
                S.set("U1", U)
            S.step(warmup)
            S.sync()
            it0 = S.scalars()["PCG_ITSUM"]
            t0 = time.perf_counter()
            S.step(steps)
            S.sync()
            sec = time.perf_counter() - t0
            sc = S.scalars()
            runs[name] = {"ms_per_step": sec / steps * 1e3,
                          "pcg_iterations_per_step": (sc["PCG_ITSUM"] - it0) / steps,
                          "drag_fx": sc["FX1"], "lift_fy": sc["FY1"]}
            if ws:
                runs[name]["predicted_guess_used"] = f"{sc['WS_LEARNED']:.0f}/{sc['WS_SOLVES']:.0f}"
            if ws == 1:
                runs[name]["mlp_loss"] = sc["WS_LOSS"]
            S.close()
        p, q, u = runs["plain_fp64"], runs["learned_fp32"], runs.pop("plain_fp64_ulp")
        scale = max(abs(p["drag_fx"]), abs(p["lift_fy"]))
        runs["speedup"] = p["ms_per_step"] / q["ms_per_step"]
        runs["speedup_extrap"] = p["ms_per_step"] / runs["extrap_fp32"]["ms_per_step"]
        runs["network_iteration_ratio"] = (q["pcg_iterations_per_step"]
                                           / max(runs["extrap_fp32"]["pcg_iterations_per_step"], 1e-9))
        runs["drag_diff"] = abs(q["drag_fx"] - p["drag_fx"]) / scale
        runs["lift_diff"] = abs(q["lift_fy"] - p["lift_fy"]) / scale
        runs["chaos_floor_drag"] = abs(u["drag_fx"] - p["drag_fx"]) / scale
        runs["chaos_floor_lift"] = abs(u["lift_fy"] - p["lift_fy"]) / scale
        res[rule] = runs
        progress(f"ale {rule}: learned {q['ms_per_step']:.2f} ms/step, "
                 f"extrap {runs['extrap_fp32']['ms_per_step']:.2f}, plain {p['ms_per_step']:.2f}")
    res["workload"] = (f"ellipse O-grid 512x1024 ({cs.npoin} nodes, {cs.nelem} triangles), M=0.5, "
                       f"Re=5000, pitching, {warmup} untimed + {steps} timed steps")
    return res


def batch_sweep(steps: int = 100, warmup: int = 5, ncases: int = 256) -> dict:
    """The batched sweep: 256 fixed-mesh cylinder cases of uneven size (Nr in
    32..256, Ntheta in 64..512, M in [0.2, 0.85], angle of attack in [-5, 5]
    deg, Re log-uniform in [50, 5000], each at an acoustic CFL of 0.5), FP32
    element arithmetic with FP64 accumulation, every case advanced by one
    hipGraph replay per step; the patch shape is chosen by the autotuner on
    3 timed steps per candidate, then every case restarts from the free
    stream."""
    from cfd_amd.batch import Batch
    from cfd_amd.mesh import sweep_cases
    cases = sweep_cases(ncases)
    t0 = time.perf_counter()
    B = Batch(cases, precision=1)
    t_create = time.perf_counter() - t0
    cand = [(512, 256), (384, 256), (640, 320), (384, 192), (640, 256), (512, 512)]
    t0 = time.perf_counter()
    best, times = B.autotune(cand, steps=3)
    t_tune = time.perf_counter() - t0
    B.step(warmup)
    B.sync()
    t0 = time.perf_counter()
    B.step(steps)
    B.sync()
    sec = time.perf_counter() - t0
    info = B.info()
    import numpy as np
    finite = all(bool(np.isfinite(B.get(k, "U")).all()) for k in range(0, ncases, 15))
    B.close()
    return {"workload": f"{ncases} cylinder O-grid cases, {info['npoin']} nodes, {info['nelem']} "
                        f"triangles in all, fixed meshes, FP32 element / FP64 accumulation",
            "value": 4.0 * info["nelem"] * steps / sec, "unit": "triangle-RHS evals/s",
            "ms_per_step": sec / steps * 1e3, "steps": steps, "warmup": warmup,
            "patch_shape": {"elements": best[0], "threads": best[1]},
            "autotune_ms_per_step": {f"{a}x{b}": t for (a, b), t in times.items()},
            "create_s": t_create, "autotune_s": t_tune, "npatch": info["npatch"],
            "states_finite": finite}


def large_ale(steps: int = 3, warmup: int = 3, tol: float = 1e-8, maxit: int = 50000) -> dict:
    """The large ALE run: pitching ellipse O-grid 4096 x 4096 (16,777,216 nodes,
    33,546,240 triangles), M=0.7, Re=10000, the GCL volume update on; plain
    (fused FP64, previous-step guess) against learned (fused FP32 element,
    learned warm start) and extrap (the same without the network) at the
    relative PCG rule `tol` with `maxit` iterations allowed (the reference's
    1000 end every solve unconverged at this size), a few steps each (the
    mesh solves dominate; at this size one plain step takes seconds).  The
    three warm-up steps fill the three-solution history the network reads."""
    from cfd_amd.mesh import ellipse_case
    from cfd_amd.solver import Session
    cs = ellipse_case(4096, 4096, mach=0.7, re=10000.0)
    runs = {}
    for name, prec, ws in (("plain_fp64", 0, 0), ("extrap_fp32", 1, 2), ("learned_fp32", 1, 1)):
        t0 = time.perf_counter()
        S = Session(cs, precision=prec, pcg_rel=1, pcg_tol=tol, pcg_maxit=maxit, warm_start=ws,
                    gcl=1)
        t_create = time.perf_counter() - t0
        print(f"large_ale {name}: created in {t_create:.1f} s", file=sys.stderr, flush=True)
        for _ in range(warmup):
            S.step(1)
            S.sync()
            print(f"large_ale {name}: warm-up step, PCG iterations so far "
                  f"{S.scalars()['PCG_ITSUM']:.0f}", file=sys.stderr, flush=True)
        it0 = S.scalars()["PCG_ITSUM"]
        sec = 0.0
        for _ in range(steps):  # one step at a time: progress lines between
            t0 = time.perf_counter()
            S.step(1)
            S.sync()
            sec += time.perf_counter() - t0
            print(f"large_ale {name}: step, {sec:.1f} s so far", file=sys.stderr, flush=True)
        sc = S.scalars()
        runs[name] = {"ms_per_step": sec / steps * 1e3, "create_s": t_create,
                      "pcg_iterations_per_step": (sc["PCG_ITSUM"] - it0) / steps,
                      "drag_fx": sc["FX1"], "lift_fy": sc["FY1"]}
        if ws:
            runs[name]["predicted_guess_used"] = f"{sc['WS_LEARNED']:.0f}/{sc['WS_SOLVES']:.0f}"
        S.close()
    runs["speedup"] = runs["plain_fp64"]["ms_per_step"] / runs["learned_fp32"]["ms_per_step"]
    runs["speedup_extrap"] = runs["plain_fp64"]["ms_per_step"] / runs["extrap_fp32"]["ms_per_step"]
    runs["workload"] = (f"ellipse O-grid 4096x4096 ({cs.npoin} nodes, {cs.nelem} triangles), M=0.7, "
                        f"Re=10000, pitching, GCL on, relative PCG rule {tol:g} (at most {maxit} "
                        f"iterations), {warmup} untimed + {steps} timed steps")
    return runs


def progress(msg: str) -> None:
    """One line per finished section on stderr (stdout carries the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def read_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), committed under profiles/; None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.isfile(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def run_path(cs, prec: int, steps: int, warmup: int, breakdown: bool = False,
             renumber: bool = True) -> dict:
    """Times `steps` whole steps after `warmup` untimed ones; HIP events on the
    session's stream time the element kernel launches inside the timed
    region.  The fused paths run with the session's patch-order numbering
    (renumber = 1: the same results bit for bit on a fixed mesh)."""
    from cfd_amd.solver import Session
    S = Session(cs, precision=prec, renumber=1 if renumber and prec in (0, 1) else 0)
    S.step(warmup)
    S.sync()
    S.timing(True, slots=["ELEMENT"])
    S.sync()
    t0 = time.perf_counter()
    S.step(steps)
    S.sync()
    sec = time.perf_counter() - t0
    te, ne = S.kernel_time("ELEMENT")
    avg_elem_s = te / ne * 1e-3
    sc = S.scalars()
    r = {"value": 4.0 * cs.nelem * steps / sec, "ms_per_step": sec / steps * 1e3,
         "avg_elem_us": avg_elem_s * 1e6,
         "achieved": element_bytes(cs.nelem, cs.npoin, prec) / avg_elem_s / 1e9,
         "state": {"dtmin": sc["DTMIN"], "time": sc["TIME"], "fx": sc["FX1"], "fy": sc["FY1"]}}
    if breakdown:  # per-phase times: a separate, untimed pass with events on every phase
        S.timing(True)
        S.step(min(steps, 10))
        S.sync()
        kern = {}
        for k in ("ELEMENT", "STAGE", "ESTAB", "DELTAT", "PCG", "GEOM", "STEP"):
            t, n = S.kernel_time(k)
            if n:
                kern[k.lower()] = {"avg_us": round(t / n * 1e3, 2), "launches": n}
        S.timing(False)
        r["kernels"] = kern
    S.close()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--nr", type=int, default=1024)
    ap.add_argument("--nt", type=int, default=2048)
    ap.add_argument("--precision", choices=list(PRECISION), default="fp64")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the short runs of the other two element paths")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ale", action="store_true", help="skip the ALE learned-vs-plain comparison")
    ap.add_argument("--ale-steps", type=int, default=40)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sweep")
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--large-ale", action="store_true",
                    help="also run the 4096 x 4096 pitching ellipse (minutes)")
    ap.add_argument("--no-renumber", action="store_true",
                    help="fused paths in the caller's numbering (no patch-order numbering)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: the path is a single-GPU kernel set "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    from cfd_amd.mesh import cylinder_case

    cs = cylinder_case(args.nr, args.nt, mach=0.8, re=1000.0)
    prec = PRECISION[args.precision]
    rn = not args.no_renumber
    r = run_path(cs, prec, args.steps, args.warmup, breakdown=True, renumber=rn)
    DTYPE = {0: "f64", 1: "f32-element/f64-accumulate", 2: "f64"}
    PATH = {0: "fused FP64", 1: "fused FP32-element", 2: "FP64 reference-order"}
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "triangle-RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[prec],
        "data": "synthetic",
        "config": {
            "workload": f"cylinder O-grid Nr={args.nr} Ntheta={args.nt} ({cs.npoin} nodes, "
                        f"{cs.nelem} triangles), M=0.8, Re=1000, fixed mesh, {PATH[prec]} path"
                        + (", patch-order numbering inside the session" if rn and prec in (0, 1)
                           else ""),
            "nelem": cs.nelem, "npoin": cs.npoin, "calcrhs_per_step": 4,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": KERNEL[prec],
            "achieved": r["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": r["achieved"] / HBM_PEAK_GBS,
            # measured with the default numbering of each path (profiles/)
            "traffic": (read_traffic(TRAFFIC[prec] + "_cylinder1024x2048_traffic.txt")
                        if (args.nr, args.nt) == (1024, 2048) and (rn or prec == 2) else None),
            "algorithmic_bytes_per_launch": element_bytes(cs.nelem, cs.npoin, prec),
            "avg_launch_us": r["avg_elem_us"],
        },
        "kernels": r["kernels"],
        "state": r["state"],
    }
    progress(f"main path: {r['ms_per_step']:.3f} ms/step")
    if not args.no_variants:
        # the other two element paths, shorter runs of the same workload
        out["variants"] = {}
        for name, p in PRECISION.items():
            if p != prec:
                v = run_path(cs, p, 20, 20, renumber=rn)
                out["variants"][name] = {
                    "path": PATH[p], "value": v["value"], "ms_per_step": v["ms_per_step"],
                    "kernel": KERNEL[p], "avg_launch_us": v["avg_elem_us"],
                    "roofline_frac": v["achieved"] / HBM_PEAK_GBS, "steps": 20, "warmup": 20}
                progress(f"variant {name}: {v['ms_per_step']:.3f} ms/step")
    if not args.no_ale:
        out["ale"] = ale_compare(args.ale_steps)
    if not args.no_batch:
        out["batch"] = batch_sweep(args.batch_steps)
        progress("batch sweep done")
    if args.large_ale:
        out["large_ale"] = large_ale()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cs)
        progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
