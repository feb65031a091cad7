"""Benchmark: triangle-RHS evaluations per second of the explicit SUPG step.

Workload (default): viscous flow past a circle on an O-grid with Nr=1024,
Ntheta=2048 (2,097,152 nodes, 4,190,208 triangles), M=0.8, Re=1000, fixed
mesh, at an acoustic CFL of 0.5 -- the FP64 GPU path (--precision fp64:
element residual, node sums and RK stage fused per element patch; fp32: the
same with binary32 element arithmetic; fp64ref: element kernel and node
gather in the reference's operation order).  The other two paths are timed
briefly beside it ("variants").  A step is one full time step of the
reference's main loop (deltat, RK with 4 calcRHS evaluations, ESTAB, FORCES
and FORCE_VISC, the two mesh-Laplace PCG solves of fluidStructure, U = U1).
value = 4 * nelem * steps / timed seconds (every calcRHS evaluation complete,
none hoisted).

Every path is timed on a developed state: the FP64 path first runs the
configuration's 500 steps from the free stream (untimed) and goes on in the
same session with its warm-up and timed steps; the state after those is
loaded into each other timed session and is the accuracy section's FP64
reference.  The timed region starts with all inputs
resident in HBM and is bracketed by a stream synchronisation on both sides.
Per-kernel times come from HIP events recorded on the session's own stream
inside the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|fp32|fp64ref]

Beside the headline line (all in the same JSON object): "accuracy" (the FP32
element path against the FP64 path after the same steps: L2 density error,
drag and lift coefficients, with the round-off chaos floor), "ale" (the
pitching ellipse 512 x 1024, learned warm start + FP32 elements against the
plain FP64 path at equal PCG tolerance), "batch" (256 uneven cases),
"large_ale" (the 4096 x 4096 pitching ellipse, a few steps) and
"cpu_baseline" (the compiled reference solver on the host cores).
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
    the element width (FP64: 48 + 40 = 88 B; FP32: 24 + 20 = 44 B; the six
    gradients the residual uses, although the kernel reads four and forms
    the first node's pair as minus the sum of the other two); per node
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


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    """The host threads this job may use: OMP_NUM_THREADS when the
    environment sets it (the job's share of a shared host), else every core."""
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)


def cpu_baseline(cs) -> dict:
    """The reference solver itself (its Fortran compiled with flang -O3
    -fopenmp from the unmodified sources, oracle/_ref/ref_harness_omp: the
    reference's loader, start-up and time step, OpenMP element loops with
    atomic scatters) on the host cores, over whole steps of the same case, on
    all the job's threads and on one; rate = 4 * nelem * steps / the
    reference's own calcRHS timer (calcrhs_t), plus its wall time per step.
    When the compiled reference is not there, the C restatement of the same
    loops (oracle/cfd_oracle.c) is timed instead (kind "port")."""
    import subprocess
    import tempfile
    threads = cpu_threads()
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "ref_harness_omp")
    info = {"unit": "triangle-RHS evals/s", "cores": threads, "host_cpus": os.cpu_count(),
            "cpu_model": cpu_model()}
    if os.path.isfile(ref_bin) and os.access(ref_bin, os.X_OK):
        from cfd_amd.mesh import write_reference_inputs
        with tempfile.TemporaryDirectory(prefix="cfdref_") as d:
            write_reference_inputs(cs, d)
            runs = {}
            for th, n in ((threads, 3), (1, 1)):
                env = dict(os.environ, OMP_NUM_THREADS=str(th))
                r = subprocess.run([ref_bin, "time", str(n)], cwd=d, env=env, capture_output=True,
                                   text=True, timeout=600)
                line = [l for l in r.stdout.splitlines() if l.startswith("timed_steps")]
                if r.returncode != 0 or not line:
                    raise RuntimeError(f"reference timing failed: {r.stdout[-500:]} {r.stderr[-500:]}")
                f = line[0].split()
                steps, wall, calc = int(f[1]), float(f[3]), float(f[5])
                runs[th] = {"steps": steps, "calcrhs_s": calc, "wall_s": wall,
                            "value": 4.0 * cs.nelem * steps / calc,
                            "step_s": wall / steps}
                progress(f"cpu baseline (reference) on {th} threads: {wall / steps:.2f} s/step")
        a, o = runs[threads], runs[1]
        info.update(value=a["value"], kind="reference",
                    sample=f"the reference's own time step (flang -O3 -fopenmp), {a['steps']} whole "
                           f"steps of the same {cs.nelem}-triangle case on {threads} threads and "
                           f"{o['steps']} on 1 thread; the reference's calcrhs_t timer",
                    step_s=a["step_s"], one_core={"value": o["value"], "step_s": o["step_s"]})
        return info
    from oracle import oracle as orc
    out = {}
    for th, n in ((threads, 3), (1, 1)):
        orc.set_threads(th)
        o = orc.OracleSolver(cs)
        o.step()                                  # first touch
        t0 = time.time()
        for _ in range(n):
            o.step()
        out[th] = (time.time() - t0) / n
        del o
    orc.set_threads(1)
    info.update(value=4.0 * cs.nelem / out[threads], kind="port",
                sample=f"whole steps of the same {cs.nelem}-triangle case, oracle/cfd_oracle.c -O3 "
                       f"-fopenmp, {threads} threads and 1 thread (rate over whole steps)",
                step_s=out[threads], one_core={"value": 4.0 * cs.nelem / out[1], "step_s": out[1]})
    return info


def force_coefficients(sc: dict, cs, length: float = 1.0) -> dict:
    """Drag and lift of force set 1 (the body).  FORCE_VISC's F_VX, F_VY
    (ns2DComp.ALE.f90:819-893) integrate the traction of the full stress
    tensor (pressure and viscous) with the edge normals pointing into the
    body, i.e. the force the body exerts on the fluid; the force on the body
    is its negative.  Coefficients over 1/2 rho_inf |U_inf|^2 times the body
    length (diameter 1, chord 1).  Also the pressure-only force of FORCES."""
    q = 0.5 * cs.rho_inf * (cs.u_inf ** 2 + cs.v_inf ** 2) * length
    return {"cd": -sc["FVX1"] / q, "cl": -sc["FVY1"] / q, "cd_pressure": sc["FX1"] / q,
            "cl_pressure": sc["FY1"] / q}


PITCH_AMP = 5.0 * 3.141592653589793 / 180.0   # alpha(t) = 5 deg sin(2 k t), k = 0.1
PITCH_OMEGA = 0.2


def near_body_nodes(cs, k: int = 5) -> list:
    """k nodes (0-based) spread over the first node ring off the body (ring 1
    of the O-grid, not a boundary node): a one-ulp change there reaches the
    body forces within a step, so the chaos floor is measured where the
    forces are formed (a change far out would not reach the body within a
    short run)."""
    nt = cs.mesh.nt
    return [nt + (j * nt) // k + 3 for j in range(k)]


def _ale_run(cs, prec, ws, rel, tol, maxit, warmup, steps, ulp_node=None) -> dict:
    import numpy as np
    from cfd_amd.solver import Session
    S = Session(cs, precision=prec, pcg_rel=rel, pcg_tol=tol, pcg_maxit=maxit, warm_start=ws)
    if ulp_node is not None:  # one-ulp change of one node's density
        U = S.get("U")
        U[ulp_node, 0] = np.nextafter(U[ulp_node, 0], 2.0)
        S.set("U", U)
        S.set("U1", U)
    S.step(warmup)
    S.sync()
    sc0 = S.scalars()
    t0 = time.perf_counter()
    S.step(steps)
    S.sync()
    sec = time.perf_counter() - t0
    sc = S.scalars()
    r = {"ms_per_step": sec / steps * 1e3,
         "pcg_iterations_per_step": (sc["PCG_ITSUM"] - sc0["PCG_ITSUM"]) / steps,
         "unconverged_solves": int(sc["PCG_UNCONV"]),
         **force_coefficients(sc, cs)}
    if ws:
        r["predicted_guess_used"] = f"{sc['WS_LEARNED']:.0f}/{sc['WS_SOLVES']:.0f}"
    if ws == 1:
        r["mlp_loss"] = sc["WS_LOSS"]
    S.close()
    return r


def _rel(a: float, b: float) -> float:
    return abs(a - b) / abs(b) if b != 0.0 else abs(a - b)


def ale_compare(steps: int = 40, warmup: int = 5) -> dict:
    """ALE step time, learned vs plain, at equal PCG tolerance: the pitching
    ellipse O-grid 512 x 1024 (524,288 nodes, 1,046,528 triangles, M=0.5,
    Re=5000), alpha(t) = 5 deg sin(0.2 t) about x = -0.25.  plain = fused
    FP64 path, each mesh solve started from the previous step's solution (the
    reference); learned = fused FP32-element path, mesh solves started from
    the learned prediction (the rates' linear extrapolation plus the MLP's
    correction, matrix cores, with fallback); extrap = the same without the
    network (the extrapolation alone, the network's ablation).  Under the
    relative PCG rules 1e-10 (the configuration's), 1e-8 and 1e-6 (r.z <=
    tol^2 r0.z of the guess that is zero off the fixed nodes), with up to
    20,000 iterations per solve so that every solve converges (the count of
    solves that stopped at the cap is reported), and under the reference's
    absolute rule (r.z < 1e-10, 1000 iterations).

    Drag and lift: coefficients of the total force on the body (pressure and
    viscous, FORCE_VISC) after the same steps; each difference is relative to
    the plain run's own value.  The flow is chaotic at round-off level, so
    under the configuration's rule the plain run is repeated from five states
    with a one-ulp change of the density of one node next to the body: the
    spread of those differences (min / median / max, "chaos floor") is what
    any differently-rounded run shows."""
    import numpy as np
    from cfd_amd.mesh import ellipse_case
    cs = ellipse_case(motion=1, motion_amp=PITCH_AMP, motion_omega=PITCH_OMEGA)
    res = {}
    for rule, rel, tol, maxit in (("relative_1e-10", 1, 1e-10, 20000),
                                  ("relative_1e-8", 1, 1e-8, 20000),
                                  ("relative_1e-6", 1, 1e-6, 20000),
                                  ("absolute_1e-10", 0, 0.0, 0)):
        runs = {name: _ale_run(cs, prec, ws, rel, tol, maxit, warmup, steps)
                for name, prec, ws in (("plain_fp64", 0, 0), ("extrap_fp32", 1, 2),
                                       ("learned_fp32", 1, 1))}
        p, q, e = runs["plain_fp64"], runs["learned_fp32"], runs["extrap_fp32"]
        runs["speedup"] = p["ms_per_step"] / q["ms_per_step"]
        runs["speedup_extrap"] = p["ms_per_step"] / e["ms_per_step"]
        runs["network_iteration_ratio"] = (q["pcg_iterations_per_step"]
                                           / max(e["pcg_iterations_per_step"], 1e-9))
        runs["network_time_ratio"] = q["ms_per_step"] / e["ms_per_step"]
        runs["cd_diff"] = _rel(q["cd"], p["cd"])
        runs["cl_diff"] = _rel(q["cl"], p["cl"])
        if rule == "relative_1e-10":
            fl = [_ale_run(cs, 0, 0, rel, tol, maxit, warmup, steps, ulp_node=k)
                  for k in near_body_nodes(cs)]
            for key in ("cd", "cl"):
                d = np.array([_rel(f[key], p[key]) for f in fl])
                runs[f"chaos_floor_{key}"] = {"min": float(d.min()), "median": float(np.median(d)),
                                              "max": float(d.max()), "samples": len(d)}
        res[rule] = runs
        progress(f"ale {rule}: learned {q['ms_per_step']:.2f} ms/step, "
                 f"extrap {e['ms_per_step']:.2f}, plain {p['ms_per_step']:.2f}, unconverged "
                 f"{p['unconverged_solves']}/{e['unconverged_solves']}/{q['unconverged_solves']}")
    res["workload"] = (f"ellipse O-grid 512x1024 ({cs.npoin} nodes, {cs.nelem} triangles), M=0.5, "
                       f"Re=5000, alpha(t) = 5 deg sin(0.2 t) about x = -0.25, {warmup} untimed + "
                       f"{steps} timed steps")
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


LARGE_RUNS = ("plain_fp64", "extrap_fp32", "learned_fp32")
# opt-in: the plain run with the density of one node next to the body changed
# by one ulp (the chaos floor of the force histories)
LARGE_FLOOR = "plain_fp64_ulp"


def large_ale(steps: int = 10, warmup: int = 3, tol: float = 1e-10, maxit: int = 100000,
              runs_wanted=LARGE_RUNS) -> dict:
    """The large ALE run: pitching ellipse O-grid 4096 x 4096 (16,777,216 nodes,
    33,546,240 triangles), M=0.7, Re=10000, alpha(t) = 5 deg sin(0.2 t), the
    GCL volume update on; plain (fused FP64, previous-step guess) against
    learned (fused FP32 element, learned warm start) and extrap (the same
    without the network) at the relative PCG rule `tol` with `maxit`
    iterations allowed (the reference's 1000 end every solve unconverged at
    this size), a few steps each (the mesh solves dominate; at this size one
    plain step takes about a second, and the first solve from the start-up
    guess tens of seconds).  The three warm-up steps fill the three-solution
    history the network reads.  Each run records its drag and lift
    coefficients after every timed step; the histories of the other runs are
    compared with the plain run's step by step (largest relative difference;
    for lift also the largest absolute one), and so is the one-ulp floor run
    when asked for (LARGE_FLOOR)."""
    from cfd_amd.mesh import ellipse_case
    from cfd_amd.solver import Session
    cs = ellipse_case(4096, 4096, mach=0.7, re=10000.0, motion=1, motion_amp=PITCH_AMP,
                      motion_omega=PITCH_OMEGA)
    import numpy as np
    runs, hist = {}, {}
    for name, prec, ws in (("plain_fp64", 0, 0), ("extrap_fp32", 1, 2), ("learned_fp32", 1, 1),
                           (LARGE_FLOOR, 0, 0)):
        if name not in runs_wanted:
            continue
        t0 = time.perf_counter()
        S = Session(cs, precision=prec, pcg_rel=1, pcg_tol=tol, pcg_maxit=maxit, warm_start=ws,
                    gcl=1)
        if name == LARGE_FLOOR:
            U = S.get("U")
            k = near_body_nodes(cs, 1)[0]
            U[k, 0] = np.nextafter(U[k, 0], 2.0)
            S.set("U", U)
            S.set("U1", U)
            del U
        t_create = time.perf_counter() - t0
        progress(f"large_ale {name}: created in {t_create:.1f} s")
        for _ in range(warmup):
            S.step(1)
            S.sync()
            progress(f"large_ale {name}: warm-up step, PCG iterations so far "
                     f"{S.scalars()['PCG_ITSUM']:.0f}")
        it0 = S.scalars()["PCG_ITSUM"]
        sec = 0.0
        hcd, hcl = [], []
        for j in range(steps):  # one step at a time: progress lines between
            t0 = time.perf_counter()
            S.step(1)
            S.sync()
            sec += time.perf_counter() - t0
            fc = force_coefficients(S.scalars(), cs)
            hcd.append(fc["cd"])
            hcl.append(fc["cl"])
            if j % 10 == 9 or j == steps - 1:
                progress(f"large_ale {name}: {j + 1} steps, {sec:.1f} s so far, cd {fc['cd']:.6f}")
        hist[name] = (np.array(hcd), np.array(hcl))
        sc = S.scalars()
        runs[name] = {"ms_per_step": sec / steps * 1e3, "create_s": t_create,
                      "pcg_iterations_per_step": (sc["PCG_ITSUM"] - it0) / steps,
                      "unconverged_solves": int(sc["PCG_UNCONV"]), **force_coefficients(sc, cs)}
        if ws:
            runs[name]["predicted_guess_used"] = f"{sc['WS_LEARNED']:.0f}/{sc['WS_SOLVES']:.0f}"
        S.close()
    if all(k in runs for k in LARGE_RUNS):
        p = runs["plain_fp64"]
        runs["speedup"] = p["ms_per_step"] / runs["learned_fp32"]["ms_per_step"]
        runs["speedup_extrap"] = p["ms_per_step"] / runs["extrap_fp32"]["ms_per_step"]
        for k in ("extrap_fp32", "learned_fp32"):
            runs[f"{k}_cd_diff"] = _rel(runs[k]["cd"], p["cd"])
            runs[f"{k}_cl_diff"] = _rel(runs[k]["cl"], p["cl"])
    if "plain_fp64" in hist:
        pcd, pcl = hist["plain_fp64"]
        for k, (cd, cl) in hist.items():
            if k == "plain_fp64":
                continue
            runs[f"{k}_history"] = {
                "steps": int(cd.size),
                "cd_max_rel_diff": float(np.max(np.abs(cd - pcd) / np.abs(pcd))),
                "cl_max_rel_diff": float(np.max(np.abs(cl - pcl) / np.abs(pcl))),
                "cl_max_abs_diff": float(np.max(np.abs(cl - pcl)))}
        runs["plain_fp64_history"] = {"cd_first_last": [float(pcd[0]), float(pcd[-1])],
                                      "cl_first_last": [float(pcl[0]), float(pcl[-1])]}
    runs["workload"] = (f"ellipse O-grid 4096x4096 ({cs.npoin} nodes, {cs.nelem} triangles), M=0.7, "
                        f"Re=10000, alpha(t) = 5 deg sin(0.2 t), GCL on, relative PCG rule {tol:g} "
                        f"(at most {maxit} iterations), {warmup} untimed + {steps} timed steps")
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


DEVELOP_FIELDS = ("U", "T", "VEL_X", "VEL_Y", "RHO", "E", "P", "RMACH")


def cylinder_bench_case(nr: int = 1024, nt: int = 2048):
    """The bench's cylinder O-grid (1024 x 2048 by default): M 0.8, Re 1000,
    fixed mesh, at an acoustic CFL of 0.5 (FSAFE = 0.5 M / (1 + M); deltat
    bounds only the advective and viscous rates, and at FSAFE 0.5 the M 0.8
    run develops negative temperatures within a few hundred steps)."""
    from cfd_amd.mesh import acoustic_fsafe, cylinder_case
    return cylinder_case(nr, nt, mach=0.8, re=1000.0, fsafe=acoustic_fsafe(0.8))


def _session(cs, prec, renumber=True):
    from cfd_amd.solver import Session
    return Session(cs, precision=prec, renumber=1 if renumber and prec in (0, 1) else 0)


def develop(cs, nsteps: int, prec: int = 0, renumber: bool = True, ulp_node=None):
    """nsteps from the free stream (untimed): the state (fields the next step
    reads, the step scalars) and the force coefficients at the end."""
    import numpy as np
    S = _session(cs, prec, renumber)
    if ulp_node is not None:
        U = S.get("U")
        U[ulp_node, 0] = np.nextafter(U[ulp_node, 0], 2.0)
        S.set("U", U)
        S.set("U1", U)
    S.step(nsteps)
    st, sc, finite = session_state(S)
    S.close()
    return st, sc, finite


def session_state(S):
    """The session's state (the fields the next step reads and the step
    scalars), its scalars, and whether the state is finite."""
    import numpy as np
    S.sync()
    sc = S.scalars()
    st = {k: S.get(k) for k in DEVELOP_FIELDS}
    st.update({k: sc[k] for k in ("DTMIN", "TIME", "DTMIN1", "BANDERA", "ITER")})
    return st, sc, bool(np.isfinite(st["U"]).all())


def accuracy(cs, dev: dict, dev_sc: dict, nsteps: int) -> dict:
    """The FP32-element path against the FP64 path after the same nsteps from
    the free stream: L2 density error and the drag and lift coefficients
    (total force on the body), with the differences a one-ulp change of one
    node's density makes to the FP64 run (the round-off chaos floor)."""
    import numpy as np
    out = {"steps": nsteps}
    st32, sc32, fin32 = develop(cs, nsteps, prec=1)
    stu, scu, finu = develop(cs, nsteps, prec=0, ulp_node=near_body_nodes(cs, 1)[0])
    f64, f32, fu = (force_coefficients(x, cs) for x in (dev_sc, sc32, scu))
    rho = dev["U"][:, 0]
    out["fp64"] = f64
    out["fp32"] = f32
    out["fp32_l2_rho"] = float(np.linalg.norm(st32["U"][:, 0] - rho) / np.linalg.norm(rho))
    out["fp32_cd_diff"] = _rel(f32["cd"], f64["cd"])
    out["fp32_cl_diff"] = _rel(f32["cl"], f64["cl"])
    out["fp32_cl_abs_diff"] = abs(f32["cl"] - f64["cl"])   # the symmetric flow's Cl is near 0
    out["floor_l2_rho"] = float(np.linalg.norm(stu["U"][:, 0] - rho) / np.linalg.norm(rho))
    out["floor_cd_diff"] = _rel(fu["cd"], f64["cd"])
    out["floor_cl_diff"] = _rel(fu["cl"], f64["cl"])
    out["floor_cl_abs_diff"] = abs(fu["cl"] - f64["cl"])
    out["states_finite"] = fin32 and finu
    out["note"] = ("FP64 here is the fused FP64 GPU path, which lands within 1e-12 of the FP64 "
                   "oracle on every teacher-forced step; the floor is the FP64 run against "
                   "itself with the density of one node next to the body changed by one ulp")
    return out


def run_path(cs, prec: int, steps: int, warmup: int, start: dict | None,
             breakdown: bool = False, renumber: bool = True, session=None) -> dict:
    """Loads the developed state `start` (or continues `session`, the open
    session that is developing it, and leaves it open), runs `warmup` untimed
    steps, then
    times `steps` whole steps; HIP events on the session's stream time the
    element kernel launches inside the timed region (on the fused paths one
    event pair around each step's four back-to-back stage launches, on the
    reference-order path one pair per launch; the events are created before
    the timed region).  The fused paths run
    with the session's patch-order numbering (renumber = 1: the same results
    bit for bit on a fixed mesh)."""
    if session is None:
        S = _session(cs, prec, renumber)
        S.set_state(start, fields=DEVELOP_FIELDS)
    else:
        S = session
    # timing on before the warm-up, so that the session's event pool is
    # built before it and the timed steps follow the warm-up without an idle
    # gap; the second call only resets the accumulators
    S.timing(True, slots=["ELEMENT"])
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
         "state": {"dtmin": sc["DTMIN"], "time": sc["TIME"], "iter": sc["ITER"],
                   **force_coefficients(sc, cs)}}
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
    if session is None:
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
    ap.add_argument("--no-large-ale", action="store_true",
                    help="skip the 4096 x 4096 pitching ellipse (about two minutes)")
    ap.add_argument("--large-ale-runs", default=",".join(LARGE_RUNS),
                    help="comma-separated runs of the large ALE section: "
                         f"{', '.join(LARGE_RUNS)} and, opt-in, {LARGE_FLOOR} (the plain run "
                         "with one node's density changed by one ulp: the force histories' "
                         "chaos floor)")
    ap.add_argument("--large-ale-steps", type=int, default=10)
    ap.add_argument("--large-ale-warmup", type=int, default=3)
    ap.add_argument("--only", choices=["large_ale", "ale"], default=None,
                    help="run only this section and print its JSON (profiling)")
    ap.add_argument("--develop", type=int, default=500,
                    help="untimed FP64 steps from the free stream before timing (the state timed)")
    ap.add_argument("--no-renumber", action="store_true",
                    help="fused paths in the caller's numbering (no patch-order numbering)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU: the path is a single-GPU kernel set "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if args.only == "large_ale":
        print(json.dumps({"large_ale": large_ale(args.large_ale_steps, args.large_ale_warmup,
                                                 runs_wanted=args.large_ale_runs.split(","))}))
        return
    if args.only == "ale":
        print(json.dumps({"ale": ale_compare(args.ale_steps)}))
        return
    cs = cylinder_bench_case(args.nr, args.nt)
    prec = PRECISION[args.precision]
    rn = not args.no_renumber
    if prec == 0:
        # the FP64 path's warm-up and timed steps follow its development in
        # the same session, with no idle gap between them (after an idle gap
        # the GPU runs the next few dozen steps slower); the state that the
        # accuracy section and the other paths start from is taken after the
        # timed and breakdown steps
        dev_s = _session(cs, 0, rn)
        dev_s.step(args.develop)
        r = run_path(cs, 0, args.steps, args.warmup, None, breakdown=True, renumber=rn,
                     session=dev_s)
        dev, dev_sc, dev_finite = session_state(dev_s)
        dev_s.close()
        dev_steps = args.develop + args.warmup + args.steps + min(args.steps, 10)
    else:
        dev, dev_sc, dev_finite = develop(cs, args.develop, prec=0, renumber=rn)
        dev_steps = args.develop
    if not dev_finite:
        raise RuntimeError("the developed state is not finite")
    progress(f"developed state: {dev_steps} FP64 steps, t = {dev['TIME']:.4g}")
    if prec != 0:
        r = run_path(cs, prec, args.steps, args.warmup, dev, breakdown=True, renumber=rn)
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
                        f"{cs.nelem} triangles), M=0.8, Re=1000, fixed mesh, acoustic CFL 0.5 "
                        f"(FSAFE {cs.fsafe:.4f}), {PATH[prec]} path, timed from the state after "
                        f"{args.develop} steps"
                        + (", patch-order numbering inside the session" if rn and prec in (0, 1)
                           else ""),
            "nelem": cs.nelem, "npoin": cs.npoin, "calcrhs_per_step": 4,
            "developed_steps": args.develop,
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
    out["accuracy"] = accuracy(cs, dev, dev_sc, dev_steps)
    progress(f"accuracy: FP32 L2(rho) {out['accuracy']['fp32_l2_rho']:.3e} after {dev_steps} steps")
    if not args.no_variants:
        # the other two element paths, shorter runs of the same workload
        out["variants"] = {}
        for name, p in PRECISION.items():
            if p != prec:
                v = run_path(cs, p, 20, 20, dev, renumber=rn)
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
    if not args.no_large_ale:
        out["large_ale"] = large_ale(args.large_ale_steps, args.large_ale_warmup,
                                     runs_wanted=args.large_ale_runs.split(","))
        progress("large ALE done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cs)
        progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
