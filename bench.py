"""Benchmark of the PECS per-time-step IMEX path on one MI355X.

Workload (default): refinement 8 (256 x 256 quads per region), LDG degree 2,
RT1 x DGQ1 Poisson, dt = 5e-4, constant donor doping 1.0, Beer-Lambert
generation G = exp(-x), SRH recombination (tau = 1), interface transfer rates
1e-2, bias 0 -- the "plain FP64 on one GPU at scale" configuration.  A step is
one pass of the IMEX loop body (SolarCell.cpp:2057-2075): both carrier
right-hand sides, four carrier solves, the Poisson right-hand side and solve.

value = DOF-updates/s = (4 N_c + N_P) * steps / seconds, state resident in HBM.
The roofline object is for the dominant kernel family (the supernodal
triangular solves of the semiconductor Schur factor, two right-hand sides),
timed live with hipEvents on the context stream over the timed steps.
cpu_baseline = the compiled CPU restatement of the step (oracle/cpu_step.cpp:
own C++ FP64, OpenMP on all host cores; not deal.II, which cannot be built
here) on the measured workload, a bounded number of steps, reported beside it.

time_to_steady_state: wall seconds from the equilibrium initial state until
the relative steady residual ||x^{n+1}-x^n|| / (||x^n|| dt) over the five
state vectors falls to 1e-8 (the reference has no criterion of its own),
on the measured workload and, beside the CPU restatement, on the small
workload (refinement 4, k = 1, dt = 1e-3).

batched: an I-V sweep of 16 applied biases on the measured mesh stepped
together (one read of each factor per sweep for all of them, FP64 MFMA):
aggregate step throughput and time to steady state per bias point.

sweep (--sweep N): an uneven-mesh sweep (refinements 5-8, conic geometries,
drawn biases, generation and doping), one context per instance, stepped in
waves by graph launches; time to steady state per bias point against the
plain path (each instance alone, launch by launch).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6  # MI355X dense FP64 (vector and matrix)


CPU_STEP = os.path.join(ROOT, "oracle", "_build", "pecs_cpu_step")


def _cpu_step(prm, steps, tol=0.0, timeout=1800):
    """Run the compiled CPU restatement of the step (oracle/cpu_step.cpp) on
    all host cores (OMP_NUM_THREADS if set); returns its JSON report."""
    import subprocess
    if not os.path.exists(CPU_STEP):
        raise RuntimeError(f"{CPU_STEP} not built (run __graft_entry__.build())")
    gen = {"ref": 0, "config": 1, "off": 2}[prm["gen_mode"]]
    args = [f"{k}={float(prm[k])!r}" for k in ("k", "r", "dt", "penalty", "mu_n", "mu_p", "mu_r", "mu_o", "eps_s",
                                                 "eps_e", "lam2", "k_et", "k_ht", "v_rec_n", "v_rec_p", "ne", "pe",
                                                 "rinf", "oinf", "phi_bi", "phi_app", "G0", "alpha", "ni", "tau_n",
                                                 "tau_p", "r1", "r2", "H", "L")]
    args += [f"gen_mode={gen}", f"srh={int(bool(prm['srh']))}", f"insulated={int(bool(prm['insulated']))}",
             f"schottky={int(bool(prm['schottky']))}", f"steps={steps}", f"tol={tol!r}"]
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", str(os.cpu_count() or 1))
    r = subprocess.run([CPU_STEP] + args, capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"pecs_cpu_step failed ({r.returncode}): {r.stderr[-500:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(args, gpu_steps_to_steady=None):
    """The compiled CPU restatement (own C++ FP64 step, not deal.II: OpenMP
    right-hand sides, Schur / multiplier solves with host supernodal Cholesky
    factors and level-parallel sweeps) on the measured workload itself, a
    bounded number of steps, on all host cores."""
    from pecs_amd import workload_params
    p = workload_params(args.refine, args.degree, args.dt)
    rep = _cpu_step(p, args.cpu_steps)
    out = {"value": rep["dof_updates_per_step"] / rep["step_s"], "unit": "DOF-updates/s", "cores": rep["threads"],
           "kind": "port",
           "sample": f"own C++ FP64 restatement of the step (oracle/cpu_step.cpp), not deal.II: the measured "
                     f"workload (refinement {args.refine}, k={args.degree}, dt={args.dt}), {rep['steps']} steps, "
                     f"{rep['threads']} OpenMP threads; host factorisation (setup {rep['setup_s']:.1f} s) excluded",
           "ms_per_step": rep["step_s"] * 1e3, "setup_s": rep["setup_s"]}
    if gpu_steps_to_steady:
        # the steady state needs the same number of steps on either path (same algorithm)
        out["time_to_steady_s_estimate"] = gpu_steps_to_steady * rep["step_s"]
        out["time_to_steady_basis"] = f"{gpu_steps_to_steady} steps (the GPU run's count) x the CPU step time"
    return out


def profiled_traffic(workload):
    """HBM bytes per launch of the roofline kernel from the committed PMC
    summary (FETCH_SIZE and WRITE_SIZE runs scaled by a plain-copy
    calibration), when it was measured on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "carrier_solves_r8k2_hbm_traffic.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    vals = dict(line.split(": ", 1) for line in text.splitlines()
                if line.startswith(("traffic_bytes_per_launch: ", "workload: ")))
    if vals.get("workload", "").strip() != workload:
        return None
    return float(vals["traffic_bytes_per_launch"])


def profiled_values(name, workload, keys):
    """key: value lines of a committed profile summary measured on this
    workload (else None)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        text = open(path).read()
    except OSError:
        return None
    vals = dict(line.split(": ", 1) for line in text.splitlines()
                if line.startswith(tuple(k + ": " for k in list(keys) + ["workload"])))
    if vals.get("workload", "").strip() != workload:
        return None
    return {k: float(vals[k]) for k in keys if k in vals}


def cpu_time_to_steady(tol, max_steps=20000):
    """The compiled CPU restatement on the small workload (refinement 4, k = 1,
    dt = 1e-3), stepped until the relative steady residual is at most tol."""
    from pecs_amd import workload_params
    rep = _cpu_step(workload_params(4, 1, 1e-3), max_steps, tol)
    return {"s": rep["run_s"], "steps": rep["steps"], "residual": rep["residual"], "cores": rep["threads"],
            "kind": "port", "reached": 0 <= rep["residual"] <= tol}


def gpu_time_to_steady(g, tol, check_every, max_steps):
    """Steps from the equilibrium initial state until the relative steady
    residual ||x^{n+1}-x^n|| / (||x^n|| dt) <= tol, checked every check_every
    steps (the residual is that of the last step of each block)."""
    import torch
    g.initial_state()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n, res = 0, float("inf")
    while n < max_steps and res > tol:
        k = min(check_every, max_steps - n)
        g.step(k)
        n += k
        res = g.residual()
    return {"s": time.perf_counter() - t0, "steps": n, "residual": res, "reached": res <= tol}


def assembly_rate(args, g, phase):
    """Carrier right-hand-side assembly of one step (k_rhs_carrier, both
    regions): algorithmic bytes = old densities of both carriers, the cell's
    RT field DoFs and 4 vertices read, both carriers' full right-hand sides
    written, per cell; over the sum of the two regions' phase times.  The
    kernel evaluates the geometry, the field, exp (generation) and the drift
    terms in FP64 at every quadrature point, so it sits below the HBM bound."""
    nb = (args.degree + 1) ** 2
    nrt = 2 * args.degree * (args.degree + 1)
    ncell = (g.n_semi + g.n_elec) // (3 * nb)
    nbytes = 8.0 * ncell * (2 * nb + nrt + 8 + 2 * 3 * nb)
    ms = phase[0] + phase[1]
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {"kernel": "k_rhs_carrier (both regions)", "algorithmic_bytes": nbytes, "avg_ms": ms,
           "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    wl = (f"PECS half cell, refinement {args.refine} ({1 << args.refine}x{1 << args.refine} quads/region), "
          f"LDG k={args.degree}, RT{args.degree - 1}xDGQ{args.degree - 1} Poisson, dt={args.dt}")
    pv = profiled_values("rhs_carrier_r8k2_hbm_traffic.md", wl, ["rhs_traffic_bytes_per_step"])
    if pv and ms > 0:
        t = pv["rhs_traffic_bytes_per_step"]
        out["counter"] = {"traffic_bytes": t, "achieved": t / (ms * 1e-3) / 1e9, "unit": "GB/s",
                          "frac": t / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "source": "profiles/rhs_carrier_r8k2_hbm_traffic.md (rocprofv3 --pmc FETCH_SIZE, "
                                    "WRITE_SIZE; copy-calibrated) over this run's kernel time"}
    return out


def mixed_leg(args, workload):
    """Mixed precision on a batch of bias points (pecs_set_mixed_precision:
    FP32 contractions with the cell-type tables in the Schur and Poisson cell
    kernels; FP64 state, assembly and triangular solves): the same steps from
    the same initial state with the FP64 path and the mixed path, timed as
    replayed graphs; per instance the relative L2 difference of each state
    vector against the FP64 path (which follows the oracle to 1e-10)."""
    import numpy as np
    import torch
    import pecs_amd
    B = max(args.batch, 1)
    prms = []
    for v in np.linspace(-0.5, 0.5, B):
        p = pecs_amd.workload_params(args.refine, args.degree, args.dt)
        p.update(phi_app=float(v))
        prms.append(p)
    g = pecs_amd.SolarCellGPU(prms if B > 1 else prms[0])
    n = args.mixed_steps

    def run():
        g.initial_state()
        g.step(3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(n)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / n * 1e3
        return ms, [g.get_state(i) for i in range(B)]

    ms64, s64 = run()
    g.set_mixed_precision(True)
    ms32, s32 = run()
    g.close()
    names = {"n": "electrons", "p": "holes", "r": "reductants", "o": "oxidants", "P": "poisson"}
    per_inst = []
    for a, b in zip(s32, s64):
        per_inst.append({names[c]: float(np.linalg.norm(a[c] - b[c]) / max(np.linalg.norm(b[c]), 1e-300))
                         for c in "nproP"})
    worst = {names[c]: max(e[names[c]] for e in per_inst) for c in "nproP"}
    print(f"[bench] mixed precision: {ms32:.3f} ms/step against {ms64:.3f} (FP64); worst errors {worst}",
          file=sys.stderr, flush=True)
    return {"batch": B, "workload": workload + f", {B} applied biases in [-0.5, 0.5]", "steps": n + 3,
            "contractions": "FP32 with FP32 accumulation (Schur transform and recovery, Poisson condensation and "
                            "recovery); state, assembly and triangular solves FP64",
            "ms_per_step_fp64": ms64, "ms_per_step_mixed": ms32, "speedup": ms64 / ms32 if ms32 > 0 else None,
            "rel_l2_error_vs_fp64_max_over_instances": worst,
            "rel_l2_error_vs_fp64_per_instance_max_over_fields": [max(e.values()) for e in per_inst]}


def batched_leg(args, workload):
    """I-V sweep on one mesh: B bias points (applied bias evenly over
    [-0.5, 0.5] scaled volts) stepped together, every factor read once per
    sweep for all of them (pecs_create_batch).  Reports the aggregate step
    throughput and the time to steady state per bias point."""
    import numpy as np
    import torch
    import pecs_amd
    B = args.batch
    prms = []
    for v in np.linspace(-0.5, 0.5, B):
        p = pecs_amd.workload_params(args.refine, args.degree, args.dt)
        p.update(phi_app=float(v))
        prms.append(p)
    t0 = time.perf_counter()
    g = pecs_amd.SolarCellGPU(prms)
    setup_s = time.perf_counter() - t0
    print(f"[bench] batch of {B}: setup {setup_s:.1f}s", file=sys.stderr, flush=True)
    g.initial_state()
    g.step(3)
    L = pecs_amd.lib()
    import ctypes
    L.pecs_enable_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.pecs_phase_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
    assert L.pecs_enable_timing(g.ctx, 1) == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step(args.batch_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ph = (ctypes.c_double * 8)()
    nst = ctypes.c_int()
    L.pecs_phase_times(g.ctx, ph, 8, ctypes.byref(nst))
    L.pecs_enable_timing(g.ctx, 0)
    names = ["rhs_semi", "rhs_elec", "schur_transforms", "carrier_cholesky", "recoveries", "poisson_cond",
             "chol_poisson", "poisson_recover"]
    phase = {n: round(ph[i] / max(nst.value, 1), 4) for i, n in enumerate(names)}
    U = 2 * g.n_semi + 2 * g.n_elec + g.n_poisson
    ms_timed = el / args.batch_steps * 1e3
    # the production path: the same number of steps replayed as graphs
    if not args.no_graph:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(args.batch_steps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    ms_step = el / args.batch_steps * 1e3
    # the batched carrier solves on FP64 MFMA: 2 flops per stored factor entry
    # (both copies) per right-hand side, 2B right-hand sides per factor
    fb = ctypes.c_double()
    L.pecs_factor_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    flops = 0.0
    for which in (0, 1):
        L.pecs_factor_bytes(g.ctx, which, ctypes.byref(fb))
        flops += 2.0 * (fb.value / 8.0) * 2 * B
    cms = phase["carrier_cholesky"]
    tf = flops / (cms * 1e-3) / 1e12 if cms > 0 else 0.0
    out = {"batch": B, "workload": workload + f", {B} applied biases in [-0.5, 0.5]", "setup_s": round(setup_s, 2),
           "ms_per_step_all_instances": ms_step, "ms_per_step_per_instance": ms_step / B,
           "ms_per_step_timed_phases": ms_timed,
           "dof_updates_per_s": U * B * args.batch_steps / el, "phase_ms_per_step": phase,
           "carrier_solves_mfma": {"kernel": "k_sweep_mfma_lds / k_sweep_mfma (FP64 MFMA 16x16x4) + packing and gathers, both regions, one step",
                                   "flops": flops, "avg_ms": cms, "achieved": tf, "peak": FP64_PEAK_TFS,
                                   "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFS},
           "sweep_tiles_semi": [list(t) for t in g.sweep_config(0)]}
    prof = f"carrier_solves_r8k2_batch{B}_counters.md"
    pv = profiled_values(prof, out["workload"],
                         ["mfma_flops_per_solve", "mfma_busy_cycles_per_solve", "batch_traffic_bytes_per_solve"])
    if pv and cms > 0:
        cf = pv["mfma_flops_per_solve"] / (cms * 1e-3) / 1e12
        out["carrier_solves_mfma"]["counter"] = {
            "flops": pv["mfma_flops_per_solve"], "achieved": cf, "unit": "TFLOP/s", "frac": cf / FP64_PEAK_TFS,
            "mfma_busy_frac": pv["mfma_busy_cycles_per_solve"] / (1024 * cms * 1e-3 * 2.4e9),
            "traffic_bytes": pv["batch_traffic_bytes_per_solve"],
            "hbm_frac": pv["batch_traffic_bytes_per_solve"] / (cms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "source": f"profiles/{prof} (rocprofv3 --pmc "
                      "SQ_INSTS_VALU_MFMA_MOPS_F64, SQ_VALU_MFMA_BUSY_CYCLES, FETCH_SIZE, WRITE_SIZE) over this "
                      "run's kernel time; busy cycles over 1024 SIMDs at 2.4 GHz"}
    if not args.no_steady:
        g.initial_state()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n, done = 0, np.full(B, -1)
        while n < args.steady_max_steps:
            k = min(args.steady_check_every, args.steady_max_steps - n)
            g.step(k)
            n += k
            r = g.residuals()
            done[(done < 0) & (r <= args.steady_tol)] = n
            if np.all(done >= 0):
                break
        wall = time.perf_counter() - t0
        out["time_to_steady_state"] = {"tol": args.steady_tol, "s_all": wall, "s_per_bias_point": wall / B,
                                       "steps_all": n, "steps_per_instance": done.tolist(),
                                       "reached": bool(np.all(done >= 0))}
    g.close()
    print(f"[bench] batched: {out}", file=sys.stderr, flush=True)
    return out


def sweep_leg(args):
    """Uneven-mesh I-V sweep: args.sweep instances with refinement drawn from
    {args.sweep_rmin..args.sweep_rmax}, conic half-angle U[0, 15] degrees
    (r2 = r1 + H tan(angle), L = 1.5), applied bias U[-0.5, 0.5], generation
    amplitude logU[0.1, 10], donor doping logU[0.5, 5] (seeded).  Each instance
    is its own context (its own mesh, factors and tuned tile shapes); the
    instances that fit in HBM at once form a wave, stepped together by one
    graph launch per stream and step (pecs_sweep_*) until each reaches the
    steady tolerance (it then leaves the sweep).  The plain path is each
    instance stepped alone (graph-replayed steps): its time per step is
    measured per instance and multiplied by the instance's steps to steady
    state."""
    import numpy as np
    import torch
    import pecs_amd
    rng = np.random.default_rng(args.sweep_seed)
    draws = []
    for i in range(args.sweep):
        r = int(rng.integers(args.sweep_rmin, args.sweep_rmax + 1))
        ang = float(rng.uniform(0.0, 15.0))
        p = pecs_amd.workload_params(r, args.degree, args.dt)
        p.update(r1=0.5, r2=0.5 + float(np.tan(np.radians(ang))), L=1.5, phi_bi=0.4,
                 phi_app=float(rng.uniform(-0.5, 0.5)), G0=float(10 ** rng.uniform(-1, 1)),
                 ne=float(10 ** rng.uniform(np.log10(0.5), np.log10(5.0))))
        draws.append((r, ang, p))
    order = sorted(range(len(draws)), key=lambda i: -draws[i][0])
    reserve = args.sweep_reserve_gib * (1 << 30)
    total_setup = total_wall = total_plain = 0.0
    steps = [None] * len(draws)
    plain_ms = [None] * len(draws)
    waves = []
    pos = 0
    while pos < len(order):
        members, idx = [], []
        t0 = time.perf_counter()
        while pos < len(order):
            free, _ = torch.cuda.mem_get_info()
            if members and free < reserve:
                break
            i = order[pos]
            members.append(pecs_amd.SolarCellGPU(draws[i][2]))
            idx.append(i)
            pos += 1
            print(f"[bench] sweep: instance {i} (refinement {draws[i][0]}) set up, {free / (1 << 30):.0f} GiB "
                  f"were free", file=sys.stderr, flush=True)
        setup = time.perf_counter() - t0
        total_setup += setup
        # plain path: each member alone, graph-replayed steps
        for g, i in zip(members, idx):
            g.initial_state()
            g.step(2)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            g.step(args.sweep_plain_steps)
            torch.cuda.synchronize()
            plain_ms[i] = (time.perf_counter() - t1) / args.sweep_plain_steps * 1e3
            g.initial_state()
        sw = pecs_amd.SolarCellSweep(members, args.sweep_streams)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        active = np.ones(len(members), bool)
        n = 0
        while active.any() and n < args.steady_max_steps:
            k = min(args.steady_check_every, args.steady_max_steps - n)
            sw.step(k)
            n += k
            r = np.array([g.residual() for g in members])
            newly = active & (r <= args.steady_tol)
            for j in np.nonzero(newly)[0]:
                steps[idx[j]] = n
            if newly.any():
                active &= ~newly
                sw.set_active(active)
            if n % 500 == 0:
                print(f"[bench] sweep: step {n}, {int(active.sum())} of {len(members)} still stepping",
                      file=sys.stderr, flush=True)
        wall = time.perf_counter() - t1
        total_wall += wall
        for j in np.nonzero(active)[0]:
            steps[idx[j]] = -n
        sw.close()
        for g in members:
            g.close()
        waves.append({"instances": len(members), "refinements": [draws[i][0] for i in idx],
                      "setup_s": round(setup, 2), "sweep_s": round(wall, 3), "steps": n})
        print(f"[bench] sweep wave {len(waves)}: {waves[-1]}", file=sys.stderr, flush=True)
    for i in range(len(draws)):
        total_plain += abs(steps[i]) * plain_ms[i] * 1e-3
    N = len(draws)
    return {"instances": N, "refinements": [d[0] for d in draws], "half_angles_deg": [round(d[1], 2) for d in draws],
            "streams": args.sweep_streams, "waves": waves, "setup_s": round(total_setup, 2),
            "tol": args.steady_tol, "steps_to_steady": steps,
            "reached": all(s is not None and s > 0 for s in steps),
            "s_per_bias_point": total_wall / N, "plain_s_per_bias_point": total_plain / N,
            "speedup_vs_plain": total_plain / total_wall if total_wall > 0 else None,
            "plain_ms_per_step": [round(m, 4) for m in plain_ms]}


def local_refinement_leg(args):
    """The reference input's mesh with its boundary-layer local refinement
    (input_file.prm: radii 0.3 / 0.6, boundary layer 0.1, one local
    refinement) at global refinement args.lr_refine, the measured physics:
    ms per step (graph-replayed) and the phase times, beside the same
    geometry refined uniformly one level more (similar DoF count)."""
    import ctypes
    import torch
    import pecs_amd
    L = pecs_amd.lib()
    L.pecs_enable_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.pecs_phase_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
    names = ["rhs_semi", "rhs_elec", "schur_transforms", "carrier_cholesky", "recoveries", "poisson_cond",
             "chol_poisson", "poisson_recover"]
    res = {"geometry": "r1=0.3, r2=0.6, H=1, L=1 (input_file.prm)", "degree": args.degree, "dt": args.dt}
    for key, r, local in (("local", args.lr_refine, 1), ("uniform", args.lr_refine + 1, 0)):
        p = pecs_amd.workload_params(r, args.degree, args.dt)
        p.update(r1=0.3, r2=0.6, H=1.0, L=1.0)
        if local:
            p.update(local_refine=1, bl=0.1)
        t0 = time.perf_counter()
        g = pecs_amd.SolarCellGPU(p)
        setup = time.perf_counter() - t0
        g.initial_state()
        g.step(args.warmup)
        assert L.pecs_enable_timing(g.ctx, 1) == 0
        g.step(args.lr_steps)
        ph = (ctypes.c_double * 8)()
        nst = ctypes.c_int()
        L.pecs_phase_times(g.ctx, ph, 8, ctypes.byref(nst))
        L.pecs_enable_timing(g.ctx, 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(args.lr_steps)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.lr_steps * 1e3
        U = 2 * g.n_semi + 2 * g.n_elec + g.n_poisson
        res[key] = {"global_refinement": r, "local_refinement": local, "boundary_layer": 0.1 if local else 0.0,
                    "n_carrier_semi": g.n_semi, "n_carrier_elec": g.n_elec, "n_poisson": g.n_poisson,
                    "setup_s": round(setup, 2), "ms_per_step": ms, "dof_updates_per_s": U / (ms * 1e-3),
                    "phase_ms_per_step": {n: round(ph[i] / max(nst.value, 1), 4) for i, n in enumerate(names)}}
        g.close()
        print(f"[bench] local refinement leg: {key}: {res[key]}", file=sys.stderr, flush=True)
    return res


def warm_start_leg(args):
    """Learned warm start: steady states of args.ws_train_groups x 16 draws
    (refinement and cone angle per group, drawn from --ws-refines; applied
    bias, generation amplitude and doping per instance) train the encoder +
    convolutional decoder (FP32, on the GPU) on 64 x 64 density grids per
    region; on args.ws_test_groups x 16 held-out draws the time to steady
    state from the network's prediction (BF16 inference, projected onto each
    instance's mesh, predicted-residual fallback) is compared with the cold
    start from the equilibrium state (both batched per group and stepped
    together by pecs_sweep) and with the plain path: each draw alone,
    graph-replayed, its steps to steady state times its measured time per
    step."""
    import numpy as np
    import torch
    from pecs_amd import SolarCellGPU
    from pecs_amd import warmstart as ws
    rng = np.random.default_rng(args.ws_seed)
    log = lambda m: print(f"[bench] warm start: {m}", file=sys.stderr, flush=True)
    k, dt = args.degree, args.dt
    refines = tuple(int(r) for r in args.ws_refines.split(","))
    train_groups = ws.draw_groups(rng, args.ws_train_groups, 16, refines)
    test_groups = ws.draw_groups(rng, args.ws_test_groups, 16, refines)
    log(f"generating {16 * len(train_groups)} training steady states (refinements {refines})")
    Xtr, Ytr, steps_tr, wall_tr, setup_tr = ws.generate(train_groups, k, dt, args.steady_tol,
                                                        args.steady_check_every, args.steady_max_steps, log=log)
    log(f"training set done: {wall_tr:.1f} s stepping, {setup_tr:.1f} s setup")
    t0 = time.perf_counter()
    net, norm, hist = ws.train(Xtr, Ytr, epochs=args.ws_epochs, log=log)
    t_train = time.perf_counter() - t0
    log(f"cold start of {16 * len(test_groups)} held-out draws")
    steps_cold, wall_cold, steady = ws.cold_start(test_groups, k, dt, args.steady_tol, args.steady_check_every,
                                                  args.steady_max_steps, log)
    log("warm start of the held-out draws")
    ev = ws.evaluate(net, norm, test_groups, k, dt, args.steady_tol, args.steady_check_every,
                     args.steady_max_steps, fallback_residual=args.ws_fallback, log=log)
    # the plain path: each draw alone, graph-replayed; its time per step
    # measured on the first 4 draws of each test group
    plain_ms = []
    for ref, ang, inst in test_groups:
        ms = []
        for d in inst[:4]:
            g = SolarCellGPU(ws.instance_params(ref, k, dt, ang, *d))
            g.initial_state()
            g.step(5)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            g.step(100)
            torch.cuda.synchronize()
            ms.append((time.perf_counter() - t1) / 100 * 1e3)
            g.close()
        plain_ms.append(float(np.mean(ms)))
    per_draw_ms = np.repeat(plain_ms, [len(inst) for _, _, inst in test_groups])
    plain_s = float(np.sum(np.abs(steps_cold) * per_draw_ms * 1e-3))
    p32 = ws.predict(net, norm, test_groups, bf16=False)
    err = np.array([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(ev["pred"], steady)])
    err32 = np.array([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(p32, steady)])
    n_te = len(steady)
    warm_total = ev["wall"] + ev["predict_s"]
    # the measured workload's refinement: one group of 16 held-out draws on
    # one planar mesh (cone angle 0: the cell-type lists of the planar
    # workload apply), a batched context warm-started from the same network
    # (its 64 x 64 grids projected onto refinement args.ws_eval_refine)
    big = None
    if args.ws_eval_refine > 0:
        grp = [(args.ws_eval_refine, 0.0, inst) for _, _, inst in ws.draw_groups(rng, 1, 16, refines)]
        log(f"refinement {args.ws_eval_refine}: cold start of 16 held-out draws")
        bc, bwall, bsteady = ws.cold_start(grp, k, dt, args.steady_tol, args.steady_check_every,
                                           args.steady_max_steps, log)
        log(f"refinement {args.ws_eval_refine}: warm start of the same draws")
        bev = ws.evaluate(net, norm, grp, k, dt, args.steady_tol, args.steady_check_every, args.steady_max_steps,
                          fallback_residual=args.ws_fallback, log=log)
        g = SolarCellGPU(ws.instance_params(args.ws_eval_refine, k, dt, grp[0][1], *grp[0][2][0]))
        g.initial_state()
        g.step(5)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g.step(50)
        torch.cuda.synchronize()
        bms = (time.perf_counter() - t1) / 50 * 1e3
        g.close()
        bplain = float(np.mean(np.abs(bc))) * bms * 1e-3
        bwarm = (bev["wall"] + bev["predict_s"]) / 16
        berr = np.array([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(bev["pred"], bsteady)])
        big = {"refinement": args.ws_eval_refine, "draws": 16, "cold_steps_mean": float(np.mean(bc)),
               "warm_steps_mean": float(bev["steps"].mean()), "fallbacks": bev["fallbacks"],
               "prediction_rel_error": {"median": float(np.median(berr)), "max": float(berr.max())},
               "cold_s_per_bias_point": bwall / 16, "warm_s_per_bias_point": bwarm,
               "plain_ms_per_step": bms, "plain_s_per_bias_point": bplain,
               "speedup_warm_vs_cold": (bwall / 16) / bwarm if bwarm > 0 else None,
               "speedup_batched_warm_vs_plain": bplain / bwarm if bwarm > 0 else None}
        log(f"refinement {args.ws_eval_refine}: {big}")
    net.close()
    d = ws.DECODER
    return {"measured_refinement": big,
            "refinements_train": sorted({g[0] for g in train_groups}),
            "refinements_test": [g[0] for g in test_groups], "train_draws": len(Ytr), "test_draws": n_te,
            "grid": f"{ws.GRID} x {ws.GRID} per region, 4 density channels",
            "network": (f"encoder 4 -> {d['layers']} x {d['hidden']} ReLU -> {d['c0']} x {d['g0']} x {d['g0']}; "
                        f"decoder {d['stages']} (channels, x2 upsampling) 3x3 convolutions; FP32 MFMA training, "
                        f"BF16 MFMA inference"),
            "epochs": args.ws_epochs, "fallback_residual": args.ws_fallback,
            "train_s": round(t_train, 2), "train_loss_first_last": [hist[0], hist[-1]],
            "prediction_rel_error_bf16": {"median": float(np.median(err)), "max": float(err.max())},
            "prediction_rel_error_fp32": {"median": float(np.median(err32)), "max": float(err32.max())},
            "predicted_residual": {"median": float(np.median(ev["predicted_residual"])),
                                   "max": float(ev["predicted_residual"].max())},
            "fallbacks": ev["fallbacks"], "tol": args.steady_tol,
            "cold_steps_mean": float(np.mean(steps_cold)), "warm_steps_mean": float(ev["steps"].mean()),
            "warm_reached": bool(np.all(ev["steps"] > 0)), "cold_reached": bool(np.all(steps_cold > 0)),
            "cold_s_per_bias_point": wall_cold / n_te,
            "warm_s_per_bias_point": warm_total / n_te,
            "warm_predict_and_project_s": ev["predict_s"],
            "speedup_warm_vs_cold": wall_cold / warm_total if warm_total > 0 else None,
            "plain_ms_per_step": plain_ms, "plain_s_per_bias_point": plain_s / n_te,
            "plain_path": "each draw alone, graph-replayed steps",
            "speedup_batched_warm_vs_plain": plain_s / warm_total if warm_total > 0 else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--refine", type=int, default=8)
    ap.add_argument("--degree", type=int, default=2)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5, help="steps of the CPU baseline on the measured workload")
    ap.add_argument("--steady-tol", type=float, default=1e-8)
    ap.add_argument("--steady-max-steps", type=int, default=8000)
    ap.add_argument("--steady-check-every", type=int, default=25)
    ap.add_argument("--no-steady", action="store_true", help="skip the time-to-steady-state legs")
    ap.add_argument("--batch", type=int, default=64, help="also run an I-V sweep of this many bias points (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-only", action="store_true", help="run only the batched leg")
    ap.add_argument("--sweep", type=int, default=16,
                    help="uneven-mesh sweep of this many instances (refinement, cone angle, bias, generation and "
                         "doping drawn per instance); 0: skip.  The configuration: --sweep 64 --sweep-rmax 8")
    ap.add_argument("--sweep-only", action="store_true", help="run only the uneven-mesh sweep")
    ap.add_argument("--sweep-rmin", type=int, default=5)
    ap.add_argument("--sweep-rmax", type=int, default=7)
    ap.add_argument("--sweep-seed", type=int, default=20261202)
    ap.add_argument("--sweep-streams", type=int, default=4)
    ap.add_argument("--sweep-plain-steps", type=int, default=20)
    ap.add_argument("--sweep-reserve-gib", type=float, default=64.0,
                    help="stop adding instances to a wave when less device memory than this is free")
    ap.add_argument("--mixed", action="store_true",
                    help="run only the mixed-precision leg (FP32 cell-local contractions) on a batch of --batch")
    ap.add_argument("--mixed-steps", type=int, default=200)
    ap.add_argument("--warm-start", action="store_true", help="run only the learned warm-start leg")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the warm-start leg of the default run")
    ap.add_argument("--ws-refines", default="5,6", help="refinements drawn per group (comma separated)")
    ap.add_argument("--ws-train-groups", type=int, default=16, help="16 draws per group (one mesh)")
    ap.add_argument("--ws-test-groups", type=int, default=2)
    ap.add_argument("--ws-epochs", type=int, default=600)
    ap.add_argument("--ws-eval-refine", type=int, default=8,
                    help="also warm-start one group of 16 held-out draws at this refinement (0: skip)")
    ap.add_argument("--ws-seed", type=int, default=20261202)
    ap.add_argument("--ws-fallback", type=float, default=1e-2,
                    help="restart a warm-started instance cold if its predicted residual (relative density "
                         "change over the first step) is above this")
    ap.add_argument("--lr-refine", type=int, default=6,
                    help="global refinement of the local-refinement leg (0: skip)")
    ap.add_argument("--lr-steps", type=int, default=100)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernel by kernel instead of replaying captured step graphs (counter collection)")
    ap.add_argument("--copy-calibration-gib", type=float, default=0.0,
                    help="only run a plain device copy of this many GiB (counter calibration) and exit")
    ap.add_argument("--mfma-calibration", type=int, default=0,
                    help="only run an FP64 matrix product of this size (torch.mm, 2 n^3 flops; MFMA counter "
                         "calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import pecs_amd

    if not torch.cuda.is_available():
        raise SystemExit("no GPU visible")
    if args.copy_calibration_gib > 0:
        n = int(args.copy_calibration_gib * (1 << 30)) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 8 * n, "bytes_written": 8 * n}), flush=True)
        return
    if args.mfma_calibration > 0:
        n = args.mfma_calibration
        a = torch.rand(n, n, dtype=torch.float64, device="cuda")
        b = torch.rand(n, n, dtype=torch.float64, device="cuda")
        for _ in range(3):
            c = torch.mm(a, b)
        torch.cuda.synchronize()
        print(json.dumps({"mfma_calibration_flops_per_product": 2.0 * n ** 3, "n": n,
                          "check": float(c[0, 0])}), flush=True)
        return
    if args.warm_start:
        print(json.dumps({"warm_start": warm_start_leg(args)}), flush=True)
        return
    if args.mixed:
        wl = (f"PECS half cell, refinement {args.refine} ({1 << args.refine}x{1 << args.refine} quads/region), "
              f"LDG k={args.degree}, RT{args.degree - 1}xDGQ{args.degree - 1} Poisson, dt={args.dt}")
        print(json.dumps({"mixed_precision": mixed_leg(args, wl)}), flush=True)
        return
    if args.sweep_only:
        print(json.dumps({"sweep": sweep_leg(args)}), flush=True)
        return
    if args.batch_only:
        wl = (f"PECS half cell, refinement {args.refine} ({1 << args.refine}x{1 << args.refine} quads/region), "
              f"LDG k={args.degree}, RT{args.degree - 1}xDGQ{args.degree - 1} Poisson, dt={args.dt}")
        print(json.dumps({"batched": batched_leg(args, wl)}), flush=True)
        return
    p = pecs_amd.workload_params(args.refine, args.degree, args.dt)
    t0 = time.perf_counter()
    g = pecs_amd.SolarCellGPU(p)
    setup_s = time.perf_counter() - t0
    if args.no_graph:
        g.use_graph(False)
    st = g.stats()
    print(f"[bench] setup {setup_s:.1f}s (factor {st['factor_s']:.1f}s) sizes semi={g.n_semi} "
          f"elec={g.n_elec} poisson={g.n_poisson}", file=sys.stderr, flush=True)
    g.initial_state()
    g.step(args.warmup)
    L = pecs_amd.lib()
    import ctypes
    L.pecs_enable_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.pecs_phase_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int)]
    L.pecs_factor_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    assert L.pecs_enable_timing(g.ctx, 1) == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    res = g.residual()
    ph = (ctypes.c_double * 8)()
    nst = ctypes.c_int()
    L.pecs_phase_times(g.ctx, ph, 8, ctypes.byref(nst))
    L.pecs_enable_timing(g.ctx, 0)
    phase = [ph[i] / max(nst.value, 1) for i in range(8)]
    # the same number of steps replayed as captured graphs (no timing events)
    ms_graph = None
    if not args.no_graph:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.step(args.steps)
        torch.cuda.synchronize()
        ms_graph = (time.perf_counter() - t0) / args.steps * 1e3
    names = ["rhs_semi", "rhs_elec", "schur_transforms", "carrier_cholesky", "recoveries", "poisson_cond",
             "chol_poisson", "poisson_recover"]
    print("[bench] per-step phase ms: " + ", ".join(f"{n}={v:.3f}" for n, v in zip(names, phase)),
          file=sys.stderr, flush=True)
    # the dominant phase: forward + backward sweeps with both regions' Schur
    # factors (run concurrently), two right-hand sides per factor read;
    # vectors per region: 2 rhs x (read b, write y, read y, write x)
    fb = ctypes.c_double()
    chol_bytes = 0.0
    # (the workload's carriers of a region have equal mobilities, so each
    # region has one Schur factor serving both of its carriers)
    for which, n_schur in ((0, g.n_semi // 3), (1, g.n_elec // 3)):
        L.pecs_factor_bytes(g.ctx, which, ctypes.byref(fb))
        chol_bytes += fb.value + 2 * 4 * 8.0 * n_schur
    chol_ms = phase[3]
    achieved = chol_bytes / (chol_ms * 1e-3) / 1e9 if chol_ms > 0 else 0.0
    workload = (f"PECS half cell, refinement {args.refine} ({1 << args.refine}x{1 << args.refine} quads/region), "
                f"LDG k={args.degree}, RT{args.degree - 1}xDGQ{args.degree - 1} Poisson, dt={args.dt}")
    U = 2 * g.n_semi + 2 * g.n_elec + g.n_poisson
    # the production path (pecs_step replaying the captured step graphs, the
    # two regions' chains on two streams) gives value; the run with
    # per-phase events (kernel by kernel, phases one after another) gives the
    # phase times and the roofline's kernel time
    ms_timed = elapsed / args.steps * 1e3
    ms_per_step = ms_graph if ms_graph is not None else ms_timed
    value = U / (ms_per_step * 1e-3)
    out = {
        "metric": "time-to-steady-state per bias point (s) + DOF-updates/s, GPU vs plain path",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: equilibrium initial state, configuration physics",
        "config": {"workload": workload,
                   "dofs_per_step": U, "n_carrier_semi": g.n_semi, "n_carrier_elec": g.n_elec,
                   "n_poisson": g.n_poisson, "setup_s": round(setup_s, 2),
                   "steady_residual_last_step": res,
                   "ms_per_step_graph_replay": ms_graph,
                   "ms_per_step_timed_phases": ms_timed,
                   "factor_stats": {k: v for k, v in st.items() if k not in ("setup_s",)},
                   "phase_ms_per_step": {n: round(v, 4) for n, v in zip(names, phase)}},
        "roofline": {"kernel": "supernodal triangular sweeps (fwd+bwd) with the semiconductor and electrolyte "
                               "Schur factors, 2 rhs each, one step",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(workload),
                     "traffic_source": "profiles/carrier_solves_r8k2_hbm_traffic.md (rocprofv3 --pmc)",
                     "algorithmic_bytes_per_launch": chol_bytes, "avg_ms": chol_ms},
    }
    out["assembly"] = assembly_rate(args, g, phase)
    out["config"]["sweep_tiles_semi"] = [list(t) for t in g.sweep_config(0)]
    if not args.no_steady:
        tts = gpu_time_to_steady(g, args.steady_tol, args.steady_check_every, args.steady_max_steps)
        tts["workload"] = out["config"]["workload"]
        print(f"[bench] time to steady state: {tts}", file=sys.stderr, flush=True)
        out["time_to_steady_state"] = {"tol": args.steady_tol, "gpu": tts}
    g.close()
    if not args.no_steady:
        small = pecs_amd.SolarCellGPU(pecs_amd.workload_params(4, 1, 1e-3))
        t_small = gpu_time_to_steady(small, args.steady_tol, args.steady_check_every, 20000)
        small.close()
        t_small["workload"] = "refinement 4 (16x16 quads/region), LDG k=1, RT0xDGQ0 Poisson, dt=0.001"
        out["time_to_steady_state"]["small_gpu"] = t_small
        if not args.no_cpu_baseline:
            c = cpu_time_to_steady(args.steady_tol)
            c["workload"] = t_small["workload"]
            out["time_to_steady_state"]["small_cpu"] = c
        print(f"[bench] small workload time to steady state: {out['time_to_steady_state']}", file=sys.stderr,
              flush=True)
    if args.batch > 0:
        out["batched"] = batched_leg(args, workload)
    if args.sweep > 0:
        out["sweep"] = sweep_leg(args)
    if not args.no_warm_start:
        out["warm_start"] = warm_start_leg(args)
    if args.lr_refine > 0:
        out["local_refinement"] = local_refinement_leg(args)
    # every speed-up against the plain path: one instance alone, graph-replayed
    plain = out.get("time_to_steady_state", {}).get("gpu", {})
    if plain.get("reached"):
        b = out.get("batched", {}).get("time_to_steady_state", {})
        if b.get("reached"):
            out["batched"]["speedup_vs_plain_graph_replay"] = plain["s"] / b["s_per_bias_point"]
    if not args.no_cpu_baseline:
        try:
            tts = out.get("time_to_steady_state", {}).get("gpu", {})
            out["cpu_baseline"] = cpu_baseline(args, tts.get("steps") if tts.get("reached") else None)
        except Exception as e:  # report, do not hide
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
