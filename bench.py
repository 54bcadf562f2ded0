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
cpu_baseline = the numpy/scipy oracle (a port of the reference algorithm,
single core) on a bounded sample, reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_baseline(seconds_budget=20.0):
    """Oracle (numpy/scipy, one core) on a bounded sample: refinement 5, k = 2."""
    from oracle.solarcell import OracleSolarCell
    from pecs_amd import workload_params
    p = workload_params(5, 2, 5e-4)
    o = OracleSolarCell(p)
    st = o.initial_state()
    st = o.step(st)
    n = 0
    t0 = time.perf_counter()
    while True:
        st = o.step(st)
        n += 1
        if time.perf_counter() - t0 > seconds_budget:
            break
    dt = time.perf_counter() - t0
    U = 4 * o.Nc + o.pfe.n
    return {"value": U * n / dt, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle (numpy/scipy SuperLU, 1 core), refinement 5 (32x32 cells/region), k=2, "
                      f"{n} steps, U={U} DOFs/step, setup excluded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--refine", type=int, default=8)
    ap.add_argument("--degree", type=int, default=2)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration-gib", type=float, default=0.0,
                    help="only run a plain device copy of this many GiB (counter calibration) and exit")
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
    p = pecs_amd.workload_params(args.refine, args.degree, args.dt)
    t0 = time.perf_counter()
    g = pecs_amd.SolarCellGPU(p)
    setup_s = time.perf_counter() - t0
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
    ph = (ctypes.c_double * 11)()
    nst = ctypes.c_int()
    L.pecs_phase_times(g.ctx, ph, 11, ctypes.byref(nst))
    phase = [ph[i] / max(nst.value, 1) for i in range(11)]
    names = ["rhs_semi", "rhs_elec", "schur_rhs_semi", "chol_semi", "recover_semi", "schur_rhs_elec",
             "chol_elec", "recover_elec", "poisson_cond", "chol_poisson", "poisson_recover"]
    print("[bench] per-step phase ms: " + ", ".join(f"{n}={v:.3f}" for n, v in zip(names, phase)),
          file=sys.stderr, flush=True)
    fb = ctypes.c_double()
    L.pecs_factor_bytes(g.ctx, 0, ctypes.byref(fb))
    # two right-hand sides share each factor read; vectors: 2 rhs x (read b, write y, read y, write x)
    n_schur = st["n_schur_semi"]
    chol_bytes = fb.value + 2 * 4 * 8.0 * n_schur
    chol_ms = phase[3]
    achieved = chol_bytes / (chol_ms * 1e-3) / 1e9 if chol_ms > 0 else 0.0
    U = 2 * g.n_semi + 2 * g.n_elec + g.n_poisson
    ms_per_step = elapsed / args.steps * 1e3
    value = U * args.steps / elapsed
    out = {
        "metric": "DOF-updates/s (IMEX step: 4 carrier + Poisson unknowns advanced per step)",
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
        "config": {"workload": f"PECS half cell, refinement {args.refine} "
                               f"({1 << args.refine}x{1 << args.refine} quads/region), LDG k={args.degree}, "
                               f"RT{args.degree - 1}xDGQ{args.degree - 1} Poisson, dt={args.dt}",
                   "dofs_per_step": U, "n_carrier_semi": g.n_semi, "n_carrier_elec": g.n_elec,
                   "n_poisson": g.n_poisson, "setup_s": round(setup_s, 2),
                   "steady_residual_last_step": res,
                   "phase_ms_per_step": {n: round(v, 4) for n, v in zip(names, phase)}},
        "roofline": {"kernel": "supernodal triangular solves (fwd+bwd) of the semiconductor Schur factor, 2 rhs",
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": chol_bytes, "avg_ms": chol_ms},
    }
    g.close()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        except Exception as e:  # report, do not hide
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
