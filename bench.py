"""Benchmark: seconds per Newton step of the F-bar Newton-step hot path at ~3M DOF.

Workload (default): 100^3 Hex8 F-bar block, +-10% interior jitter (seed 0), J2 with
linear isotropic hardening (E=200 GPa, nu=0.3, sy=250 MPa, H=1 GPa), uniaxial
compression to 20% in 40 load steps, PCG + SA-AMG to rtol 1e-8 -- 1,030,301 nodes,
3,090,903 DOF.  Synthetic mesh (no dataset); inputs resident in HBM.

A bench "step" is one load step of newton! plus its commit (Solver.jl:455-472); the
metric value is the timed wall time divided by the number of Newton iterations
(residual evaluations, the reference's SolveResult.iters) in the timed load steps, so
each Newton step carries its share of the per-load-step AMG set-up.

python bench.py [--steps K] [--warmup W] [--precision fp64|fp32_ir] [--n 100]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s/Newton step @3M DOF, rtol 1e-8: FP64 CG+AMG vs FP32-IR+learned PC"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def spmv_bytes(nnzb, nb, ndof, s=8):
    """Algorithmic bytes of one BSR(3x3) SpMV: values + column index per block, row
    pointers, x read once and y written once."""
    return nnzb * (9 * s + 4) + (nb + 1) * 4 + 2 * ndof * s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(wl, newton_iters_per_step, krylov_per_newton, budget_s=30.0):
    """Oracle (C/OpenMP restatement of the reference algorithm) on the host cores: one
    full-size assembly + Dirichlet imposition + norm of the same workload, timed.  The
    linear-solve share is estimated from the oracle's CSC SpMV time times the GPU's
    Krylov count per Newton step times the SpMV-equivalents of one PCG iteration."""
    import numpy as np

    import oracle as O
    nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    nthreads = max(1, min(nthreads, 16))
    t0 = time.perf_counter()
    M = O.OracleModel(wl.mesh.nodes, wl.mesh.elements, O.Material(**wl.materials[0]), nthreads=nthreads)
    M.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    bcs = M.freeze_bcs()
    t_build = time.perf_counter() - t0
    U = np.zeros(M.ndof)
    st = M._new_state(8 * M.ne)
    t1 = time.perf_counter()
    K, R = M.assemble(U, st)
    t_asm = time.perf_counter() - t1
    t2 = time.perf_counter()
    (rd, rv), (fd, fv) = bcs
    O.impose_dirichlet(K, R, rd, rv, True, 1 / 40, U)
    O.impose_dirichlet(K, R, fd, fv, False, 1 / 40, U)
    float(np.sqrt((R ** 2).sum()))
    t_bc = time.perf_counter() - t2
    x = np.ones(M.ndof)
    y = np.empty(M.ndof)
    reps = 3
    t3 = time.perf_counter()
    for _ in range(reps):
        O.lib.or_csc_spmv_omp(M.ndof, K.indptr, K.indices, K.data, x, y, nthreads)
    t_spmv = (time.perf_counter() - t3) / reps
    spmv_equiv = 6.0  # PCG SpMV + Chebyshev(2) pre/post + two residuals per V-cycle (fine level)
    t_solve = krylov_per_newton * spmv_equiv * t_spmv
    per_newton = t_asm + t_bc + t_solve
    return {
        "value": per_newton, "unit": "s/Newton step", "cores": nthreads, "kind": "port",
        "sample": (f"oracle (C/OpenMP) on the full 100^3 workload: 1 assembly {t_asm:.2f}s + BC/norm "
                   f"{t_bc:.2f}s, CSC SpMV {t_spmv * 1e3:.1f} ms x {krylov_per_newton:.1f} PCG its x "
                   f"{spmv_equiv:.0f} SpMV-equivalents (solve share estimated, AMG set-up excluded); "
                   f"model build {t_build:.1f}s untimed"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed load steps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed load steps")
    ap.add_argument("--n", type=int, default=100, help="elements per edge")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_ir"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--amg-rebuild", type=int, default=1, help="0: once per load step, 1: every Newton iteration")
    ap.add_argument("--verbose", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 "
              "(see DESIGN.md for how the path would shard)", file=sys.stderr)
        sys.exit(2)

    import plasticity_amd as pa
    from plasticity_amd.workloads import workload

    wl = workload(1, n=args.n)
    log(f"[bench] workload: {wl.name}, nodes {wl.mesh.nnodes}, dof {3 * wl.mesh.nnodes}")
    opts = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, precision=args.precision,
                              verbose=args.verbose, amg_rebuild=args.amg_rebuild)
    t0 = time.perf_counter()
    m = pa.Model(wl.mesh, pa.J2Material(**wl.materials[0]), options=opts)
    m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    m.reset()
    m._freeze()
    log(f"[bench] model set-up {time.perf_counter() - t0:.2f}s")

    def load_step(k):
        lam = k / wl.nsteps
        conv, hist, kry, res = m.newton(lam, tol=wl.newton_tol, maxiter=25)
        if not conv:
            raise RuntimeError(f"load step {k} did not converge: {hist}")
        m.commit()
        return res

    for k in range(1, args.warmup + 1):
        res = load_step(k)
        log(f"[bench] warmup load step {k}: newton {res.iters} krylov {list(res.krylov_iters[:res.iters])} "
            f"t={res.t_total:.3f}s")
    s0 = m.stats()
    newton = krylov = 0
    t_asm = t_setup = t_solve = 0.0
    t0 = time.perf_counter()
    for k in range(args.warmup + 1, args.warmup + args.steps + 1):
        res = load_step(k)
        newton += res.iters
        krylov += sum(res.krylov_iters[:res.iters])
        t_asm += res.t_assemble
        t_setup += res.t_setup
        t_solve += res.t_solve
        log(f"[bench] load step {k}: newton {res.iters} krylov {list(res.krylov_iters[:res.iters])} "
            f"solver {list(res.solver_used[:res.iters - 1])} t={res.t_total:.3f}s")
    elapsed = time.perf_counter() - t0
    s1 = m.stats()
    per_newton = elapsed / newton
    nnzb, nb, ndof = s1.nnzb, s1.nnodes, s1.ndof
    n_sp = s1.n_pcg_spmv - s0.n_pcg_spmv
    t_sp = (s1.t_pcg_spmv - s0.t_pcg_spmv) / max(n_sp, 1)
    bytes_sp = spmv_bytes(nnzb, nb, ndof, 8 if args.precision == "fp64" else 4)
    achieved = bytes_sp / t_sp / 1e9 if n_sp else None
    n_asm = s1.n_assemble - s0.n_assemble
    gp_ms = (s1.t_gp_kernel - s0.t_gp_kernel) / max(n_asm, 1) * 1e3
    sc_ms = (s1.t_scatter_kernel - s0.t_scatter_kernel) / max(n_asm, 1) * 1e3
    nset = max(s1.n_setup - s0.n_setup, 1)
    log(f"[bench] AMG set-up per rebuild: rho {(s1.t_setup_rho - s0.t_setup_rho) / nset * 1e3:.1f} ms, "
        f"P/RAP {(s1.t_setup_galerkin - s0.t_setup_galerkin) / nset * 1e3:.1f} ms, "
        f"coarse inverse {(s1.t_setup_coarse - s0.t_setup_coarse) / nset * 1e3:.1f} ms ({nset} rebuilds)")
    log(f"[bench] timed {args.steps} load steps: {elapsed:.3f}s, {newton} Newton steps, {krylov} Krylov its; "
        f"assemble {t_asm:.3f}s (gp kernel {gp_ms:.2f} ms + scatter {sc_ms:.2f} ms per assembly), "
        f"AMG set-up {t_setup:.3f}s, solve {t_solve:.3f}s; fine SpMV {t_sp * 1e3:.3f} ms = {achieved} GB/s; "
        f"AMG levels {list(s1.amg_ndof[:s1.amg_levels])} OC {s1.amg_operator_complexity:.3f}")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(wl, newton / args.steps, krylov / max(newton - args.steps, 1))
            log(f"[bench] cpu baseline {cpu}")
        except Exception as e:  # the baseline never blocks the GPU line
            log(f"[bench] cpu baseline failed: {e!r}")
    out = {
        "metric": METRIC, "value": per_newton, "unit": "s/Newton step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64-IR",
        "data": "synthetic (generated jittered 100^3 Hex8 mesh)",
        "config": {"workload": f"{wl.name}: {wl.mesh.nelem} Hex8 F-bar elements, {ndof} DOF, "
                               f"PCG+SA-AMG rtol {wl.krylov_rtol:g}, precision {args.precision}, "
                               f"AMG rebuilt {'every Newton iteration' if args.amg_rebuild else 'once per load step'}",
                   "newton_steps_timed": newton, "krylov_its_timed": krylov,
                   "t_assemble_s": t_asm, "t_amg_setup_s": t_setup, "t_solve_s": t_solve,
                   "amg_levels": list(s1.amg_ndof[:s1.amg_levels])},
        "roofline": {"kernel": "fine BSR(3x3) SpMV (sell_spmv_kernel, FP64) in PCG", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "bytes_per_launch": bytes_sp, "avg_launch_ms": t_sp * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
