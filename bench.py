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

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s/Newton step @3M DOF, rtol 1e-8: FP64 CG+AMG vs FP32-IR+learned PC"
HBM_PEAK_GBS = 8000.0
# FP64 operations per Hex8 F-bar element with all 8 Gauss points plastic, counted with a
# counting-double build (oracle/flopcount.cpp, tests/test_flopcount.py): ELEMENT_FLOPS for
# the structure-exploiting form of the element (the device kernels' algorithm without
# their redundant work; same Fe and Ke as the oracle to 1e-15), ELEMENT_FLOPS_DENSE for
# the oracle's routine as written (dense 9x9 A times the 9x24 Geff, unit-dF columns).
ELEMENT_FLOPS = 83913
ELEMENT_FLOPS_DENSE = 160073
# HBM bytes per fine-SpMV launch at 100^3 (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors
# from a 1 GiB device copy), measured with rocprofv3 --pmc on `--mode spmv --n 100`:
# 1,034,757.0 KiB fetched + 24,148.0 KiB written per dispatch.
SPMV_TRAFFIC_100 = (1034757.0 * 2.0 + 24148.0) * 1024
FP64_PEAK_TFLOPS = 78.6


def spmv_bytes(nnzb, nb, ndof, s=8):
    """Algorithmic bytes of one BSR(3x3) SpMV: values + column index per block, row
    pointers, x read once and y written once."""
    return nnzb * (9 * s + 4) + (nb + 1) * 4 + 2 * ndof * s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    """Host cores this process may use, the machine's count and the CPU model."""
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, os.cpu_count(), model


def cpu_baseline(n=100, maxiter=12):
    """The oracle -- the C/OpenMP restatement of the reference algorithm: colour-ordered
    FP64 assembly, impose_dirichlet!, PCG preconditioned by one SA-AMG V-cycle (rigid-body
    near-null space, damped Jacobi smoothing) -- run on all host cores for the first load
    step of the same workload at full size (100^3, 3,090,903 DOF, Krylov rtol 1e-8 as the
    GPU run), with its own Newton and PCG iteration counts.  The hierarchy is built from
    every Newton iteration's tangent: kept from the first iteration, as the reference does
    (Solver.jl:374), it makes PCG break down on the plastic F-bar tangent (where the
    reference's F-bar path would use UMFPACK).  s/Newton step = wall time of the load step
    / Newton iterations (residual evaluations), as on the GPU."""
    import numpy as np

    import oracle as O
    from plasticity_amd.workloads import workload
    threads, nproc, model = host_cpu()
    wl = workload(1, n=n)
    M = O.OracleModel(wl.mesh.nodes, wl.mesh.elements, O.Material(**wl.materials[0]), nthreads=threads)
    M.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    bcs = M.freeze_bcs()
    (rd, _), (fd, _) = bcs
    isbc = np.zeros(M.ndof, np.uint8)
    isbc[rd] = 1
    isbc[fd] = 1
    st = {"amg": None, "t_setup": 0.0, "t_solve": 0.0, "its": []}

    def solver(K, b, eta):
        t = time.perf_counter()
        st["amg"] = None  # the previous hierarchy is freed before the next is built
        st["amg"] = O.OracleAMG(K, M.nodes, isbc, coarse_max=1000, nthreads=threads)
        st["t_setup"] += time.perf_counter() - t
        t1 = time.perf_counter()
        x, it = st["amg"].solve(b, wl.krylov_rtol, itmax=200)
        st["t_solve"] += time.perf_counter() - t1
        st["its"].append(it)
        return x, abs(it)

    t0 = time.perf_counter()
    conv, hist, _ = M.newton(1 / wl.nsteps, np.zeros(M.ndof), bcs, tol=wl.newton_tol, maxiter=maxiter,
                             solver=solver)
    elapsed = time.perf_counter() - t0
    return {
        "value": elapsed / len(hist), "unit": "s/Newton step", "cores": threads, "kind": "port",
        "nproc": nproc, "cpu_model": model, "threads": threads, "converged": bool(conv),
        "newton_steps": len(hist), "pcg_its": st["its"],
        "sample": (f"oracle (C/OpenMP, {threads} threads of {nproc} logical CPUs, {model}): load step 1 of the "
                   f"same workload at {n}^3 ({M.ndof} DOF{', full size' if n == 100 else ''}), {len(hist)} Newton steps "
                   f"({'converged' if conv else 'not converged'}), PCG its {st['its']} (negative: breakdown), "
                   f"{elapsed:.1f}s wall of which AMG set-up {st['t_setup']:.1f}s and PCG {st['t_solve']:.1f}s"),
    }


def measured_peaks():
    """Achievable device peaks (pfem_device_peaks): best-of-10 STREAM-style copy of 2 GiB
    (read + write bytes), a read-only stream of the same 2 GiB and an FP64 FMA loop.  Reported beside the published peaks the
    roofline fractions use."""
    import ctypes as C
    from plasticity_amd import _lib as L
    gbs, rd, tf = C.c_double(), C.c_double(), C.c_double()
    L.check(L.lib().pfem_device_peaks(2 << 30, 10, C.byref(gbs), C.byref(rd), C.byref(tf)), "pfem_device_peaks")
    return {"copy_GBps": gbs.value, "copy_frac_of_hbm_peak": gbs.value / HBM_PEAK_GBS,
            "read_GBps": rd.value, "read_frac_of_hbm_peak": rd.value / HBM_PEAK_GBS,
            "fp64_fma_TFLOPs": tf.value, "fma_frac_of_fp64_peak": tf.value / FP64_PEAK_TFLOPS}


def copy_calibration(nbytes=1 << 30, reps=10):
    """Plain device-to-device copy of a known size through the HIP runtime (its blit
    kernel reads nbytes and writes nbytes per call): the reference point for the
    FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: {reps} x {nbytes} B device-to-device")


def vcycle_figures(m, prec, reps=20):
    """Live HIP-event timing of the AMG cycle and of one fused fine-level Chebyshev step
    (SELL product with the smoother update as its epilogue) on the model's current
    hierarchy, against their algorithmic bytes (every launch's operands read once and
    outputs written once)."""
    r = m.bench_vcycle(prec=prec, reps=reps)
    cyc_gbs = r["cycle_bytes"] / r["cycle_s"] / 1e9
    st_gbs = r["step_bytes"] / r["step_s"] / 1e9
    return {"precision": "fp32" if prec else "fp64",
            "cycle_ms": r["cycle_s"] * 1e3, "cycle_bytes": r["cycle_bytes"], "cycle_GBps": cyc_gbs,
            "cycle_frac": cyc_gbs / HBM_PEAK_GBS,
            "smoother_step_ms": r["step_s"] * 1e3, "smoother_step_bytes": r["step_bytes"],
            "smoother_step_GBps": st_gbs, "smoother_step_frac": st_gbs / HBM_PEAK_GBS}


def fp32_operator_error(pa, wl, args, report=(1, 2, 5, 10, 20, 40)):
    """For the first Newton system of selected load steps of the workload's path: the
    correction dU (FP64 solve to rtol 1e-10) and e = ||(fl32(K) - K) dU|| / ||K dU||, the
    relative residual that an exact solve with the FP32-rounded K leaves -- the contraction
    per refinement sweep of a Krylov iteration on the FP32 K.  A random vector is shown
    beside it: rounding K to FP32 moves K p by ~1e-7 relative for rough p, far more for
    the smooth corrections of an ill-conditioned tangent."""
    import scipy.sparse as sps
    opts = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol)
    mats = [pa.J2Material(**d) for d in wl.materials]
    m = pa.Model(wl.mesh, mats if len(mats) > 1 else mats[0], options=opts, material_id=wl.material_id)
    m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    m.reset()
    m._freeze()
    rng = np.random.default_rng(0)
    rows = []
    done = 0
    for k in sorted(r for r in report if r <= wl.nsteps):
        for j in range(done + 1, k):
            m.advance(j / wl.nsteps, wl.nsteps, tol=wl.newton_tol, maxiter=args.maxiter, cutback=args.cutback)
        done = k - 1
        lam = k / wl.nsteps
        m.assemble()
        m.subtract_external(lam)
        m.impose_dirichlet(lam)
        b = -m.residual()
        dU = m.linear_solve(1e-10)[0]
        rp, ci, v = m.get_K_bsr()
        K = sps.bsr_matrix((v, ci, rp), shape=(m.ndof, m.ndof))
        K32 = sps.bsr_matrix((v.astype(np.float32).astype(np.float64), ci, rp), shape=(m.ndof, m.ndof))
        y = K @ dU
        p = rng.standard_normal(m.ndof)
        row = {"load_step": k, "solve_relres": float(np.linalg.norm(b - y) / np.linalg.norm(b)),
               "fp32_K_error_on_dU": float(np.linalg.norm(K32 @ dU - y) / np.linalg.norm(y)),
               "fp32_K_error_on_random": float(np.linalg.norm(K32 @ p - K @ p) / np.linalg.norm(K @ p))}
        del K, K32, v
        rows.append(row)
        log(f"[bench] fp32 operator, load step {k}: {row}")
    print(json.dumps({"workload": wl.name, "ndof": m.ndof, "rows": rows}), flush=True)
    m.close()


class PathTotals:
    """Sums over the load increments of timed load steps (advance records)."""

    def __init__(self):
        self.newton = self.accepted = self.krylov = self.increments = self.cutbacks = 0
        self.t_asm = self.t_setup = self.t_solve = 0.0

    def add(self, recs):
        for r in recs:
            self.increments += 1
            self.cutbacks += 0 if r["converged"] else 1
            self.newton += r["iters"]
            self.accepted += r["iters"] if r["converged"] else 0
            self.krylov += sum(r["krylov"])
            t = r.get("times") or {}
            self.t_asm += t.get("assemble", 0.0)
            self.t_setup += t.get("setup", 0.0)
            self.t_solve += t.get("solve", 0.0)


def time_variant(pa, wl, precision, learned, warmup, steps, args):
    """The same load steps of the workload on a fresh model with another linear solver:
    precision "fp32_ir", learned = path of U-net weights (GCR with the network, AMG
    fallback) or None.  Returns s/Newton step over the timed load steps and the counts."""
    opts = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, precision=precision,
                              amg_rebuild=args.amg_rebuild, newton_diverge=args.diverge,
                              fp32_krylov=args.fp32_krylov, assembly_chunk=args.assembly_chunk,
                              smoother_unfused=args.smoother_unfused, coarse_degree=args.coarse_degree,
                              krylov_sync=args.krylov_sync, split_operator=args.split_operator,
                              vcycle_eager=args.vcycle_eager)
    mats = [pa.J2Material(**d) for d in wl.materials]
    m = pa.Model(wl.mesh, mats if len(mats) > 1 else mats[0], options=opts, material_id=wl.material_id)
    m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    m.reset()
    m._freeze()
    if learned:
        from plasticity_amd.learned import UNet, grid_of
        net = UNet(grid_of(wl.mesh)).load(learned)
        m.attach_net(net, "precondition")
    tot = PathTotals()
    t0 = None
    for k in range(1, warmup + steps + 1):
        if k == warmup + 1:
            t0 = time.perf_counter()
        t1 = time.perf_counter()
        recs = m.advance(k / wl.nsteps, wl.nsteps, tol=wl.newton_tol, maxiter=args.maxiter, cutback=args.cutback,
                         grow_iters=args.grow_iters)
        if k > warmup:
            tot.add(recs)
        log(f"[bench] {precision}{' + learned' if learned else ''} load step {k}: "
            f"{sum(r['iters'] for r in recs)} Newton, {sum(sum(r['krylov']) for r in recs)} Krylov, "
            f"{time.perf_counter() - t1:.3f}s")
    elapsed = time.perf_counter() - t0
    U, abar = m.U, m.get_state()["abar"]
    vc = None
    if not learned:
        try:
            vc = vcycle_figures(m, 1 if precision == "fp32_ir" else 0)
        except Exception as e:
            log(f"[bench] V-cycle timing failed: {e!r}")
    if learned:
        m.attach_net(None, "off")
    m.close()
    return {"value": elapsed / tot.newton, "newton_steps": tot.newton, "krylov_its": tot.krylov,
            "increments": tot.increments, "cutbacks": tot.cutbacks, "seconds": elapsed,
            "t_assemble_s": tot.t_asm, "t_amg_setup_s": tot.t_setup, "t_solve_s": tot.t_solve,
            **({"vcycle": vc} if vc else {})}, U, abar


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed load steps (default: the rest of the workload's load path after the warm-up, "
                         "39 of workload 1's 40)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed load steps")
    ap.add_argument("--n", type=int, default=None, help="elements per edge (default: the workload's own)")
    ap.add_argument("--workload", type=int, default=1, choices=[1, 2, 4],
                    help="1: 100^3 jittered compression (the metric), 2: 200^3 two-phase tension, "
                         "4: 150^3 notched bar tension")
    ap.add_argument("--nsteps", type=int, default=None,
                    help="load steps of the full load path (default: the workload's own)")
    ap.add_argument("--line-search", type=int, default=0, metavar="HALVINGS",
                    help="Newton backtracking on ||R|| (not in the reference; 0 = full Newton steps)")
    ap.add_argument("--rho-iters", type=int, nargs=2, default=(0, 0), metavar=("COLD", "WARM"),
                    help="power-iteration steps of the smoother's rho estimate (0 0: library defaults)")
    ap.add_argument("--compare", action="store_true",
                    help="after the timed run, repeat the same load steps untimed with the plain FP64 path "
                         "and report the relative displacement and equivalent-plastic-strain differences")
    ap.add_argument("--learned", default=None, metavar="WEIGHTS",
                    help="precondition with the learned U-net (weights .npz) in flexible GCR, AMG fallback")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_ir"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=100,
                    help="elements per edge of the CPU baseline's run (default: the workload's full 100^3)")
    ap.add_argument("--no-variants", action="store_true",
                    help="FP64 run of workload 1: skip timing the FP32-IR and FP32-IR + learned solvers "
                         "on the same load steps")
    ap.add_argument("--amg-rebuild", type=int, default=7,
                    help="0: once per load step (the reference), 1: every Newton iteration, 4: at the first "
                         "Newton iteration of a load step, then all but the prolongators at every iteration, "
                         "5: the prolongators once, everything else at every Newton iteration, 6: as 5, the "
                         "prolongators rebuilt after a solve that needed a fallback, 7: as 6, each Newton "
                         "iteration but the second reuses the coarse levels of the same iteration of an earlier "
                         "load step")
    ap.add_argument("--smoother", default="chebyshev", choices=["chebyshev", "chebyshev4", "jacobi"])
    ap.add_argument("--smoother-degree", type=int, default=0,
                    help="Chebyshev degree / Jacobi sweeps (0: library default, 2 / 1)")
    ap.add_argument("--cycle", default="v", choices=["v", "w"], help="AMG cycle")
    ap.add_argument("--coarse-degree", type=int, default=0,
                    help="Chebyshev degree on the levels below the finest (0: the fine level's)")
    ap.add_argument("--fallback", default="gcr", choices=["bicgstab", "gcr"],
                    help="Krylov method after a CG breakdown")
    ap.add_argument("--cutback", type=int, default=6,
                    help="load-step bisection: a failed load increment is undone and halved, at most this many "
                         "times below the nominal increment (0: stop at the first failed load step, the "
                         "reference's solve!)")
    ap.add_argument("--diverge", type=float, default=0.0,
                    help="end a diverging Newton loop early when ||R|| exceeds this multiple of its best value "
                         "(0: run maxiter iterations, the reference)")
    ap.add_argument("--maxiter", type=int, default=25, help="Newton iterations per load increment")
    ap.add_argument("--newton-tol", type=float, default=None,
                    help="Newton tolerance on the reference-scaled residual (default: the workload's, 1e-8)")
    ap.add_argument("--grow-iters", type=int, default=6,
                    help="a cut increment doubles again only after two converged increments that each took at "
                         "most this many Newton iterations")
    ap.add_argument("--verbose", type=int, default=0)
    ap.add_argument("--progress", action="store_true", help="one stderr line per load increment")
    ap.add_argument("--show-residuals", action="store_true",
                    help="print each increment's Newton residual history in the per-load-step lines")
    ap.add_argument("--vcycle-eager", type=int, default=0, choices=[0, 1],
                    help="1: launch the AMG cycle's kernels one by one instead of replaying a HIP graph")
    ap.add_argument("--split-operator", type=int, default=1, choices=[0, 1],
                    help="FP32-IR: 1 = flexible CG's K p from the FP32 copy of K plus a BF16 correction "
                         "(default); 0 = from the FP64 K")
    ap.add_argument("--krylov-sync", type=int, default=0, choices=[0, 1],
                    help="1: PCG / flexible CG read each iteration's scalars before launching the next "
                         "product (0: the next product is queued first; same results)")
    ap.add_argument("--smoother-unfused", type=int, default=0, choices=[0, 1],
                    help="1: fine-level Chebyshev steps as a product launch plus a vector launch")
    ap.add_argument("--assembly-chunk", type=int, default=0,
                    help="> 0: pipelined assembly, the Gauss-point kernel of one chunk of this many elements "
                         "beside the scatter of the previous chunk (0: whole-mesh Gauss-point kernel, then "
                         "the scatter)")
    ap.add_argument("--fp32-krylov", type=int, default=0, choices=[0, 1],
                    help="FP32-IR: 1 = the Krylov iteration itself in FP32 inside FP64 iterative refinement, "
                         "0 = FP64 flexible CG with the FP32 V-cycle")
    ap.add_argument("--fp16-fine", action="store_true",
                    help="FP32-IR: fine-level smoother products with an FP16 copy of the scaled K")
    ap.add_argument("--mode", default="newton", choices=["newton", "spmv", "assemble", "copy", "fp32-operator"],
                    help="newton: the metric; spmv: only repeated fine SpMVs (for counter runs); assemble: only "
                         "repeated tangent assemblies at a plastic state (for counter runs); "
                         "copy: a plain 1 GiB device-to-device copy, the calibration of byte counters; "
                         "fp32-operator: how far the FP32-rounded tangent moves K dU for the Newton corrections of "
                         "the load path (the limit of a Krylov iteration on the FP32 K)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this implementation runs the Newton-step path on one GPU only; --gpus must be 1 "
              "(multi-GPU sharding by element slabs with node halos is designed but not built)",
              file=sys.stderr)
        sys.exit(2)

    if args.mode == "copy":
        return copy_calibration()

    import plasticity_amd as pa
    from plasticity_amd.workloads import workload

    t0 = time.perf_counter()
    wl = workload(args.workload, n=args.n, nsteps=args.nsteps, device=True)
    if args.newton_tol is not None:
        wl.newton_tol = args.newton_tol
    if args.steps is None:  # the rest of the load path after the warm-up
        args.steps = max(1, wl.nsteps - args.warmup)
    log(f"[bench] workload generated in {time.perf_counter() - t0:.2f}s (mesh on the device)")
    if args.mode == "spmv":
        m = pa.Model(wl.mesh, pa.J2Material(**wl.materials[0]), options=pa.default_options())
        m.assemble(np.zeros(m.ndof))
        st = m.stats()
        # FP64, the FP32 copy, and FP64 from the FP32 + BF16 split copy (values 4 + 2 bytes,
        # vectors 8 bytes)
        for prec, name in ((0, "FP64"), (1, "FP32"), (2, "split FP32+BF16")):
            sec = m.bench_spmv(prec=prec, reps=20)
            if prec == 2:
                b = st.nnzb * (9 * 6 + 4) + (st.nnodes + 1) * 4 + 2 * st.ndof * 8
            else:
                b = spmv_bytes(st.nnzb, st.nnodes, st.ndof, 8 if prec == 0 else 4)
            log(f"[bench] fine SpMV {name} x20: {sec * 1e3:.3f} ms each, {b / sec / 1e9:.0f} GB/s algorithmic "
                f"({b} B)")
        return
    if args.mode == "assemble":
        m = pa.Model(wl.mesh, pa.J2Material(**wl.materials[0]), options=pa.default_options())
        # homogeneous 1% compression (every Gauss point plastic: yield strain 0.125%)
        U = (np.array([-0.01, 0.003, 0.003]) * wl.mesh.nodes).ravel()
        m.assemble(U)
        s0 = m.stats()
        reps = 5
        for _ in range(reps):
            m.assemble(U)
        s1 = m.stats()
        gp = (s1.t_gp_kernel - s0.t_gp_kernel) / reps
        sc = (s1.t_scatter_kernel - s0.t_scatter_kernel) / reps
        log(f"[bench] assembly x{reps} at 1% compression: gp kernel {gp * 1e3:.3f} ms + scatter {sc * 1e3:.3f} ms "
            f"per assembly ({wl.mesh.nelem} elements)")
        return
    if args.mode == "fp32-operator":
        return fp32_operator_error(pa, wl, args)
    log(f"[bench] workload: {wl.name}, nodes {wl.mesh.nnodes}, dof {3 * wl.mesh.nnodes}")
    opts = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, precision=args.precision,
                              verbose=args.verbose, amg_rebuild=args.amg_rebuild, fp16_fine=args.fp16_fine,
                              line_search=args.line_search, rho_iters=args.rho_iters, smoother=args.smoother,
                              smoother_degree=args.smoother_degree,
                              wcycle=args.cycle == "w",
                              gcr_fallback=args.fallback == "gcr", newton_diverge=args.diverge,
                              fp32_krylov=args.fp32_krylov, assembly_chunk=args.assembly_chunk,
                              smoother_unfused=args.smoother_unfused, coarse_degree=args.coarse_degree,
                              krylov_sync=args.krylov_sync, split_operator=args.split_operator,
                              vcycle_eager=args.vcycle_eager)
    t0 = time.perf_counter()
    mats = [pa.J2Material(**d) for d in wl.materials]
    m = pa.Model(wl.mesh, mats if len(mats) > 1 else mats[0], options=opts, material_id=wl.material_id)
    m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    m.reset()
    m._freeze()
    net = None
    if args.learned:
        from plasticity_amd.learned import UNet, grid_of
        net = UNet(grid_of(wl.mesh)).load(args.learned)
        m.attach_net(net, "precondition")
    st0 = m.stats()
    log(f"[bench] model set-up {time.perf_counter() - t0:.2f}s (pattern, slots, colours "
        f"{st0.t_create_pattern:.3f}s, built on the device: {'pattern' if st0.built_on_device & 1 else 'host'}"
        f"/{'colours' if st0.built_on_device & 2 else 'host'})")

    def load_step(k):
        # newton! + commit, with the failed increments bisected (Model.advance)
        return m.advance(k / wl.nsteps, wl.nsteps, tol=wl.newton_tol, maxiter=args.maxiter, cutback=args.cutback,
                         grow_iters=args.grow_iters, verbose=args.progress)

    def show(k, recs, t):
        inc = ", ".join(f"{r['lam'] * wl.nsteps:.4g}:{r['iters']}{'' if r['converged'] else 'x'}" for r in recs)
        kry = [r["krylov"] for r in recs if r["converged"]]
        sol = [r["solver_used"] for r in recs if r["converged"]]
        res = ""
        if args.show_residuals:
            res = " residuals " + str([[float(f"{x:.3e}") for x in r["residuals"]] for r in recs])
        return f"load step {k}: increments [{inc}] krylov {kry} solver {sol}{res} t={t:.3f}s"

    path = []  # load factors of the committed increments, replayed by --compare

    for k in range(1, args.warmup + 1):
        t1 = time.perf_counter()
        recs = load_step(k)
        path += [r["lam"] for r in recs if r["converged"]]
        log("[bench] warmup " + show(k, recs, time.perf_counter() - t1))
    s0 = m.stats()
    tot = PathTotals()
    t0 = time.perf_counter()
    for k in range(args.warmup + 1, args.warmup + args.steps + 1):
        t1 = time.perf_counter()
        recs = load_step(k)
        tot.add(recs)
        path += [r["lam"] for r in recs if r["converged"]]
        log("[bench] " + show(k, recs, time.perf_counter() - t1))
    elapsed = time.perf_counter() - t0
    newton, krylov, t_asm, t_setup, t_solve = tot.newton, tot.krylov, tot.t_asm, tot.t_setup, tot.t_solve
    s1 = m.stats()
    per_newton = elapsed / newton
    nnzb, nb, ndof = s1.nnzb, s1.nnodes, s1.ndof
    n_sp = s1.n_pcg_spmv - s0.n_pcg_spmv
    t_sp = (s1.t_pcg_spmv - s0.t_pcg_spmv) / max(n_sp, 1)
    # the Krylov operator: the FP64 K, or under FP32-IR with the split operator the FP32
    # copy of K plus its BF16 correction (values 4 + 2 bytes, vectors FP64)
    split = args.precision == "fp32_ir" and args.split_operator == 1 and net is None and not args.fp32_krylov
    if split:
        bytes_sp = nnzb * (9 * 6 + 4) + (nb + 1) * 4 + 2 * ndof * 8
        sp_kernel = "fine BSR(3x3) product from the FP32 + BF16 split copy of K (sell_split_kernel<2>) in flexible CG"
    else:
        bytes_sp = spmv_bytes(nnzb, nb, ndof, 8)
        sp_kernel = "fine BSR(3x3) SpMV (sell_lds_kernel<double,double,2>, FP64) in PCG"
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
    n32 = s1.n_krylov_spmv32 - s0.n_krylov_spmv32
    spmv32 = None
    if n32:
        t32 = (s1.t_krylov_spmv32 - s0.t_krylov_spmv32) / n32
        b32 = spmv_bytes(nnzb, nb, ndof, 4)
        spmv32 = {"avg_launch_ms": t32 * 1e3, "GBps": b32 / t32 / 1e9, "bytes_per_launch": b32, "launches": n32}
        log(f"[bench] FP32 Krylov SpMV {t32 * 1e3:.3f} ms = {b32 / t32 / 1e9:.0f} GB/s ({n32} launches)")
    vcyc = None
    if net is None:
        try:
            vcyc = vcycle_figures(m, 1 if args.precision == "fp32_ir" else 0)
            log(f"[bench] AMG cycle ({vcyc['precision']}): {vcyc['cycle_ms']:.3f} ms = {vcyc['cycle_GBps']:.0f} GB/s; "
                f"fine smoother step {vcyc['smoother_step_ms']:.3f} ms = {vcyc['smoother_step_GBps']:.0f} GB/s")
        except Exception as e:  # the measurement never blocks the line
            log(f"[bench] V-cycle timing failed: {e!r}")
    compare = None
    if args.compare:
        U_run, a_run = m.U, m.get_state()["abar"]
        m.attach_net(None, "off") if net is not None else None
        m.close()
        o64 = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, amg_rebuild=args.amg_rebuild,
                                 line_search=args.line_search, newton_diverge=args.diverge,
                                 assembly_chunk=args.assembly_chunk)
        m64 = pa.Model(wl.mesh, mats if len(mats) > 1 else mats[0], options=o64, material_id=wl.material_id)
        m64.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
        m64.reset()
        m64._freeze()
        # the same committed load increments as the run above (plasticity is path dependent):
        # the difference is the precision's, not the increments'
        t_rep = time.perf_counter()
        for i, lam in enumerate(path):
            conv, hist, _, _ = m64.newton(lam, tol=wl.newton_tol, maxiter=args.maxiter)
            if not conv:
                raise RuntimeError(f"FP64 comparison run did not converge at lambda={lam}")
            m64.commit()
            if (i + 1) % 10 == 0 or i + 1 == len(path):
                log(f"[bench] FP64 replay: {i + 1} of {len(path)} increments, "
                    f"{time.perf_counter() - t_rep:.1f}s")
        U64, a64 = m64.U, m64.get_state()["abar"]
        compare = {"reference": "plain FP64 GPU path (PCG + SA-AMG), the same load increments",
                   "increments": len(path), "newton_tol": wl.newton_tol,
                   "u_rel_err": float(np.linalg.norm(U_run - U64) / np.linalg.norm(U64)),
                   "abar_rel_err": float(np.linalg.norm(a_run - a64) / max(np.linalg.norm(a64), 1e-300))}
        log(f"[bench] vs FP64 path: displacement rel err {compare['u_rel_err']:.2e}, "
            f"abar rel err {compare['abar_rel_err']:.2e}")
        m64.close()
    variants = None
    if (args.workload == 1 and args.precision == "fp64" and not args.learned and not args.no_variants
            and not args.compare):
        # the metric compares the FP64 drop-in with FP32-IR + the learned preconditioner:
        # time both on the same load steps, each on a fresh model
        m.close()
        weights = os.path.join(ROOT, "plasticity_amd", "weights", "unet_16_32_64_64.npz")
        variants = {}
        for k, learned in (("fp32_ir", None), ("fp32_ir_learned", weights if os.path.exists(weights) else None)):
            if k == "fp32_ir_learned" and learned is None:
                continue
            try:
                v, _, _ = time_variant(pa, wl, "fp32_ir", learned, args.warmup, args.steps, args)
            except Exception as e:  # a failed variant is reported, the FP64 line still printed
                variants[k] = {"error": repr(e)}
                log(f"[bench] {k} failed: {e!r}")
                continue
            v["speedup_vs_fp64"] = per_newton / v["value"]
            variants[k] = v
            log(f"[bench] {k}: {v['value']:.4f} s/Newton step ({v['newton_steps']} Newton, {v['krylov_its']} "
                f"Krylov), {v['speedup_vs_fp64']:.2f}x the FP64 path")
    peaks = measured_peaks()
    log(f"[bench] measured peaks: copy {peaks['copy_GBps']:.0f} GB/s, read {peaks['read_GBps']:.0f} GB/s, "
        f"FP64 FMA {peaks['fp64_fma_TFLOPs']:.1f} TFLOP/s")
    cpu = None
    if not args.no_cpu and args.workload == 1:
        try:
            cpu = cpu_baseline(n=args.cpu_n)
            log(f"[bench] cpu baseline {cpu}")
        except Exception as e:  # the baseline never blocks the GPU line
            log(f"[bench] cpu baseline failed: {e!r}")
    out = {
        "metric": METRIC, "value": per_newton, "unit": "s/Newton step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64-IR",
        "data": f"synthetic (generated {wl.name} Hex8 mesh)",
        "config": {"workload": f"{wl.name}: {wl.mesh.nelem} Hex8 F-bar elements, {ndof} DOF, "
                               f"{'GCR + learned U-net/block-Jacobi, AMG fallback' if net else 'PCG+SA-AMG'} "
                               f"rtol {wl.krylov_rtol:g}, precision {args.precision}, "
                               f"AMG rebuild policy {args.amg_rebuild} ("
                               + {0: "once per load step", 1: "every Newton iteration", 2: "first two iterations",
                                  3: "fine smoother refreshed", 4: "prolongators kept within a load step, the rest "
                                  "rebuilt every Newton iteration", 5: "prolongators built once, the rest rebuilt "
                                  "every Newton iteration", 6: "prolongators kept until a solve falls back, the rest "
                                  "rebuilt every Newton iteration", 7: "as 6, with the coarse levels of the same Newton iteration "
                                  "of an earlier load step reused except at the second iteration"}.get(args.amg_rebuild, "?")
                               + ")",
                   "amg_cycle": args.cycle.upper(), "coarse_degree": args.coarse_degree,
                   "newton_steps_timed": newton, "newton_steps_accepted": tot.accepted,
                   "krylov_its_timed": krylov,
                   "load_increments_timed": tot.increments, "cutbacks_timed": tot.cutbacks,
                   "cutback_limit": args.cutback, "newton_maxiter": args.maxiter,
                   "t_assemble_s": t_asm, "t_amg_setup_s": t_setup, "t_solve_s": t_solve,
                   "amg_levels": list(s1.amg_ndof[:s1.amg_levels]),
                   **({"fp32_krylov_spmv": spmv32} if spmv32 else {}),
                   **({"vs_fp64_path": compare} if compare else {})},
        "roofline": {"kernel": sp_kernel,
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": SPMV_TRAFFIC_100 if (args.workload == 1 and ndof == 3090903 and not split) else None,
                     "bytes_per_launch": bytes_sp, "avg_launch_ms": t_sp * 1e3},
        "cpu_baseline": cpu,
        "vcycle_roofline": vcyc,
        "solver_variants": variants,
        "measured_peaks": peaks,
        "element_roofline": {"kernel": "gp_kernel + scatter_kernel (Hex8 F-bar integration and tangent "
                                       "contraction, FP64)", "bound": "fp64",
                             "achieved": ELEMENT_FLOPS * wl.mesh.nelem / ((gp_ms + sc_ms) * 1e-3) / 1e12,
                             "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": ELEMENT_FLOPS * wl.mesh.nelem / ((gp_ms + sc_ms) * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                             "flops_per_element": ELEMENT_FLOPS,
                             "flops_per_element_dense": ELEMENT_FLOPS_DENSE,
                             "achieved_dense_count": ELEMENT_FLOPS_DENSE * wl.mesh.nelem / ((gp_ms + sc_ms) * 1e-3) / 1e12,
                             "elements_per_s": wl.mesh.nelem / ((gp_ms + sc_ms) * 1e-3),
                             "gp_kernel_ms": gp_ms, "scatter_ms": sc_ms,
                             "assembly_chunk": args.assembly_chunk},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
