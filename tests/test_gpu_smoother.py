"""Chebyshev steps fused with their product (sell_cheb_kernel on the fine level,
bsr_cheb_kernel on the 6x6 coarse levels) against the separate product + step launches
(option smoother_unfused = 1): the same arithmetic in the same order, so the
preconditioned solves agree bit for bit -- for both precisions,
the first- and fourth-kind Chebyshev smoothers, and degree 2 (one step per smoothing:
the fused launch skips the dead r and d stores) as well as degrees 3 and 4 (the new
direction goes to the spare buffer and the two are swapped)."""
import numpy as np
import pytest

import plasticity_amd as pa
from plasticity_amd.workloads import uniaxial_constraints

pytestmark = pytest.mark.gpu
MAT = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9)


def _solve(precision, smoother, degree, unfused, **extra):
    mesh = pa.box_mesh(1, 1, 1, 10, 10, 10, jitter=0.1, seed=2)
    opts = pa.default_options(precision=precision, smoother=smoother, smoother_degree=degree,
                              smoother_unfused=unfused, coarse_max=60, **extra)
    m = pa.Model(mesh, pa.J2Material(**MAT), options=opts)
    d, v, r = uniaxial_constraints(mesh, -0.02)
    m.constrain(d, v, r)
    m._freeze()
    U = (np.array([-0.004, 0.0015, 0.0015]) * mesh.nodes).ravel()
    m.assemble(U, commit=True)
    m.U = 1.5 * U
    m.assemble(1.5 * U)
    m.impose_dirichlet(1.0)
    dU, it, used, relres = m.linear_solve(1e-10)
    m.reset()  # a load step of the compression from the undeformed state
    conv, hist, kry, _ = m.newton(1.0, tol=1e-8, maxiter=25)
    out = (dU.copy(), it, used, relres, conv, hist, kry, m.U.copy())
    m.close()
    return out


@pytest.mark.parametrize("precision,smoother,degree", [
    ("fp64", "chebyshev", 2), ("fp32_ir", "chebyshev", 2), ("fp64", "chebyshev", 3),
    ("fp32_ir", "chebyshev", 4), ("fp64", "chebyshev4", 3), ("fp32_ir", "chebyshev4", 2)])
def test_fused_chebyshev_step_is_bitwise_equal(precision, smoother, degree, need_gpu):
    print(f"[smoother] {precision} {smoother} degree {degree}: fused vs separate launches", flush=True)
    a = _solve(precision, smoother, degree, 0)
    b = _solve(precision, smoother, degree, 1)
    print(f"  krylov {a[1]} / {b[1]}, relres {a[3]:.2e}, newton {len(a[5])} its", flush=True)
    assert a[3] < 1e-9 and a[4]
    assert np.array_equal(a[0], b[0]) and a[1:4] == b[1:4]
    assert a[5] == b[5] and a[6] == b[6] and np.array_equal(a[7], b[7])


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_krylov_overlap_is_bitwise_equal(precision, need_gpu):
    """PCG (FP64) and flexible CG (FP32-IR) queue the next direction and its product
    before the host reads an iteration's scalars (krylov_sync = 0, the default); with
    krylov_sync = 1 the host reads them first.  The kernels and their order on the stream
    are the same, so the solves and the Newton load step agree bit for bit."""
    print(f"[krylov overlap] {precision}: queued vs host-first", flush=True)
    a = _solve(precision, "chebyshev", 2, 0, krylov_sync=0)
    b = _solve(precision, "chebyshev", 2, 0, krylov_sync=1)
    print(f"  krylov {a[1]} / {b[1]}, newton {len(a[5])} its, krylov per newton {a[6]}", flush=True)
    assert a[3] < 1e-9 and a[4]
    assert np.array_equal(a[0], b[0]) and a[1:4] == b[1:4]
    assert a[5] == b[5] and a[6] == b[6] and np.array_equal(a[7], b[7])


def test_split_operator_product(need_gpu):
    """K x from the FP32 + BF16 split copy of K (pfem_apply_K prec 2) against the FP64
    product and the exported K: every entry of the split copy carries 33 significant bits
    (2^-33 = 1.2e-10 relative), so the product is within a few 1e-10 of the FP64 one
    and far closer than the FP32 copy's (2^-24 per entry).  The FP64 product itself
    equals the exported K (CSC) times x to rounding."""
    print("[split operator] 12^3 jittered block, plastic tangent with Dirichlet rows", flush=True)
    mesh = pa.box_mesh(1, 1, 1, 12, 12, 12, jitter=0.1, seed=3)
    m = pa.Model(mesh, pa.J2Material(**MAT), options=pa.default_options(precision="fp32_ir", coarse_max=60))
    d, v, r = uniaxial_constraints(mesh, -0.02)
    m.constrain(d, v, r)
    m._freeze()
    U = (np.array([-0.004, 0.0015, 0.0015]) * mesh.nodes).ravel()
    m.assemble(U, commit=True)
    m.assemble(1.5 * U)
    m.impose_dirichlet(1.0)
    K = m.get_K_csc()
    rng = np.random.default_rng(5)
    for x in (rng.standard_normal(m.ndof), 1.5 * U + 1e-3 * rng.standard_normal(m.ndof)):
        y0, y1, y2 = (m.apply_K(x, prec=p) for p in (0, 1, 2))
        ref = K @ x
        e0 = np.linalg.norm(y0 - ref) / np.linalg.norm(ref)
        e1 = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
        e2 = np.linalg.norm(y2 - y0) / np.linalg.norm(y0)
        print(f"  |y64 - Kx| {e0:.1e}, FP32 copy {e1:.1e}, split copy {e2:.1e}", flush=True)
        assert e0 < 1e-13
        assert e2 < 2e-9 and e2 < 1e-2 * e1
    m.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_vcycle_timing_entry_point(precision, need_gpu):
    """pfem_bench_vcycle times the cycle and the fused fine smoother step on the current
    hierarchy; its byte count of the step is the SELL operator (values + column index per
    block, row pointers), the block-diagonal inverse and six vectors (d, r twice, x twice,
    d'); the cycle holds at least the four fine-level products of a degree-2 V-cycle.  A
    solve after the timing is unchanged (the timing only uses the cycle's own work
    vectors)."""
    print(f"[vcycle timing] {precision}", flush=True)
    mesh = pa.box_mesh(1, 1, 1, 10, 10, 10, jitter=0.1, seed=2)
    m = pa.Model(mesh, pa.J2Material(**MAT), options=pa.default_options(precision=precision, coarse_max=60))
    d, v, r = uniaxial_constraints(mesh, -0.02)
    m.constrain(d, v, r)
    m._freeze()
    U = (np.array([-0.004, 0.0015, 0.0015]) * mesh.nodes).ravel()
    m.assemble(U)
    m.impose_dirichlet(1.0)
    dU0 = m.linear_solve(1e-10, rebuild_pc=True)[0].copy()
    st = m.stats()
    for prec in ([0, 1] if precision == "fp32_ir" else [0]):
        s = 8 if prec == 0 else 4
        f = m.bench_vcycle(prec=prec, reps=5)
        step = st.nnzb * (9 * s + 4) + 4 * (st.nnodes + 1) + 6 * st.ndof * s + 9 * st.nnodes * s
        print(f"  prec {prec}: cycle {f['cycle_s'] * 1e6:.1f} us {f['cycle_bytes']:.0f} B, "
              f"step {f['step_s'] * 1e6:.1f} us {f['step_bytes']:.0f} B", flush=True)
        assert f["cycle_s"] > 0 and f["step_s"] > 0
        assert f["step_bytes"] == step
        assert f["cycle_bytes"] > 4 * (st.nnzb * (9 * s + 4))
    dU1 = m.linear_solve(1e-10, rebuild_pc=False)[0]
    assert np.array_equal(dU0, dU1)
    m.close()
