"""The device path at the full sizes of the benchmark workloads, where the oracle is
either run at that size (assembly: the oracle's OpenMP C assembly of 1M elements takes
seconds) or replaced by properties that do not depend on the mesh size:

* 100^3 jittered block (3,090,903 DOF): K and R of a plastic state equal the oracle's
  to 1e-10 (max-relative); a homogeneous deformation leaves zero force on interior nodes
  (patch test); the free tangent at U = 0 annihilates the six rigid-body modes; one load
  step of the workload under FP32-IR ends within 1e-8 of the FP64 path (displacements and
  equivalent plastic strain) after the same Newton iterations.
* 200^3 two-phase block (24,361,803 DOF): after three load steps of 1/160 (0.1875%
  tension) the matrix (sigma_y 250 MPa) has yielded and every Gauss point of the
  inclusions (sigma_y 750 MPa, same elastic moduli, stress at most E * 0.19% = 375 MPa)
  is still elastic.
* 150^3 notched bar (10,328,853 DOF): after four load steps of 1/1000 the equivalent
  plastic strain is largest at the notch (mid-length), where the section is smallest.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import plasticity_amd as pa
from plasticity_amd.workloads import workload

pytestmark = pytest.mark.gpu


def _rigid_modes(X):
    n = X.shape[0]
    B = np.zeros((3 * n, 6))
    c = X - X.mean(axis=0)
    for d in range(3):
        B[d::3, d] = 1.0
    # infinitesimal rotations about x, y, z: w x X
    B[1::3, 3], B[2::3, 3] = -c[:, 2], c[:, 1]
    B[0::3, 4], B[2::3, 4] = c[:, 2], -c[:, 0]
    B[0::3, 5], B[1::3, 5] = -c[:, 1], c[:, 0]
    return B


def test_full_size_assembly_matches_oracle_100(need_gpu):
    wl = workload(1, device=True)
    mesh = wl.mesh
    print(f"[full size] 100^3: {mesh.nelem} elements, {3 * mesh.nnodes} DOF", flush=True)
    m = pa.Model(mesh, pa.J2Material(**wl.materials[0]))
    rng = np.random.default_rng(11)
    # 1% axial compression with lateral expansion and a small random field: plastic at
    # every Gauss point from the virgin state (yield strain 0.125%)
    U = (np.array([-0.01, 0.004, 0.004]) * mesh.nodes).ravel() + 2e-5 * rng.standard_normal(3 * mesh.nnodes)
    Rg = m.assemble(U)
    Kg = m.get_K_csc()
    tau_g = m.trial_stress()
    print("  device assembly done; oracle assembly (OpenMP)...", flush=True)
    om = O.OracleModel(mesh.nodes, mesh.elements, O.Material(**wl.materials[0]), element=wl.element)
    st = om._new_state(8 * om.ne)
    Ko, Ro = om.assemble(U, st)
    assert np.array_equal(Kg.indptr, Ko.indptr) and np.array_equal(Kg.indices, Ko.indices)
    dk = np.abs(Kg.data - Ko.data).max() / np.abs(Ko.data).max()
    dr = np.abs(Rg - Ro).max() / np.abs(Ro).max()
    dt = np.abs(tau_g - st["sig"]).max() / np.abs(st["sig"]).max()
    print(f"  nnz {Kg.nnz}: K {dk:.2e}  R {dr:.2e}  tau {dt:.2e}", flush=True)
    assert dk < 1e-10 and dr < 1e-10 and dt < 1e-10
    del Ko, Ro, om, st
    # patch test: a homogeneous deformation gradient (no noise) loads only boundary nodes
    F = np.array([[0.99, 0.002, 0.0], [0.001, 1.004, 0.0], [0.0, 0.003, 1.004]])
    Uh = (mesh.nodes @ F.T - mesh.nodes).ravel()
    Rh = m.assemble(Uh).reshape(-1, 3)
    X = mesh.nodes
    tol = 1e-8
    interior = np.all((X > tol) & (X < 1.0 - tol), axis=1)
    ri, rb = np.abs(Rh[interior]).max(), np.abs(Rh[~interior]).max()
    print(f"  patch test: max |R| interior {ri:.2e}, boundary {rb:.2e}", flush=True)
    assert ri <= 1e-10 * rb
    # rigid-body modes of the free tangent at U = 0
    m.assemble(np.zeros(3 * mesh.nnodes))
    K0 = m.get_K_csc()
    B = _rigid_modes(X)
    KB = K0 @ B
    knorm = sp.linalg.norm(K0, ord=np.inf)
    rel = [np.abs(KB[:, j]).max() / (knorm * np.abs(B[:, j]).max()) for j in range(6)]
    print(f"  |K b| / (|K| |b|) for the six rigid modes: {['%.1e' % r for r in rel]}", flush=True)
    assert max(rel) < 1e-12
    m.close()


def test_full_size_fp32_ir_load_step_matches_fp64_100(need_gpu):
    wl = workload(1, device=True)
    out = {}
    for prec in ("fp64", "fp32_ir"):
        o = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, precision=prec)
        m = pa.Model(wl.mesh, pa.J2Material(**wl.materials[0]), options=o)
        m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
        m.reset()
        m._freeze()
        conv, hist, kry, res = m.newton(1 / wl.nsteps, tol=wl.newton_tol)
        assert conv
        m.commit()
        out[prec] = (m.U, m.get_state()["abar"], hist)
        print(f"[full size] 100^3 load step 1 {prec}: |R| {['%.2e' % h for h in hist]} krylov {kry}", flush=True)
        m.close()
    (U64, a64, h64), (U32, a32, h32) = out["fp64"], out["fp32_ir"]
    eu = np.linalg.norm(U32 - U64) / np.linalg.norm(U64)
    ea = np.linalg.norm(a32 - a64) / max(np.linalg.norm(a64), 1e-300)
    print(f"  FP32-IR vs FP64: U {eu:.2e}, abar {ea:.2e}", flush=True)
    assert a64.max() > 0
    # both runs solve every Newton system to the Krylov rtol 1e-8 and stop at the Newton
    # tolerance 1e-8: the displacement bar of the mixed-precision path is 1e-8
    assert len(h64) == len(h32)
    assert eu < 1e-8 and ea < 1e-8


def test_full_size_policy7_fp32_ir_matches_fp64_100(need_gpu):
    """Workload 1 at full size, load steps 1-4, under the benchmark's solver settings (AMG
    policy 7: load steps 2-4 restore the coarse levels kept at load step 1; FP32-IR with
    the split operator) against the FP64 path (policy 7 too), both at Newton tolerance
    1e-11 (at 1e-10 the two runs stop at different iterations and differ by 1.4e-8):
    displacements and equivalent plastic strain within 1e-8, and load steps 2-4 restore
    kept coarse levels instead of setting up."""
    wl = workload(1, device=True)
    out = {}
    for prec in ("fp64", "fp32_ir"):
        o = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, precision=prec,
                               amg_rebuild="reuse_coarse")
        m = pa.Model(wl.mesh, pa.J2Material(**wl.materials[0]), options=o)
        m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
        res = m.solve(nsteps=wl.nsteps, tol=1e-11, steps=range(1, 5))
        assert res.converged
        st = m.stats()
        out[prec] = (m.U, m.get_state()["abar"])
        print(f"[full size] 100^3 load steps 1-4 {prec}, policy 7: newton {res.iters} krylov "
              f"{[sum(k) for k in res.cg_iters]}, set-ups {st.n_setup} for {st.n_solve} solves", flush=True)
        assert st.n_setup <= st.n_solve - 4  # load steps 2-4 restore at their first iteration at least
        m.close()
    (U64, a64), (U32, a32) = out["fp64"], out["fp32_ir"]
    eu = np.linalg.norm(U32 - U64) / np.linalg.norm(U64)
    ea = np.linalg.norm(a32 - a64) / max(np.linalg.norm(a64), 1e-300)
    print(f"  FP32-IR vs FP64: U {eu:.2e}, abar {ea:.2e}", flush=True)
    assert a64.max() > 0 and eu < 1e-8 and ea < 1e-8


def test_full_size_two_phase_inclusions_stay_elastic_200(need_gpu):
    wl = workload(2, nsteps=160, device=True)
    print(f"[full size] 200^3 two-phase: {wl.mesh.nelem} elements, {3 * wl.mesh.nnodes} DOF, "
          f"inclusion fraction {wl.material_id.mean():.3f}", flush=True)
    o = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, precision="fp32_ir")
    m = pa.Model(wl.mesh, [pa.J2Material(**d) for d in wl.materials], options=o, material_id=wl.material_id)
    m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    res = m.solve(nsteps=wl.nsteps, tol=wl.newton_tol, steps=range(1, 4))
    print(f"  newton {res.iters} krylov {res.cg_iters}", flush=True)
    assert res.converged
    abar = m.get_state()["abar"].reshape(-1, 8)
    inc = wl.material_id == 1
    print(f"  abar: matrix mean {abar[~inc].mean():.2e} (yielded GPs {(abar[~inc] > 0).mean():.3f}), "
          f"inclusions max {abar[inc].max():.2e}", flush=True)
    assert (abar[~inc] > 0).mean() > 0.5
    assert abar[inc].max() == 0.0
    m.close()


def test_full_size_notched_bar_yields_at_the_notch_150(need_gpu):
    wl = workload(4, nsteps=1000, device=False)
    print(f"[full size] 150^3 notched bar: {wl.mesh.nelem} elements, {3 * wl.mesh.nnodes} DOF", flush=True)
    o = pa.default_options(eta_min=wl.krylov_rtol, eta_max=wl.krylov_rtol, precision="fp32_ir")
    m = pa.Model(wl.mesh, pa.J2Material(**wl.materials[0]), options=o)
    m.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    res = m.solve(nsteps=wl.nsteps, tol=wl.newton_tol, steps=range(1, 5))
    print(f"  newton {res.iters} krylov {res.cg_iters}", flush=True)
    assert res.converged
    abar = m.get_state()["abar"].reshape(-1, 8).max(axis=1)
    xc = wl.mesh.nodes[wl.mesh.elements, 0].mean(axis=1)
    e = int(np.argmax(abar))
    near = np.abs(xc - 0.5) < 0.1
    print(f"  max abar {abar[e]:.2e} at x = {xc[e]:.3f}; yielded elements near the notch "
          f"{(abar[near] > 0).mean():.3f}, far from it {(abar[np.abs(xc - 0.5) > 0.3] > 0).mean():.3f}", flush=True)
    assert abar[e] > 0 and abs(xc[e] - 0.5) < 0.1
    m.close()
