"""Device assembly (colour-ordered BSR scatter), Dirichlet imposition and residual norm
against the oracle: pattern and colouring bit-exact, K and R within 1e-10 relative."""
import numpy as np
import pytest

import oracle as O
import plasticity_amd as pa
from plasticity_amd.workloads import uniaxial_constraints

pytestmark = pytest.mark.gpu
MAT = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9)


def build(n=5, jit=0.1, seed=4, element="finite_fbar"):
    mesh = pa.box_mesh(1, 1, 1, n, n, n, jitter=jit, seed=seed)
    gm = pa.Model(mesh, pa.J2Material(**MAT), element=element)
    om = O.OracleModel(mesh.nodes, mesh.elements, O.Material(**MAT), element=element)
    return mesh, gm, om


def plastic_history(om, U, seed):
    st = om._new_state(8 * om.ne)
    om.assemble(U, st, commit=True)
    return st


def test_pattern_and_colours_bit_exact(need_gpu):
    print("[pattern] 5^3 jittered", flush=True)
    mesh, gm, om = build()
    K = gm.get_K_csc()
    assert np.array_equal(K.indptr, om.colptr) and np.array_equal(K.indices, om.rowval)
    col, nc = gm.colors()
    assert nc == om.ncolors and np.array_equal(col, om.colors)


@pytest.mark.parametrize("element", ["finite", "finite_fbar"])
def test_assembly_matches_oracle(element, need_gpu):
    print(f"[assembly] {element} 5^3 jittered", flush=True)
    mesh, gm, om = build(element=element)
    rng = np.random.default_rng(1)
    U1 = (np.array([-0.01, 0.004, 0.004]) * mesh.nodes).ravel() + 1e-4 * rng.normal(size=gm.ndof)
    st = plastic_history(om, U1, 1)
    assert st["abar"].max() > 0
    gm.set_state(st["epsp"], st["abar"], st["cpinv"])
    U2 = U1 + (np.array([-0.005, 0.002, 0.002]) * mesh.nodes).ravel() + 1e-4 * rng.normal(size=gm.ndof)
    Rg = gm.assemble(U2)
    st2 = {k: v.copy() for k, v in st.items()}
    Ko, Ro = om.assemble(U2, st2)
    Kg = gm.get_K_csc()
    dk = np.abs(Kg.data - Ko.data).max() / np.abs(Ko.data).max()
    dr = np.abs(Rg - Ro).max() / np.abs(Ro).max()
    print(f"  K {dk:.2e}  R {dr:.2e}", flush=True)
    assert dk < 1e-10 and dr < 1e-10
    tau = gm.trial_stress()
    assert np.abs(tau - st2["sig"]).max() / np.abs(st2["sig"]).max() < 1e-10


def test_impose_dirichlet_and_norm_match_oracle(need_gpu):
    mesh, gm, om = build(n=4)
    d, v, r = uniaxial_constraints(mesh, -0.2)
    gm.constrain(d, v, r)
    gm._freeze()
    om.constrain(d, v, r)
    (rd, rv), (fd, fv) = om.freeze_bcs()
    rng = np.random.default_rng(2)
    U = 1e-3 * rng.normal(size=gm.ndof)
    lam = 0.3
    gm.assemble(U)
    gm.subtract_external(lam)
    gm.impose_dirichlet(lam)
    Ko, Ro = om.assemble(U, om._new_state(8 * om.ne))
    O.impose_dirichlet(Ko, Ro, rd, rv, True, lam, U)
    O.impose_dirichlet(Ko, Ro, fd, fv, False, lam, U)
    Kg, Rg = gm.get_K_csc(), gm.residual()
    assert np.abs(Kg.data - Ko.data).max() <= 1e-10 * np.abs(Ko.data).max()
    assert np.abs(Rg - Ro).max() <= 1e-10 * np.abs(Ro).max()
    assert gm.residual_norm() == pytest.approx(np.linalg.norm(Ro), rel=1e-12)


def test_inversion_reported_by_assembly(need_gpu):
    mesh, gm, om = build(n=2, jit=0.0)
    U = np.zeros(gm.ndof)
    U[0::3] = -2.5 * mesh.nodes[:, 0]  # x -> -1.5 x: every element inverted
    with pytest.raises(pa.ElementInversionError) as ei:
        gm.assemble(U)
    assert ei.value.J < 0 and ei.value.e == 0 and ei.value.g == 0
