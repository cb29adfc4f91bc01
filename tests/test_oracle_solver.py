"""The oracle's PCG + SA-AMG (the reference's CG+AMG linear solve, Solver.jl:235-290)
against direct solves: C1 of test/test_scaling.jl (CG+AMG == direct within 10 rtol),
with an iteration bound in the spirit of C2 (mesh-independent counts)."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from plasticity_amd.workloads import workload


def imposed_system(n, element="finite"):
    wl = workload(1, n=n)
    M = O.OracleModel(wl.mesh.nodes, wl.mesh.elements, O.Material(**wl.materials[0]), element=element)
    M.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    (rd, rv), (fd, fv) = M.freeze_bcs()
    isbc = np.zeros(M.ndof, np.uint8)
    isbc[rd] = 1
    isbc[fd] = 1
    U = np.zeros(M.ndof)
    K, R = M.assemble(U, M._new_state(8 * M.ne))
    O.impose_dirichlet(K, R, rd, rv, True, 1 / 40, U)
    O.impose_dirichlet(K, R, fd, fv, False, 1 / 40, U)
    return M, K, -R, isbc


@pytest.mark.parametrize("n", [8, 12])
def test_C1_pcg_amg_matches_direct(n):
    M, K, b, isbc = imposed_system(n)
    amg = O.OracleAMG(K, M.nodes, isbc, coarse_max=500)
    assert len(amg.levels()) >= 2
    rtol = 1e-8
    x, it = amg.solve(b, rtol)
    xd = spla.spsolve(K.tocsc(), b)
    assert it > 0
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 10 * rtol
    assert it <= 40


def test_vcycle_is_symmetric_for_symmetric_K():
    M, K, b, isbc = imposed_system(6)
    amg = O.OracleAMG(K, M.nodes, isbc, coarse_max=100)
    rng = np.random.default_rng(0)
    u, v = rng.normal(size=M.ndof), rng.normal(size=M.ndof)
    assert abs(u @ amg.vcycle(v) - v @ amg.vcycle(u)) <= 1e-10 * abs(u @ amg.vcycle(u))


def test_forcing_term_known_values():
    """Eisenstat-Walker choice 2 (Solver.jl:159-171), values worked by hand:
    eta = 0.9 (r_k / r_{k-1})^1.5, raised to the safeguard 0.9 eta_prev^1.5 when that
    exceeds 0.1, clamped to [eta_min, eta_max]."""
    assert O.forcing_term(1, 5.0, 0.0, 0.1) == 0.1                     # first iteration: eta_max
    assert O.forcing_term(3, 5.0, 0.0, 0.1) == 0.1                     # no previous residual
    # ratio 0.01: 0.9 * 1e-3 = 9e-4; safeguard 0.9 * 0.1^1.5 = 0.02846 is below 0.1
    assert O.forcing_term(2, 1e-2, 1.0, 0.1) == pytest.approx(9e-4, rel=1e-14)
    # ratio 0.25: 0.9 * 0.125 = 0.1125 -> clamped to eta_max = 0.1
    assert O.forcing_term(2, 0.25, 1.0, 0.1) == 0.1
    # safeguard active: eta_prev 0.64 -> 0.9 * 0.512 = 0.4608 > 0.1 beats 0.9 * 0.001
    assert O.forcing_term(2, 1e-2, 1.0, 0.64, eta_max=1.0) == pytest.approx(0.4608, rel=1e-14)
    # quadratic phase: ratio 1e-8 -> 0.9e-12 -> clamped to eta_min
    assert O.forcing_term(4, 1e-8, 1.0, 1e-3) == 1e-8
    # gamma and alpha are honoured: 0.5 * 0.01^2 = 5e-5
    assert O.forcing_term(2, 1e-2, 1.0, 1e-3, gamma=0.5, alpha=2.0) == pytest.approx(5e-5, rel=1e-14)
