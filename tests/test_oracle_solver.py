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
