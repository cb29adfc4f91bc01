"""The oracle reproduces its committed golden fixtures (guards the checker itself)."""
import os

import numpy as np

import oracle as O
from plasticity_amd.workloads import workload

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_reproduces_element_fixture():
    g = np.load(os.path.join(HERE, "golden", "elements.npz"))
    mat = O.Material(200e9, 0.3, 250e6, Hiso=1e9)
    for e in range(g["X"].shape[0]):
        Fe, Ke, s = O.element_finite(1, mat, g["X"][e], g["ue"][e], g["epsp"][e].copy(), g["abar"][e].copy(),
                                     g["cpinv"][e].copy())
        assert np.array_equal(Fe, g["Fe"][e]) and np.array_equal(Ke, g["Ke"][e])


def test_oracle_reproduces_config0_first_steps():
    """Step 1 of the 20-step run (lambda = 1/20) equals a 1-step run to 1/20 of the
    prescribed displacement, and both equal the committed fixture."""
    g = np.load(os.path.join(HERE, "golden", "config0.npz"))
    wl = workload(0, nsteps=20)
    M = O.OracleModel(wl.mesh.nodes, wl.mesh.elements, O.Material(**wl.materials[0]))
    M.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    (rd, rv), (fd, fv) = M.freeze_bcs()
    M.U[:] = 0.0
    M.st_c = M._new_state(8 * M.ne)
    conv, _, _ = M.newton(1 / 20, np.zeros(M.ndof), ((rd, rv), (fd, fv)), tol=wl.newton_tol)
    assert conv
    U_step1 = M.U.copy()
    M2 = O.OracleModel(wl.mesh.nodes, wl.mesh.elements, O.Material(**wl.materials[0]))
    M2.constrain(wl.dir_dofs, wl.dir_vals / 20, wl.dir_ramp)
    M2.solve(nsteps=1, tol=wl.newton_tol)
    ref = np.linalg.norm(g["U_1"])
    assert np.linalg.norm(M2.U - g["U_1"]) <= 1e-12 * ref
    assert np.linalg.norm(U_step1 - g["U_1"]) <= 1e-12 * ref
