"""Pins of the oracle's pattern, colouring, assembly, Dirichlet imposition and Newton
driver: brute-force sparse(I,J) pattern (test/test_assembly.jl:26-35), assembled K ==
independent COO sum (B9), six zero modes of the free-free tangent (B8), colour validity
and equality with the parity colours (Assembly.jl:172-188), a dense restatement of
impose_dirichlet!, and the finite-strain uniaxial closed form through solve!."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O


def coo_assembly(nodes, conn, mat, fbar, U):
    ndof = 3 * nodes.shape[0]
    rows, cols, vals = [], [], []
    R = np.zeros(ndof)
    for e in range(conn.shape[0]):
        ed = (3 * conn[e][:, None] + np.arange(3)).ravel()
        Fe, Ke, _ = O.element_finite(fbar, mat, nodes[conn[e]], U[ed])
        rows.append(np.repeat(ed, 24))
        cols.append(np.tile(ed, 24))
        vals.append(Ke.ravel())
        np.add.at(R, ed, Fe)
    K = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(ndof, ndof)).tocsc()
    return K, R


def test_pattern_equals_bruteforce_sparse_IJ():
    nodes, conn = O.box_mesh(1, 1, 1, 3, 2, 4)
    colptr, rowval = O.build_csc(nodes.shape[0], conn)
    ndof = 3 * nodes.shape[0]
    I, J = [], []
    for e in range(conn.shape[0]):
        ed = (3 * conn[e][:, None] + np.arange(3)).ravel()
        I += list(np.repeat(ed, 24))
        J += list(np.tile(ed, 24))
    B = sp.csc_matrix((np.ones(len(I)), (I, J)), shape=(ndof, ndof))
    B.sum_duplicates()
    B.sort_indices()
    assert np.array_equal(B.indptr, colptr) and np.array_equal(B.indices, rowval)
    # nnzb formula on an n^3 box: (3(n+1)-2)^3 blocks (27 neighbours per interior node)
    n = 4
    nodes, conn = O.box_mesh(1, 1, 1, n, n, n)
    colptr, _ = O.build_csc(nodes.shape[0], conn)
    assert colptr[-1] == 9 * (3 * (n + 1) - 2) ** 3


@pytest.mark.parametrize("dims", [(8, 8, 8), (10, 10, 10), (5, 7, 3)])
def test_greedy_colours_equal_parity_colours(dims):
    nx, ny, nz = dims
    nodes, conn = O.box_mesh(1, 1, 1, nx, ny, nz, jitter=0.1, seed=1)
    col, nc = O.greedy_colors(nodes.shape[0], conn)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    parity = ((i & 1) + 2 * (j & 1) + 4 * (k & 1)).ravel()
    assert nc == 8 and np.array_equal(col, parity)
    # validity: no two same-colour elements share a node
    for c in range(nc):
        els = conn[col == c].ravel()
        assert len(np.unique(els)) == len(els)


@pytest.mark.parametrize("fbar", [0, 1])
def test_B9_assembly_equals_independent_coo(fbar):
    nodes, conn = O.box_mesh(1, 1, 1, 3, 3, 3, jitter=0.1, seed=4)
    mat = O.Material(200e9, 0.3, 250e6, Hiso=1e9)
    M = O.OracleModel(nodes, conn, mat, element="finite_fbar" if fbar else "finite")
    rng = np.random.default_rng(2)
    U = 0.004 * rng.normal(size=M.ndof)
    K, R = M.assemble(U, M._new_state(8 * M.ne))
    Kc, Rc = coo_assembly(nodes, conn, mat, fbar, U)
    d = (K - Kc).tocsc()
    assert abs(d).max() <= 1e-12 * abs(Kc).max()
    assert np.abs(R - Rc).max() <= 1e-12 * np.abs(Rc).max()


def test_B8_free_free_tangent_has_six_zero_modes():
    nodes, conn = O.box_mesh(1, 1, 1, 2, 2, 2, jitter=0.1, seed=3)
    mat = O.Material(210e3, 0.3, 250.0, Hiso=1000.0)
    M = O.OracleModel(nodes, conn, mat, element="finite")
    K, _ = M.assemble(np.zeros(M.ndof), M._new_state(8 * M.ne))
    w = np.linalg.eigvalsh(K.toarray())
    assert np.sum(np.abs(w) < 1e-8 * w.max()) == 6
    # the rigid-body modes (Solver.jl:117-129) span that null space
    B = O.rigid_body_modes(nodes)
    assert np.abs(K @ B).max() < 1e-8 * w.max()


def test_impose_dirichlet_matches_dense_elimination():
    nodes, conn = O.box_mesh(1, 1, 1, 2, 2, 2)
    mat = O.Material(210e3, 0.3, 250.0)
    M = O.OracleModel(nodes, conn, mat)
    rng = np.random.default_rng(0)
    U = 1e-3 * rng.normal(size=M.ndof)
    K, R = M.assemble(U, M._new_state(8 * M.ne))
    Kd, Rd = K.toarray(), R.copy()
    dofs = np.array([0, 4, 17, 22, 50], dtype=np.int64)
    vals = np.array([0.01, -0.02, 0.0, 0.03, 0.005])
    lam = 0.4
    O.impose_dirichlet(K, R, dofs, vals, True, lam, U)
    dg = np.zeros(M.ndof)
    dg[dofs] = lam * vals - U[dofs]
    isbc = np.zeros(M.ndof, bool)
    isbc[dofs] = True
    Rd[~isbc] += Kd[np.ix_(~isbc, isbc)] @ dg[isbc]
    Kd[isbc, :] = 0
    Kd[:, isbc] = 0
    Kd[dofs, dofs] = 1.0
    Rd[dofs] = -dg[dofs]
    assert np.allclose(K.toarray(), Kd, rtol=0, atol=1e-12 * abs(Kd).max())
    assert np.allclose(R, Rd, rtol=0, atol=1e-12 * np.abs(Rd).max())
    # the solve then yields dU[d] = dg[d]
    dU = np.linalg.solve(K.toarray(), -R)
    assert np.allclose(dU[dofs], dg[dofs], atol=1e-14)


@pytest.mark.parametrize("element", ["finite", "finite_fbar"])
def test_uniaxial_finite_strain_closed_form(element):
    """Roller cube in uniaxial tension: homogeneous state, so F-bar == standard and the
    Kirchhoff stress follows the log-strain closed form tau = sy0 + Ht (ln(1+e) - sy0/E)."""
    E, sy0, H = 210e3, 250.0, 1000.0
    nodes, conn = O.box_mesh(1, 1, 1, 2, 2, 2)
    M = O.OracleModel(nodes, conn, O.Material(E, 0.3, sy0, Hiso=H), element=element)
    M.fix(O.on_face(nodes, "xmin"), "x")
    M.fix(O.on_face(nodes, "ymin"), "y")
    M.fix(O.on_face(nodes, "zmin"), "z")
    M.prescribe(O.on_face(nodes, "xmax"), "x", 0.05)
    ok, iters, res, _ = M.solve(nsteps=5, tol=1e-10)
    assert ok
    tau = M.st_c["sig"]
    Ht = E * H / (E + H)
    assert np.allclose(tau[:, 0], sy0 + Ht * (np.log(1.05) - sy0 / E), rtol=1e-8)
    assert np.abs(tau[:, 1:]).max() < 1e-8 * tau[0, 0]
    # abar = plastic log strain, uniform
    assert np.allclose(M.st_c["abar"], np.log(1.05) - tau[0, 0] / E, rtol=1e-8)
    # Newton converges quadratically: final ratios super-linear (T15 / F2')
    r = res[-1]
    assert len(r) <= 8
    if len(r) >= 4:
        assert r[-1] < r[-2] ** 1.5 + 1e-12
