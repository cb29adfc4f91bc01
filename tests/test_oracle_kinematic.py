"""Pins of the oracle's objective kinematic hardening at finite strain and of its
small-strain Hex8 element.

Reference tests restated: F2 (tangent vs finite differences with Hkin = 500),
F4b (objectivity at 120 degrees with a stored back stress), F9b (consistent but
non-symmetric tangent with Hkin = 2000) in test/test_finite_strain.jl; T11/T12 and
the patch/rigid-mode checks of test/test_element.jl for the small-strain element.
Independent pins: the polar decomposition and its derivative against their defining
identities and finite differences, dtau/dbeta against finite differences of the
return map, and the coaxial reduction (diagonal F: R = I, so the finite-strain update
along a diagonal path is the small-strain return map on the Hencky strain)."""
import numpy as np
import pytest

import oracle as O

I6 = np.array([1.0, 1, 1, 0, 0, 0])
Z6 = np.zeros(6)
KIN500 = dict(E=210e3, nu=0.3, sy0=250.0, Hiso=1000.0, Hkin=500.0)
KIN2000 = dict(E=210e3, nu=0.3, sy0=250.0, Hiso=200.0, Hkin=2000.0)


def unit_element():
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    return nodes[conn[0]]


def ue_from_F(X, F):
    return (X @ np.asarray(F).T - X).ravel()


def fresh():
    return np.zeros((8, 6)), np.zeros((8, 6)), np.zeros(8), np.tile(I6, (8, 1))


def elem(fbar, mat, X, ue, st, commit=False):
    ep, bt, ab, cp = st
    return O.element_finite(fbar, mat, X, ue, ep, ab, cp, commit=commit, beta=bt)


def copy(st):
    return tuple(a.copy() for a in st)


def fd_tangent(fbar, mat, X, ue, st, h=1e-7):
    _, Ke, _ = elem(fbar, mat, X, ue, copy(st))
    Kfd = np.zeros((24, 24))
    for j in range(24):
        up, um = ue.copy(), ue.copy()
        up[j] += h
        um[j] -= h
        Kfd[:, j] = (elem(fbar, mat, X, up, copy(st))[0] - elem(fbar, mat, X, um, copy(st))[0]) / (2 * h)
    return np.linalg.norm(Kfd - Ke) / np.linalg.norm(Ke), np.linalg.norm(Ke - Ke.T) / np.linalg.norm(Ke)


def rand_F(rng, amp=0.15):
    return np.eye(3) + amp * rng.standard_normal((3, 3))


def test_polar_decomposition_identities():
    rng = np.random.default_rng(1)
    for _ in range(20):
        F = rand_F(rng, 0.3)
        if np.linalg.det(F) <= 0.1:
            continue
        R, U = O.polar_RU(F)
        assert np.allclose(R @ U, F, atol=1e-13)
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-13)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-13)
        assert np.allclose(U, U.T, atol=1e-14) and np.linalg.eigvalsh(U).min() > 0
    # F = Q diag(s): R = Q, U = diag(s)
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    R, U = O.polar_RU(Q @ np.diag([1.2, 0.9, 1.05]))
    assert np.allclose(R, Q, atol=1e-14) and np.allclose(U, np.diag([1.2, 0.9, 1.05]), atol=1e-14)


def test_dR_polar_vs_finite_differences():
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(5):
        F = rand_F(rng)
        R, U = O.polar_RU(F)
        for p in range(3):
            for q in range(3):
                dF = np.zeros((3, 3))
                dF[p, q] = 1.0
                fd = (O.polar_RU(F + h * dF)[0] - O.polar_RU(F - h * dF)[0]) / (2 * h)
                an = O.dR_polar(R, U, dF)
                assert np.linalg.norm(an - fd) <= 1e-8 * max(1.0, np.linalg.norm(fd))
                # R^T dR is skew
                assert np.allclose(R.T @ an + an.T @ R, 0, atol=1e-12)


def test_dtau_dbeta_vs_finite_differences_of_the_return_map():
    mat = O.Material(**KIN2000)
    ee = np.array([0.004, -0.002, -0.001, 0.003, -0.001, 0.002])
    bsp = np.array([20.0, -12.0, -8.0, 5.0, 3.0, -4.0])
    an = 0.002
    T = O.dtau_dbeta(mat, ee, bsp, an)
    assert np.abs(T).max() > 0
    h = 1e-4
    for j in range(6):
        bp, bm = bsp.copy(), bsp.copy()
        bp[j] += h
        bm[j] -= h
        fd = (O.return_map(mat, ee, Z6, bp, an)[0] - O.return_map(mat, ee, Z6, bm, an)[0]) / (2 * h)
        assert np.allclose(T[:, j], fd, rtol=1e-7, atol=1e-9)
    # elastic branch: zero
    assert np.all(O.dtau_dbeta(mat, 1e-6 * ee, Z6, 0.0) == 0)


def test_hkin_zero_equals_isotropic_path_bitwise():
    """Hkin = 0 with a (necessarily zero) back stress is the isotropic update exactly."""
    mat = O.Material(E=210e3, nu=0.3, sy0=250.0, Hiso=1000.0)
    X = unit_element()
    F = np.array([[1.05, 0.04, 0.01], [0.025, 0.98, 0.015], [0.005, 0.01, 1.02]])
    st = fresh()
    f1, k1, s1 = elem(1, mat, X, ue_from_F(X, F), st, commit=True)
    assert np.all(st[1] == 0)
    f2, k2, s2 = O.element_finite(1, mat, X, ue_from_F(X, F), np.zeros((8, 6)), np.zeros(8),
                                  np.tile(I6, (8, 1)))
    assert np.array_equal(f1, f2) and np.array_equal(k1, k2) and np.array_equal(s1, s2)


@pytest.mark.parametrize("fbar", [0, 1])
def test_F2_kinematic_tangent_vs_fd(fbar):
    """F2 (a)-(d) with Hkin = 500: elastic, incremental plastic, non-coaxial plastic."""
    mat = O.Material(**KIN500)
    X = unit_element()
    for F in (np.diag([1.1, 0.95, 1.03]),
              np.array([[1.08, 0.03, 0.01], [0.04, 0.97, 0.0], [0.0, 0.02, 1.05]])):
        assert fd_tangent(fbar, mat, X, ue_from_F(X, F), fresh())[0] < 1e-6
    st = fresh()
    for k in range(1, 9):
        Fk = np.diag([1 + 0.02 * k, 1 - 0.008 * k, 1 - 0.008 * k])
        assert fd_tangent(fbar, mat, X, ue_from_F(X, Fk), st)[0] < 1e-6
        elem(fbar, mat, X, ue_from_F(X, Fk), st, commit=True)
    assert st[2].max() > 0.05
    st = fresh()
    F1 = np.array([[1.05, 0.025, 0.005], [0.04, 0.98, 0.01], [0.01, 0.015, 1.02]])
    F2 = np.array([[1.1, 0.05, 0.01], [0.08, 0.96, 0.02], [0.02, 0.03, 1.04]])
    elem(fbar, mat, X, ue_from_F(X, F1), st, commit=True)
    assert np.abs(st[1]).max() > 0
    assert fd_tangent(fbar, mat, X, ue_from_F(X, F2), st)[0] < 1e-6


def test_F9b_kinematic_tangent_consistent_and_nonsymmetric():
    mat = O.Material(**KIN2000)
    X = unit_element()
    st = fresh()
    F1 = np.array([[1.1, 0.03, 0.0], [0.04, 0.95, 0.01], [0.01, 0.02, 0.97]])
    F2 = np.array([[1.18, 0.06, 0.02], [0.09, 0.92, 0.03], [0.03, 0.04, 1.0]])
    elem(0, mat, X, ue_from_F(X, F1), st, commit=True)
    assert st[2].max() > 0
    rel, sym = fd_tangent(0, mat, X, ue_from_F(X, F2), st)
    assert rel < 1e-6
    assert sym > 1e-6
    # without the back-stress rotation term the tangent is NOT consistent (the term matters)
    k = O.finite_kinematics(F2, st[3][0])
    tau, _, _, D, _, _, bsp, R, U = O.finite_stress_update(mat, k, st[0][0], st[2][0], F=F2, bref=st[1][0],
                                                           full=True)
    A_full = O.dPdF(k, st[3][0], D, tau, F2, st[1][0], R, U, O.dtau_dbeta(mat, np.array(k.ee[:]), bsp,
                                                                         st[2][0]))
    A_iso = O.dPdF(k, st[3][0], D, tau, F2)
    assert np.linalg.norm(A_full - A_iso) > 1e-4 * np.linalg.norm(A_full)


def test_F4b_objectivity_with_back_stress():
    mat = O.Material(**KIN2000)
    F1 = np.array([[1.2, 0.03, 0.0], [0.05, 0.92, 0.0], [0.0, 0.0, 0.95]])
    k1 = O.finite_kinematics(F1, I6)
    _, ep1, a1, _, cpi1, bref1, _, _, _ = O.finite_stress_update(mat, k1, Z6, 0.0, F=F1, full=True)
    assert a1 > 0 and np.linalg.norm(bref1) > 0
    F2 = np.array([[1.3, 0.06, 0.0], [0.08, 0.9, 0.02], [0.02, 0.01, 0.97]])

    def sigma(F):
        k = O.finite_kinematics(F, cpi1)
        tau = O.finite_stress_update(mat, k, ep1, a1, F=F, bref=bref1)[0]
        T = np.array([[tau[0], tau[3], tau[5]], [tau[3], tau[1], tau[4]], [tau[5], tau[4], tau[2]]])
        return T / np.linalg.det(F)

    s1 = sigma(F2)
    for th in (0.1, np.deg2rad(120.0), 3.0):
        Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        assert np.linalg.norm(sigma(Q @ F2) - Q @ s1 @ Q.T) / np.linalg.norm(s1) < 1e-10


def test_coaxial_path_reduces_to_small_strain_return_map():
    """Diagonal F: R = I, Cp^-1 diagonal, so each finite-strain update is the small-strain
    return map on eps = ln(diag F) with the accumulated eps_p, beta, abar -- through a
    tension-compression cycle with reverse yielding (Bauschinger)."""
    mat = O.Material(E=70e3, nu=0.33, sy0=120.0, Hiso=0.0, Hkin=900.0)
    path = [1.004, 1.01, 1.02, 1.005, 0.99, 0.98]
    cpi, ep_f, br, af = I6.copy(), Z6.copy(), Z6.copy(), 0.0
    ep_s, bs, as_ = Z6.copy(), Z6.copy(), 0.0
    yielded = []
    for s in path:
        F = np.diag([s, 1.0, 1.0])
        k = O.finite_kinematics(F, cpi)
        tau, ep_f, a_new, D, cpi, br_new, bsp, R, U = O.finite_stress_update(mat, k, ep_f, af, F=F, bref=br,
                                                                              full=True)
        assert np.allclose(R, np.eye(3), atol=1e-15)
        eps = np.array([np.log(s), 0, 0, 0, 0, 0])
        sig, ep_s, bs, a_s_new, Ds = O.return_map(mat, eps, ep_s, bs, as_)
        assert np.allclose(tau, sig, rtol=1e-10, atol=1e-9)
        assert np.allclose(br_new, bs, rtol=1e-10, atol=1e-10)
        assert np.allclose(D, Ds, rtol=1e-10, atol=1e-6)
        yielded.append(a_new > af)
        br, af, as_ = br_new, a_new, a_s_new
    assert yielded[1] and yielded[2] and yielded[-1]  # forward and reverse plastic flow


# ------------------------------------------------------------------ small strain
def test_small_element_tangent_vs_fd_and_symmetric():
    mat = O.Material(**KIN500)
    X = unit_element() + 0.05 * np.sin(np.arange(24)).reshape(8, 3)
    rng = np.random.default_rng(3)
    ep, bt, ab = np.zeros((8, 6)), np.zeros((8, 6)), np.zeros(8)
    u1 = 4e-3 * rng.standard_normal(24)
    O.element_small(mat, X, u1, ep, bt, ab, commit=True)
    assert ab.max() > 0
    u2 = 1.5 * u1
    _, Ke, _ = O.element_small(mat, X, u2, ep.copy(), bt.copy(), ab.copy())
    h = 1e-8
    Kfd = np.zeros((24, 24))
    for j in range(24):
        up, um = u2.copy(), u2.copy()
        up[j] += h
        um[j] -= h
        Kfd[:, j] = (O.element_small(mat, X, up, ep.copy(), bt.copy(), ab.copy())[0] -
                     O.element_small(mat, X, um, ep.copy(), bt.copy(), ab.copy())[0]) / (2 * h)
    assert np.linalg.norm(Kfd - Ke) / np.linalg.norm(Ke) < 1e-6
    assert np.linalg.norm(Ke - Ke.T) / np.linalg.norm(Ke) < 1e-12


def test_small_element_rigid_modes_and_patch():
    """Six rigid-body modes in the kernel of the elastic Ke; a homogeneous strain gives the
    closed-form stress C eps at every Gauss point (patch)."""
    mat = O.Material(E=200e3, nu=0.3, sy0=1e9)
    X = unit_element() + 0.04 * np.cos(np.arange(24)).reshape(8, 3)
    _, Ke, _ = O.element_small(mat, X, np.zeros(24))
    w = np.linalg.eigvalsh(0.5 * (Ke + Ke.T))
    assert np.sum(np.abs(w) < 1e-9 * w.max()) == 6
    Hg = np.array([[1e-3, 2e-4, 0.0], [-1e-4, -5e-4, 3e-4], [2e-4, 0.0, 4e-4]])
    ue = (X @ Hg.T).ravel()
    _, _, sig = O.element_small(mat, X, ue)
    eps = np.array([Hg[0, 0], Hg[1, 1], Hg[2, 2], Hg[0, 1] + Hg[1, 0], Hg[1, 2] + Hg[2, 1], Hg[0, 2] + Hg[2, 0]])
    ref = mat.Cmat @ eps
    assert np.allclose(sig, np.tile(ref, (8, 1)), rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_small_strain_limit_of_finite_element():
    """F1: at a tiny load the finite-strain element's force approaches the small-strain one."""
    mat = O.Material(**KIN500)
    X = unit_element()
    rng = np.random.default_rng(4)
    ue = 1e-7 * rng.standard_normal(24)
    fs = O.element_small(mat, X, ue)[0]
    ff = O.element_finite(0, mat, X, ue)[0]
    assert np.linalg.norm(ff - fs) / np.linalg.norm(fs) < 1e-5


def test_oracle_model_small_strain_uniaxial_closed_form():
    """Uniaxial stress bar on a small-strain box (T2 analogue): sigma = sy0 + E H/(E+H)(eps - sy0/E)."""
    E, sy0, H = 200e3, 250.0, 1000.0
    nodes, conn = O.box_mesh(1, 1, 1, 2, 2, 2)
    M = O.OracleModel(nodes, conn, O.Material(E, 0.3, sy0, Hiso=H), element="small")
    M.fix(np.where(np.abs(nodes[:, 0]) < 1e-12)[0], "x")
    M.fix(np.where(np.abs(nodes[:, 1]) < 1e-12)[0], "y")
    M.fix(np.where(np.abs(nodes[:, 2]) < 1e-12)[0], "z")
    eps = 0.004
    M.prescribe(np.where(np.abs(nodes[:, 0] - 1) < 1e-12)[0], "x", eps)
    conv, iters, *_ = M.solve(nsteps=4, tol=1e-10)
    assert conv
    sxx = M.st_c["sig"][:, 0]
    ref = sy0 + E * H / (E + H) * (eps - sy0 / E)
    assert np.allclose(sxx, ref, rtol=1e-8)
