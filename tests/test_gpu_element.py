"""GPU element kernels (gp_kernel + tangent contraction, through pfem_element_batch)
against the oracle: Fe, Ke, Kirchhoff stress and committed history within 1e-10
relative; FD gate on the GPU tangent; inversion error; degenerate-eigenvalue branch."""
import os

import numpy as np
import pytest

import oracle as O
import plasticity_amd as pa

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MAT = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def random_states(n, seed, fbar, plastic=True, distort=0.06):
    rng = np.random.default_rng(seed)
    mat = O.Material(**MAT)
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X0 = nodes[conn[0]]
    Xs, ues, eps, abs_, cps = [], [], [], [], []
    for _ in range(n):
        X = X0 * (0.5 + rng.random(3)) + distort * rng.normal(size=(8, 3))
        F1 = np.eye(3) + 0.03 * rng.normal(size=(3, 3))
        F2 = F1 + 0.02 * rng.normal(size=(3, 3))
        e, a, c = np.zeros((8, 6)), np.zeros(8), np.tile([1.0, 1, 1, 0, 0, 0], (8, 1))
        if plastic:
            O.element_finite(fbar, mat, X.ravel(), (X @ F1.T - X).ravel(), e, a, c, commit=True)
        Xs.append(X.ravel())
        ues.append((X @ F2.T - X).ravel() + 1e-3 * rng.normal(size=24))
        eps.append(e); abs_.append(a); cps.append(c)
    return np.array(Xs), np.array(ues), np.array(eps), np.array(abs_), np.array(cps)


@pytest.mark.parametrize("fbar", [0, 1])
def test_element_batch_matches_oracle(fbar, need_gpu):
    print(f"[element batch] fbar={fbar} n=64", flush=True)
    X, ue, ep, ab, cp = random_states(64, 11 + fbar, fbar)
    mat = O.Material(**MAT)
    Fe, Ke, sig = pa.element_batch("finite_fbar" if fbar else "finite", pa.J2Material(**MAT), X, ue,
                                   ep.copy(), ab.copy(), cp.copy())
    worst = 0.0
    for e in range(X.shape[0]):
        f, K, s = O.element_finite(fbar, mat, X[e], ue[e], ep[e].copy(), ab[e].copy(), cp[e].copy())
        worst = max(worst, rel(Fe[e], f), rel(Ke[e], K), rel(sig[e], s))
    print(f"  worst relative difference {worst:.2e}")
    assert worst < 1e-10


def test_element_batch_commit_matches_oracle(need_gpu):
    X, ue, ep, ab, cp = random_states(32, 5, 1)
    mat = O.Material(**MAT)
    e2, a2, c2 = ep.copy(), ab.copy(), cp.copy()
    pa.element_batch("finite_fbar", pa.J2Material(**MAT), X, ue, e2, a2, c2, commit=True)
    for e in range(X.shape[0]):
        eo, ao, co = ep[e].copy(), ab[e].copy(), cp[e].copy()
        O.element_finite(1, mat, X[e], ue[e], eo, ao, co, commit=True)
        assert rel(a2[e], ao) < 1e-10 and rel(c2[e], co) < 1e-12 and rel(e2[e], eo) < 1e-10


def test_golden_elements(need_gpu):
    g = np.load(os.path.join(HERE, "golden", "elements.npz"))
    Fe, Ke, sig = pa.element_batch("finite_fbar", pa.J2Material(**MAT), g["X"], g["ue"], g["epsp"].copy(),
                                   g["abar"].copy(), g["cpinv"].copy())
    assert rel(Fe, g["Fe"]) < 1e-10 and rel(Ke, g["Ke"]) < 1e-10 and rel(sig, g["tau"]) < 1e-10


def test_gpu_tangent_vs_fd_of_gpu_force(need_gpu):
    """F2 master gate on the device kernels: Ke vs central differences of Fe (<1e-6)."""
    X, ue, ep, ab, cp = random_states(1, 3, 1)
    mat = pa.J2Material(**MAT)
    _, Ke, _ = pa.element_batch("finite_fbar", mat, X, ue, ep.copy(), ab.copy(), cp.copy())
    h = 1e-9 * np.abs(ue).max() * 100
    Up = np.repeat(ue, 24, axis=0) + h * np.eye(24)
    Um = np.repeat(ue, 24, axis=0) - h * np.eye(24)
    rep = lambda a: np.ascontiguousarray(np.repeat(a, 24, axis=0))  # noqa: E731
    Fp, _, _ = pa.element_batch("finite_fbar", mat, rep(X), Up, rep(ep), rep(ab), rep(cp))
    Fm, _, _ = pa.element_batch("finite_fbar", mat, rep(X), Um, rep(ep), rep(ab), rep(cp))
    Kfd = ((Fp - Fm) / (2 * h)).T
    assert rel(Ke[0], Kfd) < 1e-6


def test_gpu_inversion_error(need_gpu):
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X = nodes[conn[0]]
    ue = (X @ np.diag([-1.1, 1.0, 1.0]).T - X).ravel()
    for kind in ("finite", "finite_fbar"):
        with pytest.raises(pa.ElementInversionError) as ei:
            pa.element_batch(kind, pa.J2Material(**MAT), X.ravel()[None], ue[None])
        assert ei.value.J < 0 and ei.value.e == 0


def test_degenerate_eigenvalue_branch(need_gpu):
    """Homogeneous uniaxial states have exactly repeated lateral eigenvalues: the
    EIG_TOL switch of _dhalflog_contract (FiniteStrain.jl:471-472) is taken."""
    mat = O.Material(**MAT)
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X = nodes[conn[0]]
    Xs, ues = [], []
    for s in (1.001, 1.01, 1.05, 0.97):
        F = np.diag([s, 1 - 0.4 * (s - 1), 1 - 0.4 * (s - 1)])
        Xs.append(X.ravel())
        ues.append((X @ F.T - X).ravel())
    Fe, Ke, sig = pa.element_batch("finite_fbar", pa.J2Material(**MAT), np.array(Xs), np.array(ues))
    for e in range(len(Xs)):
        f, K, s = O.element_finite(1, mat, Xs[e], ues[e])
        assert rel(Ke[e], K) < 1e-10 and rel(Fe[e], f) < 1e-10


def test_voce_hardening_matches_oracle(need_gpu):
    """Voce saturation hardening (Materials.jl:117-127 with the local Newton of :201-213)
    through the device return map: force, tangent, stress and committed history equal the
    oracle's to 1e-10, and the hardening is really nonlinear (history beyond the linear
    law's prediction)."""
    voce = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9, ssat=450e6, delta=60.0)
    mat = O.Material(**voce)
    rng = np.random.default_rng(3)
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X0 = nodes[conn[0]]
    Xs, ues, eps, abs_, cps = [], [], [], [], []
    for _ in range(32):
        X = X0 * (0.5 + rng.random(3)) + 0.05 * rng.normal(size=(8, 3))
        F1 = np.eye(3) + 0.04 * rng.normal(size=(3, 3))
        e, a, c = np.zeros((8, 6)), np.zeros(8), np.tile([1.0, 1, 1, 0, 0, 0], (8, 1))
        O.element_finite(1, mat, X.ravel(), (X @ F1.T - X).ravel(), e, a, c, commit=True)
        F2 = F1 + 0.02 * rng.normal(size=(3, 3))
        Xs.append(X.ravel()); ues.append((X @ F2.T - X).ravel())
        eps.append(e); abs_.append(a); cps.append(c)
    X, ue, ep, ab, cp = map(np.array, (Xs, ues, eps, abs_, cps))
    assert ab.max() > 1e-3  # plastic history with the saturation term active
    e2, a2, c2 = ep.copy(), ab.copy(), cp.copy()
    Fe, Ke, sig = pa.element_batch("finite_fbar", pa.J2Material(**voce), X, ue, e2, a2, c2, commit=True)
    worst = 0.0
    for e in range(X.shape[0]):
        e1, a1, c1 = ep[e].copy(), ab[e].copy(), cp[e].copy()
        f, K, s = O.element_finite(1, mat, X[e], ue[e], e1, a1, c1, commit=True)
        worst = max(worst, rel(Fe[e], f), rel(Ke[e], K), rel(sig[e], s), rel(a2[e], a1), rel(c2[e], c1))
    print(f"[voce] worst relative difference {worst:.2e}", flush=True)
    assert worst < 1e-10


def _kin_states(n, seed, fbar, mat_kw, amp=0.05):
    """Elements with a committed plastic state that holds a back stress, then a
    non-coaxial second deformation (the F2(c)/F9b states of test/test_finite_strain.jl)."""
    rng = np.random.default_rng(seed)
    mat = O.Material(**mat_kw)
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X0 = nodes[conn[0]]
    out = {k: [] for k in ("X", "ue", "ep", "bt", "ab", "cp")}
    for _ in range(n):
        X = X0 * (0.5 + rng.random(3)) + 0.05 * rng.normal(size=(8, 3))
        F1 = np.eye(3) + amp * rng.normal(size=(3, 3))
        e, b, a, c = np.zeros((8, 6)), np.zeros((8, 6)), np.zeros(8), np.tile([1.0, 1, 1, 0, 0, 0], (8, 1))
        O.element_finite(fbar, mat, X.ravel(), (X @ F1.T - X).ravel(), e, a, c, commit=True, beta=b)
        F2 = F1 + 0.5 * amp * rng.normal(size=(3, 3))
        out["X"].append(X.ravel()); out["ue"].append((X @ F2.T - X).ravel())
        out["ep"].append(e); out["bt"].append(b); out["ab"].append(a); out["cp"].append(c)
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.parametrize("fbar", [0, 1])
def test_kinematic_hardening_matches_oracle(fbar, need_gpu):
    """Objective back stress (FiniteStrain.jl:172-346): Fe, the non-symmetric Ke, tau and
    the committed history (beta_ref, abar, Cp^-1) within 1e-10 relative of the oracle."""
    kw = dict(E=210e3, nu=0.3, sy0=250.0, Hiso=200.0, Hkin=2000.0)
    s = _kin_states(48, 21 + fbar, fbar, kw)
    assert np.abs(s["bt"]).max() > 0 and s["ab"].max() > 0
    print(f"[kinematic] fbar={fbar} n=48, max |beta| {np.abs(s['bt']).max():.3g}", flush=True)
    mat = O.Material(**kw)
    e2, b2, a2, c2 = s["ep"].copy(), s["bt"].copy(), s["ab"].copy(), s["cp"].copy()
    kind = "finite_fbar" if fbar else "finite"
    Fe, Ke, sig = pa.element_batch(kind, pa.J2Material(**kw), s["X"], s["ue"], e2, a2, c2, commit=True, beta=b2)
    worst, asym = 0.0, 0.0
    for e in range(s["X"].shape[0]):
        e1, b1, a1, c1 = s["ep"][e].copy(), s["bt"][e].copy(), s["ab"][e].copy(), s["cp"][e].copy()
        f, K, t = O.element_finite(fbar, mat, s["X"][e], s["ue"][e], e1, a1, c1, commit=True, beta=b1)
        worst = max(worst, rel(Fe[e], f), rel(Ke[e], K), rel(sig[e], t), rel(b2[e], b1), rel(a2[e], a1),
                    rel(c2[e], c1), rel(e2[e], e1))
        asym = max(asym, rel(Ke[e], Ke[e].T))
    print(f"  worst relative difference {worst:.2e}; max asymmetry {asym:.2e}", flush=True)
    assert worst < 1e-10
    assert asym > 1e-6  # the back-stress rotation term is present (non-symmetric tangent)


def test_small_strain_element_matches_oracle(need_gpu):
    """Hex8Small (Elements.jl:323-360) with kinematic + isotropic hardening: Fe, Ke, sigma
    and committed eps_p, beta, abar within 1e-10 relative of the oracle."""
    kw = dict(E=210e3, nu=0.3, sy0=250.0, Hiso=1000.0, Hkin=500.0)
    mat = O.Material(**kw)
    rng = np.random.default_rng(8)
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X0 = nodes[conn[0]]
    Xs, us, eps, bts, abs_ = [], [], [], [], []
    for _ in range(48):
        X = X0 * (0.5 + rng.random(3)) + 0.05 * rng.normal(size=(8, 3))
        u1 = 4e-3 * rng.normal(size=24)
        e, b, a = np.zeros((8, 6)), np.zeros((8, 6)), np.zeros(8)
        O.element_small(mat, X.ravel(), u1, e, b, a, commit=True)
        Xs.append(X.ravel()); us.append(u1 + 2e-3 * rng.normal(size=24))
        eps.append(e); bts.append(b); abs_.append(a)
    X, ue, ep, bt, ab = map(np.array, (Xs, us, eps, bts, abs_))
    assert ab.max() > 0 and np.abs(bt).max() > 0
    e2, b2, a2 = ep.copy(), bt.copy(), ab.copy()
    Fe, Ke, sig = pa.element_batch("small", pa.J2Material(**kw), X, ue, e2, a2, None, commit=True, beta=b2)
    worst = 0.0
    for e in range(X.shape[0]):
        e1, b1, a1 = ep[e].copy(), bt[e].copy(), ab[e].copy()
        f, K, s = O.element_small(mat, X[e], ue[e], e1, b1, a1, commit=True)
        worst = max(worst, rel(Fe[e], f), rel(Ke[e], K), rel(sig[e], s), rel(e2[e], e1), rel(b2[e], b1),
                    rel(a2[e], a1))
    print(f"[small strain] worst relative difference {worst:.2e}", flush=True)
    assert worst < 1e-10
