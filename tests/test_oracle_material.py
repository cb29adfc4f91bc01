"""Pins of the oracle's J2 return map (Materials.jl:159-250) against closed forms and
invariants, mirroring the reference tests T1-T8 (test/test_material.jl), T9-T15 Voce
(test/test_saturation.jl) and A1/A3/A5 (test/hard_validation.jl)."""
import numpy as np
import pytest

import oracle as O

Z6 = np.zeros(6)


def test_T1_elastic_round_trip():
    m = O.Material(210e3, 0.3, 250.0, Hiso=1000.0, Hkin=500.0)
    eps = np.array([1e-4, -2e-5, 3e-5, 1e-5, -1e-5, 2e-5])
    s, ep, bt, a, D = O.return_map(m, eps, Z6, Z6, 0.0)
    C = m.Cmat
    assert np.allclose(s, C @ eps, rtol=1e-12, atol=0)
    assert np.all(ep == 0) and np.all(bt == 0) and a == 0.0
    assert np.allclose(D, C, rtol=1e-12, atol=0)
    # Cmat closed form: lambda+2G on the normal diagonal, G (not 2G) on shear (Materials.jl:27-35)
    E, nu = 210e3, 0.3
    G, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    assert C[0, 0] == pytest.approx(lam + 2 * G, rel=1e-15) and C[3, 3] == pytest.approx(G, rel=1e-15)


def _uniaxial(m, exx_final, nsteps=50):
    """Drive one point in uniaxial stress by Newton on the lateral strains (T2 helper)."""
    ep, bt, a, elat = Z6.copy(), Z6.copy(), 0.0, 0.0
    for n in range(1, nsteps + 1):
        exx = exx_final * n / nsteps
        for _ in range(60):
            eps = np.array([exx, elat, elat, 0, 0, 0])
            s, _, _, _, D = O.return_map(m, eps, ep, bt, a)
            d = D[1, 1] + D[1, 2]
            if d == 0:
                break
            elat -= s[1] / d
            if abs(s[1]) < 1e-11:
                break
        eps = np.array([exx, elat, elat, 0, 0, 0])
        s, ep, bt, a, _ = O.return_map(m, eps, ep, bt, a)
    return s, ep, a


@pytest.mark.parametrize("exx", [0.004, 0.007, 0.01])
def test_T2_uniaxial_closed_form(exx):
    E, sy0, H = 210e3, 250.0, 1000.0
    m = O.Material(E, 0.3, sy0, Hiso=H)
    s, ep, a = _uniaxial(m, exx)
    Ht = E * H / (E + H)
    assert s[0] == pytest.approx(sy0 + Ht * (exx - sy0 / E), rel=1e-6)
    assert abs(s[1]) < 1e-6 * s[0]
    # abar is the equivalent plastic strain (Materials.jl:183-193): = eps_p axial
    assert a == pytest.approx(ep[0], rel=1e-9)
    # 258.77 MPa anchor (lagrangian/docs/DESIGN.md:829, lagrangian/README.md:60 of the reference)
    if exx == 0.01:
        assert s[0] == pytest.approx(258.77, abs=5e-3)


def test_T3_perfect_plasticity_cap():
    m = O.Material(210e3, 0.3, 250.0)
    s, _, _ = _uniaxial(m, 0.02, nsteps=200)
    assert s[0] <= 250.0 * (1 + 1e-10)
    assert s[0] == pytest.approx(250.0, rel=1e-3)


def test_A3_combined_hardening_uniaxial():
    E, sy0, Hi, Hk = 210e3, 250.0, 600.0, 400.0
    m = O.Material(E, 0.3, sy0, Hiso=Hi, Hkin=Hk)
    s, _, _ = _uniaxial(m, 0.01)
    Ht = E * (Hi + Hk) / (E + Hi + Hk)
    assert s[0] == pytest.approx(sy0 + Ht * (0.01 - sy0 / E), rel=1e-6)


def test_T4_bauschinger_reverse_yield():
    """pure kinematic hardening: after loading to s_max, reverse yield at s_max - 2 sy0."""
    E, sy0, Hk = 210e3, 250.0, 2000.0
    m = O.Material(E, 0.3, sy0, Hkin=Hk)
    ep, bt, a, elat = Z6.copy(), Z6.copy(), 0.0, 0.0

    def step(exx, ep, bt, a, elat):
        for _ in range(60):
            s, _, _, _, D = O.return_map(m, np.array([exx, elat, elat, 0, 0, 0]), ep, bt, a)
            elat -= s[1] / (D[1, 1] + D[1, 2])
            if abs(s[1]) < 1e-11:
                break
        s, ep2, bt2, a2, _ = O.return_map(m, np.array([exx, elat, elat, 0, 0, 0]), ep, bt, a)
        return s, ep2, bt2, a2, elat

    for n in range(1, 101):
        s, ep, bt, a, elat = step(0.01 * n / 100, ep, bt, a, elat)
    smax = s[0]
    exx = 0.01
    prev = s[0]
    while True:  # unload in small strain decrements until plastic flow restarts
        exx -= 1e-6
        s, ep2, bt2, a2, elat = step(exx, ep, bt, a, elat)
        if a2 > a:
            break
        prev = s[0]
        ep, bt = ep2, bt2
    assert prev == pytest.approx(smax - 2 * sy0, abs=E * 1e-6 * 1.01)


def _fd_D(m, eps, ep, bt, a, h=1e-8):
    D = np.zeros((6, 6))
    for j in range(6):
        e1, e2 = eps.copy(), eps.copy()
        e1[j] += h
        e2[j] -= h
        D[:, j] = (O.return_map(m, e1, ep, bt, a)[0] - O.return_map(m, e2, ep, bt, a)[0]) / (2 * h)
    return D


@pytest.mark.parametrize("seed", range(6))
def test_A1_tangent_vs_fd_many_states(seed):
    rng = np.random.default_rng(seed)
    m = O.Material(210e3, 0.3, 250.0, Hiso=1000.0, Hkin=300.0 * (seed % 2))
    ep = rng.normal(scale=1e-3, size=6)
    ep[:3] -= ep[:3].mean()
    bt = rng.normal(scale=20.0, size=6)
    bt[:3] -= bt[:3].mean()
    a = abs(rng.normal(scale=1e-3))
    eps = ep + rng.normal(scale=4e-3, size=6)
    s, _, _, a_new, D = O.return_map(m, eps, ep, bt, a)
    assert a_new > a  # plastic state
    Dfd = _fd_D(m, eps, ep, bt, a)
    assert np.linalg.norm(D - Dfd) / np.linalg.norm(D) < 1e-5
    assert np.linalg.norm(D - D.T) / np.linalg.norm(D) < 1e-12


def test_A5_incompressibility_and_isotropy():
    m = O.Material(210e3, 0.3, 250.0, Hiso=1000.0)
    eps = np.array([5e-3, -1e-3, 2e-3, 1e-3, 0, 0])
    s, ep, _, _, _ = O.return_map(m, eps, Z6, Z6, 0.0)
    assert abs(ep[0] + ep[1] + ep[2]) < 1e-12 * np.abs(ep).max()
    # mean stress unchanged by the plastic corrector
    st = m.Cmat @ eps
    assert (s[:3].sum() / 3) == pytest.approx(st[:3].sum() / 3, rel=1e-12)
    # permutation isotropy: swapping x<->y in the input swaps the output
    P = [1, 0, 2, 3, 5, 4]
    s2, _, _, _, _ = O.return_map(m, eps[P], Z6, Z6, 0.0)
    assert np.allclose(s2, s[P], rtol=1e-12, atol=1e-12 * np.abs(s).max())
    # on the yield surface after return
    p = s[:3].mean()
    dev = s - np.r_[p, p, p, 0, 0, 0]
    q = np.sqrt(1.5 * (dev[0] ** 2 + dev[1] ** 2 + dev[2] ** 2 + 2 * (dev[3] ** 2 + dev[4] ** 2 + dev[5] ** 2)))
    a = O.return_map(m, eps, Z6, Z6, 0.0)[3]
    assert q == pytest.approx(250.0 + 1000.0 * a, rel=1e-12)


def test_voce_reduces_to_linear_bitwise():
    """T9 (test/test_saturation.jl): ssat == sy0 or delta == 0 gives the linear law bit-for-bit."""
    eps = np.array([6e-3, -2e-3, -1e-3, 1e-3, 5e-4, -3e-4])
    lin = O.return_map(O.Material(210e3, 0.3, 250.0, Hiso=800.0), eps, Z6, Z6, 0.01)
    for m in (O.Material(210e3, 0.3, 250.0, Hiso=800.0, ssat=250.0, delta=15.0),
              O.Material(210e3, 0.3, 250.0, Hiso=800.0, ssat=400.0, delta=0.0)):
        v = O.return_map(m, eps, Z6, Z6, 0.01)
        for x, y in zip(lin, v):
            assert np.array_equal(np.asarray(x), np.asarray(y))


def test_voce_on_surface_and_fd():
    m = O.Material(210e3, 0.3, 250.0, Hiso=100.0, ssat=450.0, delta=20.0)
    eps = np.array([8e-3, -3e-3, -2e-3, 1e-3, 0, 2e-3])
    s, ep, bt, a, D = O.return_map(m, eps, Z6, Z6, 0.0)
    p = s[:3].mean()
    dev = s - np.r_[p, p, p, 0, 0, 0]
    q = np.sqrt(1.5 * (dev[0] ** 2 + dev[1] ** 2 + dev[2] ** 2 + 2 * (dev[3] ** 2 + dev[4] ** 2 + dev[5] ** 2)))
    ys = 250.0 + 200.0 * (1 - np.exp(-20.0 * a)) + 100.0 * a
    assert q == pytest.approx(ys, rel=1e-12)
    assert np.linalg.norm(D - _fd_D(m, eps, Z6, Z6, 0.0)) / np.linalg.norm(D) < 1e-6


def test_material_guards():
    for kw in ({"E": -1}, {"nu": 0.5}, {"sy0": 0}, {"Hiso": -1}):
        args = {"E": 210e3, "nu": 0.3, "sy0": 250.0}
        args.update(kw)
        with pytest.raises(ValueError):
            O.Material(**args)
