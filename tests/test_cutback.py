"""The load-step bisection rule (lagrangian/docs/SCALING.md:268-274) on the CPU: the host
driver of the device path (Model.advance / Model.solve) and the oracle's solve take the
same increments for the same failures.  Newton itself is replaced by a stub that fails
whenever the increment exceeds a limit, so the attempted load factors can be written down
by hand:

  nsteps = 4 (nominal increment 0.25), limit 0.1, grow after 2 successes, each failure
  doubling the successes needed (at most 16), a full nominal increment resetting them:
  0.25 x  0.125 x  0.0625 ok x8 (grow: 0.125)  0.875 x ...
The stub's own sequence is checked against the hand-worked one below."""
import numpy as np
import pytest

import oracle as O
import plasticity_amd.model as PM


def expected(nsteps, limit, cutback, grow_after, nload, iters=2, grow_iters=6):
    """Hand-rolled restatement of the rule, independent of both implementations (every
    converged increment takes `iters` Newton iterations)."""
    nominal = 1.0 / nsteps
    lam_c, dl, streak, need, out = 0.0, nominal, 0, grow_after, []
    for n in range(1, nload + 1):
        target = n / nsteps
        while lam_c < target:
            lam = min(lam_c + dl, target)
            if abs(lam - target) <= 1e-12 * nominal:
                lam = target
            ok = lam - lam_c <= limit + 1e-15
            out.append((round(lam, 12), ok))
            if ok:
                if lam - lam_c >= nominal * (1 - 1e-9):  # a full nominal increment: back to grow_after
                    need = grow_after
                lam_c, streak = lam, (streak + 1 if iters <= grow_iters else 0)
                if streak >= need and dl < nominal:
                    dl, streak = min(2 * dl, nominal), 0
            else:
                if (lam - lam_c) / 2 < nominal / 2 ** cutback * (1 - 1e-9):
                    return out, False
                dl, streak, need = (lam - lam_c) / 2, 0, min(2 * need, 8 * grow_after)  # each failure doubles it
    return out, True


class _StubModel(PM.Model):
    """Model.advance / solve against a fake newton!: no device involved."""

    def __init__(self, limit, iters=2):
        self.limit, self.lam_c_dev, self.calls, self.iters = limit, 0.0, [], iters
        self._lam_c, self._dl, self._streak, self._need = 0.0, None, 0, None
        self.dir_dofs, self.neu_dofs = [], []

    def reset(self):
        self._lam_c, self._dl, self._streak, self._need = 0.0, None, 0, None
        self.lam_c_dev = 0.0
        return self

    def _freeze(self):
        pass

    def newton(self, lam, tol=1e-8, maxiter=25):
        ok = lam - self.lam_c_dev <= self.limit + 1e-15
        self.calls.append((round(lam, 12), ok))

        class R:
            solves, stopped = 2, 0
            solver_used = [0, 0]
            t_assemble = t_bc = t_setup = t_solve = t_total = 0.0
        if not ok and len(self.calls) % 2:
            raise PM.ElementInversionError(0, 0, -1.0)  # inversions count as failures too
        return ok, [1.0] * (self.iters - 1) + [0.0] if ok else [1.0, 2.0], [3, 4], R()

    def commit(self):
        self.lam_c_dev = self.calls[-1][0]

    def rollback(self):
        pass


@pytest.mark.parametrize("nsteps,limit,cutback,nload", [(4, 0.1, 4, 4), (20, 0.0126, 6, 3), (10, 0.1, 2, 10),
                                                        (8, 0.01, 2, 2)])
def test_host_driver_follows_the_rule(nsteps, limit, cutback, nload):
    exp, ok = expected(nsteps, limit, cutback, 2, nload)
    m = _StubModel(limit)
    res = m.solve(nsteps=nsteps, steps=range(1, nload + 1), cutback=cutback)
    assert m.calls == exp
    assert res.converged == ok
    assert res.cutbacks == sum(1 for _, c in exp if not c)


@pytest.mark.parametrize("iters", [6, 7])
def test_slow_increments_do_not_grow(iters):
    """Converged increments that needed more than grow_iters (6) Newton iterations never
    grow the increment again: no failed attempt after the first cut."""
    exp, ok = expected(4, 0.1, 4, 2, 3, iters=iters)
    m = _StubModel(0.1, iters=iters)
    m.solve(nsteps=4, steps=range(1, 4), cutback=4)
    assert m.calls == exp
    fails = sum(1 for _, c in exp if not c)
    assert (fails == 2) if iters == 7 else (fails > 2)  # 0.25 x, 0.125 x, then 0.0625 for good


def test_hand_worked_sequence():
    """nsteps 4, limit 0.1: 0.25 x, 0.125 x (8 successes now needed), eight increments of
    0.0625 up to 0.5, then 0.625 x (0.125 > 0.1; 16 needed), 0.5625 ok ..."""
    exp, _ = expected(4, 0.1, 4, 2, 3)
    lams = [round(0.0625 * k, 12) for k in range(1, 9)]
    assert exp[:11] == [(0.25, False), (0.125, False)] + [(l, True) for l in lams] + [(0.625, False)]
    assert exp[11] == (0.5625, True)


def test_without_cutback_the_first_failure_ends_the_solve():
    m = _StubModel(0.1)
    with pytest.warns(UserWarning):
        res = m.solve(nsteps=4, cutback=0)
    assert m.calls == [(0.25, False)] and not res.converged


@pytest.mark.parametrize("nsteps,limit,cutback,nload", [(4, 0.1, 4, 4), (20, 0.0126, 6, 3), (8, 0.01, 2, 2)])
def test_oracle_follows_the_rule(nsteps, limit, cutback, nload, monkeypatch):
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    M = O.OracleModel(nodes, conn, O.Material(200e9, 0.3, 250e6, 1e9))
    state = {"lam_c": 0.0}

    def newton(lam, Fext, bcs, tol=1e-8, maxiter=25, solver=None, diverge=0.0):
        ok = lam - state["lam_c"] <= limit + 1e-15
        if ok:
            state["lam_c"] = lam
        return ok, [1.0, 0.0], []

    monkeypatch.setattr(M, "newton", newton)
    monkeypatch.setattr(M, "commit", lambda: None)
    recs = []
    conv, *_ = M.solve(nsteps=nsteps, cutback=cutback, records=recs, steps=range(1, nload + 1))
    exp, ok = expected(nsteps, limit, cutback, 2, nload)
    assert [(round(lam, 12), c) for lam, c, _ in recs] == exp
    assert conv == ok


def test_oracle_gauss_stress_homogeneous_closed_form():
    """gauss_stress on a jittered mesh under a homogeneous deformation U = (F - I) X: every
    Gauss point has the same un-barred F, so sigma = tau / det F exactly where det F > 0,
    and tau itself where det F <= 0 (Model.jl:203-206); the small-strain kind copies tau."""
    nodes, conn = O.box_mesh(1, 1, 1, 3, 2, 2, jitter=0.2, seed=9)
    M = O.OracleModel(nodes, conn, O.Material(200e9, 0.3, 250e6, 1e9))
    rng = np.random.default_rng(1)
    tau = rng.standard_normal((8 * M.ne, 6))
    M.st_c["sig"] = tau
    for F in (np.array([[1.02, 0.01, 0.0], [0.003, 0.97, 0.02], [0.0, -0.01, 1.05]]),
              np.diag([-0.5, 1.0, 1.0])):
        M.U = ((nodes @ F.T) - nodes).ravel()
        sig = M.gauss_stress()
        J = np.linalg.det(F)
        np.testing.assert_allclose(sig, tau / J if J > 0 else tau, rtol=1e-12, atol=0)
    Ms = O.OracleModel(nodes, conn, O.Material(200e9, 0.3, 250e6, 1e9), element="small")
    Ms.st_c["sig"] = tau
    Ms.U = rng.standard_normal(Ms.ndof)
    assert np.array_equal(Ms.gauss_stress(), tau)
