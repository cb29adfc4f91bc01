"""Writes the golden fixtures under tests/golden/ from the CPU oracle (FP64, direct
linear solves).  The reference (Julia) cannot run in this image and ships no golden
vectors of its own, so these fixtures are the oracle's outputs; the oracle itself is
pinned by the closed-form / finite-difference / invariant tests in tests/.

  config0.npz   workload 0 (8^3 F-bar tension to 5%, 20 steps, tol 1e-10):
                U and abar after steps 1, 10 and 20, Newton residual histories
  jitter6.npz   6^3 jittered block (seed 0, +-10%), compression 20% in 80 steps, tol 1e-10
  elements.npz  per-element Fe/Ke/tau for distorted elements at plastic states
Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from plasticity_amd.workloads import workload  # noqa: E402


def run_workload(wl, keep_steps):
    mat = O.Material(**wl.materials[0])
    M = O.OracleModel(wl.mesh.nodes, wl.mesh.elements, mat, element=wl.element)
    M.constrain(wl.dir_dofs, wl.dir_vals, wl.dir_ramp)
    out = {}

    def cb(n, model):
        if n in keep_steps:
            out[f"U_{n}"] = model.U.copy()
            out[f"abar_{n}"] = model.st_c["abar"].copy()
            if n == max(keep_steps):
                out[f"tau_{n}"] = model.st_c["sig"].copy()

    ok, iters, res, _ = M.solve(nsteps=wl.nsteps, tol=wl.newton_tol, callback=cb)
    assert ok
    out["newton_iters"] = np.array(iters)
    out["res_last"] = np.array(res[-1])
    return out


def element_states():
    rng = np.random.default_rng(7)
    mat = O.Material(200e9, 0.3, 250e6, Hiso=1e9)
    nodes, conn = O.box_mesh(1, 1, 1, 1, 1, 1)
    X0 = nodes[conn[0]]
    Xs, ues, ep, ab, cp, Fe, Ke, tau = [], [], [], [], [], [], [], []
    for k in range(6):
        X = X0 * (0.5 + rng.random(3)) + 0.06 * rng.normal(size=(8, 3))
        F1 = np.eye(3) + 0.03 * rng.normal(size=(3, 3))
        F2 = F1 + 0.02 * rng.normal(size=(3, 3))
        e, a, c = np.zeros((8, 6)), np.zeros(8), np.tile([1.0, 1, 1, 0, 0, 0], (8, 1))
        O.element_finite(1, mat, X.ravel(), (X @ F1.T - X).ravel(), e, a, c, commit=True)
        ue = (X @ F2.T - X).ravel() + 1e-3 * rng.normal(size=24)
        f, K, s = O.element_finite(1, mat, X.ravel(), ue, e.copy(), a.copy(), c.copy())
        Xs.append(X.ravel()); ues.append(ue); ep.append(e); ab.append(a); cp.append(c)
        Fe.append(f); Ke.append(K); tau.append(s)
    return dict(X=np.array(Xs), ue=np.array(ues), epsp=np.array(ep), abar=np.array(ab), cpinv=np.array(cp),
                Fe=np.array(Fe), Ke=np.array(Ke), tau=np.array(tau))


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "elements.npz"), **element_states())
    np.savez_compressed(os.path.join(HERE, "config0.npz"), **run_workload(workload(0), {1, 10, 20}))
    wl = workload(1, n=6, nsteps=80)
    wl.newton_tol = wl.krylov_rtol = 1e-10
    np.savez_compressed(os.path.join(HERE, "jitter6.npz"), **run_workload(wl, {20, 40, 80}))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
