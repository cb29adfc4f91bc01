"""The benchmark workloads as plain data (mesh arrays, materials, constraint lists).

Constraint set of every uniaxial workload: the x = 0 face is held in x, its corner
(0,0,0) is pinned in x, y, z, and the corner (0, Ly, 0) is held in z -- the last one
removes the rigid rotation about the x axis through the pinned corner that the
x-face + one-corner set leaves free (it does not restrain the Poisson contraction).
The x = Lx face gets the ramped axial displacement.

  workload 0  8^3 unit cube, 512 Hex8 F-bar, E=200 GPa nu=0.3 sy=250 MPa H=1 GPa,
              tension to 5% in 20 steps, Newton tol 1e-10, Krylov rtol 1e-10
  workload 1  100^3 block, +-10% h interior jitter (seed 0), compression to 20% in
              40 steps, Krylov rtol 1e-8
  workload 2  200^3 block, jitter seed 1, two-phase (20% ellipsoidal inclusions,
              sy 750 MPa vs 250 MPa), tension to 10% in 20 steps
  workload 3  samples for the learned preconditioner: workload-1-style blocks with
              jitter seed `seed` and ellipsoidal inclusions (sy 750 MPa) covering
              `inclusion` of the volume (seeds 2-20 train, 100-104 evaluate),
              compression in 320 steps
  workload 4  150^3 notched bar (20% section reduction), +-5% jitter (seed 7),
              tension to 30% in 60 steps
Sizes can be overridden (n=...) for parity runs on small meshes.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mesh import Mesh, box_mesh, ellipsoid_inclusions, notched_bar_mesh

STEEL = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9)


@dataclass
class Workload:
    name: str
    mesh: Mesh
    materials: list
    material_id: np.ndarray | None
    dir_dofs: np.ndarray
    dir_vals: np.ndarray
    dir_ramp: np.ndarray
    nsteps: int
    newton_tol: float
    krylov_rtol: float
    element: str = "finite_fbar"
    info: dict = field(default_factory=dict)


def uniaxial_constraints(mesh: Mesh, disp: float):
    X = mesh.nodes
    lo, hi = X.min(axis=0), X.max(axis=0)
    tol = 1e-8 * max(float((hi - lo).max()), 1.0)
    x0 = np.nonzero(np.abs(X[:, 0] - lo[0]) <= tol)[0]
    x1 = np.nonzero(np.abs(X[:, 0] - hi[0]) <= tol)[0]
    at = lambda y, z: x0[(np.abs(X[x0, 1] - y) <= tol) & (np.abs(X[x0, 2] - z) <= tol)]  # noqa: E731
    c0 = at(lo[1], lo[2])
    c1 = at(hi[1], lo[2])
    dofs, vals, ramp = [], [], []
    dofs += list(3 * x0)
    vals += [0.0] * len(x0)
    ramp += [False] * len(x0)
    for c in range(3):
        dofs += list(3 * c0 + c)
        vals += [0.0] * len(c0)
        ramp += [False] * len(c0)
    dofs += list(3 * c1 + 2)
    vals += [0.0] * len(c1)
    ramp += [False] * len(c1)
    dofs += list(3 * x1)
    vals += [float(disp)] * len(x1)
    ramp += [True] * len(x1)
    return (np.asarray(dofs, dtype=np.int64), np.asarray(vals, dtype=np.float64),
            np.asarray(ramp, dtype=np.uint8))


def workload(k: int, n: int | None = None, nsteps: int | None = None, seed: int | None = None,
             inclusion: float = 0.0, device: bool = False) -> Workload:
    """Workload k at size n; device=True generates the (bit-identical) mesh on the GPU."""
    if k == 0:
        n = 8 if n is None else n
        mesh = box_mesh(1.0, 1.0, 1.0, n, n, n)
        d, v, r = uniaxial_constraints(mesh, 0.05)
        return Workload("tension cube 8^3 F-bar", mesh, [STEEL], None, d, v, r, nsteps or 20, 1e-10, 1e-10)
    if k == 1:
        n = 100 if n is None else n
        mesh = box_mesh(1.0, 1.0, 1.0, n, n, n, jitter=0.1, seed=0, device=device)
        d, v, r = uniaxial_constraints(mesh, -0.2)
        return Workload(f"jittered block {n}^3 compression 20%", mesh, [STEEL], None, d, v, r, nsteps or 40,
                        1e-8, 1e-8)
    if k == 2:
        n = 200 if n is None else n
        mesh = box_mesh(1.0, 1.0, 1.0, n, n, n, jitter=0.1, seed=1, device=device)
        mid = ellipsoid_inclusions(mesh, 0.2, seed=1)
        d, v, r = uniaxial_constraints(mesh, 0.1)
        hard = dict(STEEL, sy0=750e6)
        return Workload(f"two-phase block {n}^3 tension 10%", mesh, [STEEL, hard], mid, d, v, r, nsteps or 20,
                        1e-8, 1e-8)
    if k == 3:
        n = 64 if n is None else n
        seed = 2 if seed is None else seed
        mesh = box_mesh(1.0, 1.0, 1.0, n, n, n, jitter=0.1, seed=seed)
        mid = ellipsoid_inclusions(mesh, inclusion, seed=seed) if inclusion > 0 else None
        d, v, r = uniaxial_constraints(mesh, -0.2)
        hard = dict(STEEL, sy0=750e6)
        mats = [STEEL, hard] if mid is not None else [STEEL]
        # 320 load steps: Newton without line search diverges or inverts elements of
        # two-phase blocks in compression at 0.125% strain increments (32^3 blocks)
        return Workload(f"sample block {n}^3 seed {seed} inclusions {inclusion:.2f}", mesh, mats, mid, d, v, r,
                        nsteps or 320, 1e-8, 1e-8)
    if k == 4:
        n = 150 if n is None else n
        mesh = notched_bar_mesh(n, length=1.0, notch_depth=0.2, jitter=0.05, seed=7)
        d, v, r = uniaxial_constraints(mesh, 0.3)
        return Workload(f"notched bar {n}^3 tension 30%", mesh, [STEEL], None, d, v, r, nsteps or 60, 1e-8, 1e-8)
    raise ValueError(f"no workload {k}")
