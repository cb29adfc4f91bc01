"""Host-side mirror of the reference's Model / solve! API over libpfem (HIP, gfx950).

Reference (lagrangian/src):
  J2Material        Materials.jl:82-108
  Model             Model.jl:63-75         fix/prescribe/load Model.jl:102-145
  reset             Model.jl:157-162
  assemble          Assembly.jl:238-257    (assemble!; K stays on the device)
  linear_solve      Solver.jl:235-290      (_linear_solve!)
  newton            Solver.jl:317-389      (newton!)
  solve             Solver.jl:419-475      (solve!)
  nodal_displacements / gauss_stress / gauss_kirchhoff / equivalent_plastic_strain
                    Model.jl:169-232
Indices are 0-based.  Every compute call runs on the GPU through the C-ABI; there is
no CPU path.
"""
from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .mesh import Mesh

__all__ = ["J2Material", "Model", "SolveResult", "ElementInversionError", "element_batch",
           "default_options"]


class ElementInversionError(RuntimeError):
    """Mirror of FiniteStrain.ElementInversionError(e, g, J) (FiniteStrain.jl:42-52)."""

    def __init__(self, e, g, J):
        super().__init__(f"element {e} Gauss point {g}: det F = {J} <= 0")
        self.e, self.g, self.J = int(e), int(g), float(J)


class J2Material:
    """J2Material(E, nu, sy0, Hiso=0, Hkin=0, ssat=sy0, delta=0) (Materials.jl:82-108)."""

    def __init__(self, E, nu, sy0, Hiso=0.0, Hkin=0.0, ssat=None, delta=0.0):
        ssat = sy0 if ssat is None else ssat
        if not E > 0:
            raise ValueError(f"E must be positive (got {E})")
        if not -1 < nu < 0.5:
            raise ValueError(f"nu must satisfy -1 < nu < 0.5 (got {nu})")
        if not sy0 > 0:
            raise ValueError(f"sy0 must be positive (got {sy0})")
        if not Hiso >= 0:
            raise ValueError(f"Hiso must be >= 0 (got {Hiso})")
        if not Hkin >= 0:
            raise ValueError(f"Hkin must be >= 0 (got {Hkin})")
        if not ssat >= sy0:
            raise ValueError(f"ssat must be >= sy0 (got {ssat})")
        if not delta >= 0:
            raise ValueError(f"delta must be >= 0 (got {delta})")
        if delta > 0 and ssat > sy0 and Hkin > 0:
            raise ValueError("saturation hardening with kinematic hardening is not supported")
        self.E, self.nu, self.sy0, self.Hiso, self.Hkin = float(E), float(nu), float(sy0), float(Hiso), float(Hkin)
        self.ssat, self.delta = float(ssat), float(delta)
        self.G = E / (2 * (1 + nu))
        self.K = E / (3 * (1 - 2 * nu))
        self.lam = E * nu / ((1 + nu) * (1 - 2 * nu))

    def _c(self):
        return L.Material(self.E, self.nu, self.sy0, self.Hiso, self.Hkin, self.ssat, self.delta)


@dataclass
class SolveResult:
    """SolveResult (Solver.jl:61-69) plus per-iteration solver data and timings."""
    converged: bool
    iters: list = field(default_factory=list)
    residuals: list = field(default_factory=list)
    cg_iters: list = field(default_factory=list)
    solver_used: list = field(default_factory=list)
    times: list = field(default_factory=list)


def default_options(**kw):
    o = L.Options()
    L.lib().pfem_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "amg_rebuild":  # 0: once per load step (Solver.jl:374), 1: every Newton iteration
            o.reserved[0] = int(v)
            continue
        if k == "fcg_directions":  # directions kept by the learned preconditioner's flexible GCR
            o.reserved[1] = int(v)
            continue
        if k == "rho_iters":  # (cold, warm) power-iteration steps of the Chebyshev bound estimate
            o.reserved[4], o.reserved[5] = int(v[0]), int(v[1])
            continue
        if k == "line_search":  # Newton backtracking on ||R||: most step halvings (0 = full steps)
            o.reserved[3] = int(v)
            continue
        if k == "gcr_fallback":  # after a Krylov breakdown: flexible GCR(16) with the AMG cycle (default) or BiCGStab
            o.reserved[6] = (o.reserved[6] & ~128) if v else (o.reserved[6] | 128)
            continue
        if k == "wcycle":  # AMG cycle: True W, False / None V (the default)
            o.reserved[6] = (o.reserved[6] | 16) if v else (o.reserved[6] & ~16)
            continue
        if k == "stream_prolongation":  # fine prolongation over row chunks (default) or one thread per row
            o.reserved[6] = (o.reserved[6] & ~256) if v else (o.reserved[6] | 256)
            continue
        if k == "fp16_fine":  # FP32-IR: smoother products on the fine level with the FP16 scaled K
            o.reserved[2] = 1 if v else 0
            continue
        if k == "precision" and isinstance(v, str):
            v = {"fp64": L.PREC_FP64, "fp32_ir": L.PREC_FP32_IR}[v]
        if k == "smoother" and isinstance(v, str):
            v = {"jacobi": L.SMOOTH_JACOBI, "chebyshev": L.SMOOTH_CHEBYSHEV, "chebyshev4": L.SMOOTH_CHEBYSHEV4}[v]
        setattr(o, k, v)
    return o


_KIND = {"small": L.HEX8_SMALL, "finite": L.HEX8_FINITE, "finite_fbar": L.HEX8_FINITE_FBAR}


def _comps(comp):
    if comp == "all":
        return (0, 1, 2)
    return ("xyz".index(comp),)


class Model:
    """Model(mesh, material; element) (Model.jl:63-75).  Device-resident."""

    def __init__(self, mesh: Mesh, material, element="finite_fbar", options=None, material_id=None):
        if element not in _KIND:
            raise ValueError(f"unknown element kind {element!r} (use 'small', 'finite' or 'finite_fbar')")
        mats = list(material) if isinstance(material, (list, tuple)) else [material]
        self.mesh = mesh
        self.materials = mats
        self.material = mats[0]
        self.element = element
        self.ngp = 8 * mesh.nelem
        self.ndof = 3 * mesh.nnodes
        self._nodes = np.ascontiguousarray(mesh.nodes, dtype=np.float64)
        self._conn = np.ascontiguousarray(mesh.elements, dtype=np.int64)
        self._matid = None if material_id is None else np.ascontiguousarray(material_id, dtype=np.int32)
        cm = (L.Material * len(mats))(*[m._c() for m in mats])
        self._cm = cm
        desc = L.MeshDesc(mesh.nnodes, mesh.nelem, L.ptr(self._nodes), L.ptr(self._conn),
                          None if self._matid is None else L.ptr(self._matid), len(mats),
                          C.cast(cm, C.c_void_p), _KIND[element])
        self.options = options if options is not None else default_options()
        h = C.c_void_p()
        L.check(L.lib().pfem_create(C.byref(desc), C.byref(self.options), C.byref(h)), "pfem_create")
        self._h = h
        self.dir_dofs, self.dir_vals, self.dir_ramp = [], [], []
        self.neu_dofs, self.neu_vals = [], []

    def close(self):
        if getattr(self, "_h", None):
            L.lib().pfem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- BC builders
    def fix(self, nodes, comp="all"):
        for c in _comps(comp):
            for n in np.asarray(nodes, dtype=np.int64):
                self.dir_dofs.append(3 * int(n) + c)
                self.dir_vals.append(0.0)
                self.dir_ramp.append(False)
        return self

    def prescribe(self, nodes, comp, value, ramp=True):
        for c in _comps(comp):
            for n in np.asarray(nodes, dtype=np.int64):
                self.dir_dofs.append(3 * int(n) + c)
                self.dir_vals.append(float(value))
                self.dir_ramp.append(bool(ramp))
        return self

    def load(self, nodes, comp, value, distribute=False, ramp=True):
        nodes = np.asarray(nodes, dtype=np.int64)
        per = float(value) / len(nodes) if distribute else float(value)
        for c in _comps(comp):
            for n in nodes:
                self.neu_dofs.append(3 * int(n) + c)
                self.neu_vals.append(per)
        if not ramp:
            warnings.warn("load ramp=False is not supported; the load is ramped")
        return self

    def constrain(self, dofs, values, ramp):
        """Raw Dirichlet entries (global 0-based dofs), appended like fix/prescribe."""
        self.dir_dofs += [int(d) for d in dofs]
        self.dir_vals += [float(v) for v in values]
        self.dir_ramp += [bool(r) for r in ramp]
        return self

    def _freeze(self):
        d = np.ascontiguousarray(self.dir_dofs, dtype=np.int64)
        v = np.ascontiguousarray(self.dir_vals, dtype=np.float64)
        r = np.ascontiguousarray(self.dir_ramp, dtype=np.uint8)
        L.check(L.lib().pfem_set_dirichlet(self._h, len(d), L.ptr(d), L.ptr(v), L.ptr(r)), "set_dirichlet")
        nd = np.ascontiguousarray(self.neu_dofs, dtype=np.int64)
        nv = np.ascontiguousarray(self.neu_vals, dtype=np.float64)
        L.check(L.lib().pfem_set_neumann(self._h, len(nd), L.ptr(nd), L.ptr(nv)), "set_neumann")

    # ---------------------------------------------------------------- state
    def reset(self):
        L.check(L.lib().pfem_reset(self._h), "reset")
        return self

    @property
    def U(self):
        u = np.empty(self.ndof)
        L.check(L.lib().pfem_get_displacement(self._h, L.ptr(u)), "get_displacement")
        return u

    @U.setter
    def U(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        L.check(L.lib().pfem_set_displacement(self._h, L.ptr(u)), "set_displacement")

    def get_state(self):
        """Committed GaussState (Mesh.jl:124-163): epsp, beta (back stress), abar, cpinv and
        the last assembled stress sig (Kirchhoff for finite strain, Cauchy for small)."""
        ep, bt, ab, cp, sg = (np.empty((self.ngp, 6)), np.empty((self.ngp, 6)), np.empty(self.ngp),
                              np.empty((self.ngp, 6)), np.empty((self.ngp, 6)))
        L.check(L.lib().pfem_get_state(self._h, L.ptr(ep), L.ptr(bt), L.ptr(ab), L.ptr(cp), L.ptr(sg)),
                "get_state")
        return {"epsp": ep, "beta": bt, "abar": ab, "cpinv": cp, "sig": sg}

    def set_state(self, epsp=None, abar=None, cpinv=None, beta=None):
        args = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (epsp, beta, abar, cpinv)]
        L.check(L.lib().pfem_set_state(self._h, *[None if a is None else L.ptr(a) for a in args]), "set_state")

    def nodal_displacements(self):
        """3 x nnodes (Model.jl:169-178)."""
        return self.U.reshape(-1, 3).T.copy()

    def gauss_kirchhoff(self):
        return self.get_state()["sig"]

    def equivalent_plastic_strain(self):
        return self.get_state()["abar"]

    def gauss_stress(self):
        """Cauchy stress, ngp x 6 (Model.jl:188-208): the stored stress for the small-strain
        element, tau / det F with the unbarred F for the finite-strain elements."""
        tau = self.gauss_kirchhoff()
        if self.element == "small":
            return tau
        U = self.U.reshape(-1, 3)
        X = self._nodes
        g = 1.0 / np.sqrt(3.0)
        sgn = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                        [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
        out = np.empty_like(tau)
        Xe = X[self._conn]      # nelem x 8 x 3
        ue = U[self._conn]
        for gp in range(8):
            xi = sgn[gp] * g
            dN = np.empty((8, 3))
            for a in range(8):
                s = sgn[a]
                dN[a] = 0.125 * np.array([s[0] * (1 + s[1] * xi[1]) * (1 + s[2] * xi[2]),
                                          s[1] * (1 + s[0] * xi[0]) * (1 + s[2] * xi[2]),
                                          s[2] * (1 + s[0] * xi[0]) * (1 + s[1] * xi[1])])
            J = np.einsum("eai,aj->eij", Xe, dN)
            dNdX = np.einsum("ak,ekj->eaj", dN, np.linalg.inv(J))
            F = np.eye(3) + np.einsum("eai,eaj->eij", ue, dNdX)
            det = np.linalg.det(F)
            inv = np.where(det > 0, 1.0 / det, 1.0)
            out[gp::8] = tau[gp::8] * inv[:, None]
        return out

    # ---------------------------------------------------------------- hot path
    def assemble(self, U=None, commit=False):
        """assemble!: K (device) and R = F_int; returns R.  Raises ElementInversionError."""
        R = np.empty(self.ndof)
        f = L.Fail()
        Ua = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        rc = L.lib().pfem_assemble(self._h, None if Ua is None else L.ptr(Ua), int(commit), L.ptr(R), C.byref(f))
        if rc == L.PFEM_ERR_INVERTED:
            raise ElementInversionError(f.e, f.g, f.J)
        L.check(rc, "assemble")
        return R

    def residual(self):
        R = np.empty(self.ndof)
        L.check(L.lib().pfem_get_R(self._h, L.ptr(R)), "get_R")
        return R

    def impose_dirichlet(self, lam):
        L.check(L.lib().pfem_impose_dirichlet(self._h, float(lam)), "impose_dirichlet")

    def subtract_external(self, lam):
        L.check(L.lib().pfem_subtract_external(self._h, float(lam)), "subtract_external")

    def residual_norm(self):
        out = C.c_double()
        L.check(L.lib().pfem_residual_norm(self._h, C.byref(out)), "residual_norm")
        return out.value

    def linear_solve(self, rtol, rebuild_pc=True, strict=False):
        """_linear_solve!(dU, K, -R, ...) -> (dU, niter, solver_used, true_relres)."""
        it, used, rel = C.c_int(), C.c_int(), C.c_double()
        rc = L.lib().pfem_linear_solve(self._h, float(rtol), int(rebuild_pc), C.byref(it), C.byref(used),
                                       C.byref(rel))
        if rc in (L.PFEM_ERR_NOT_CONVERGED, L.PFEM_ERR_NON_SPD) and not strict:
            warnings.warn(f"linear solve: {L.lib().pfem_status_string(rc).decode()}")
        else:
            L.check(rc, "linear_solve")
        dU = np.empty(self.ndof)
        L.check(L.lib().pfem_get_dU(self._h, L.ptr(dU)), "get_dU")
        return dU, it.value, used.value, rel.value

    def get_K_csc(self):
        import scipy.sparse as sp
        nnz = C.c_int64()
        L.check(L.lib().pfem_get_K_csc(self._h, C.byref(nnz), None, None, None), "get_K_csc")
        cp = np.empty(self.ndof + 1, dtype=np.int64)
        rv = np.empty(nnz.value, dtype=np.int64)
        nz = np.empty(nnz.value)
        L.check(L.lib().pfem_get_K_csc(self._h, C.byref(nnz), L.ptr(cp), L.ptr(rv), L.ptr(nz)), "get_K_csc")
        return sp.csc_matrix((nz, rv, cp), shape=(self.ndof, self.ndof))

    def get_K_bsr(self):
        nnzb = C.c_int64()
        L.check(L.lib().pfem_get_K_bsr(self._h, C.byref(nnzb), None, None, None), "get_K_bsr")
        rp = np.empty(self.mesh.nnodes + 1, dtype=np.int32)
        ci = np.empty(nnzb.value, dtype=np.int32)
        v = np.empty((nnzb.value, 3, 3))
        L.check(L.lib().pfem_get_K_bsr(self._h, C.byref(nnzb), L.ptr(rp), L.ptr(ci), L.ptr(v)), "get_K_bsr")
        return rp, ci, v

    def colors(self):
        col = np.empty(self.mesh.nelem, dtype=np.int32)
        nc = C.c_int32()
        L.check(L.lib().pfem_get_colors(self._h, L.ptr(col), C.byref(nc)), "get_colors")
        return col, nc.value

    def newton(self, lam, tol=1e-8, maxiter=25):
        """newton! one load step -> (converged, residual history, krylov iterations, result)."""
        res = L.NewtonResult()
        f = L.Fail()
        rc = L.lib().pfem_newton(self._h, float(lam), float(tol), int(maxiter), C.byref(res), C.byref(f))
        if rc == L.PFEM_ERR_INVERTED:
            raise ElementInversionError(f.e, f.g, f.J)
        L.check(rc, "newton")
        n = res.iters
        hist = list(res.residuals[:n])
        kry = [res.krylov_iters[i] for i in range(n) if not (res.converged and i == n - 1)]
        return bool(res.converged), hist, kry, res

    def commit(self):
        f = L.Fail()
        rc = L.lib().pfem_commit(self._h, C.byref(f))
        if rc == L.PFEM_ERR_INVERTED:
            raise ElementInversionError(f.e, f.g, f.J)
        L.check(rc, "commit")

    def solve(self, nsteps=10, tol=1e-8, maxiter=25, verbose=False, callback=None, steps=None):
        """solve! (Solver.jl:419-475): reset, freeze BCs, ramp lambda = n/nsteps."""
        self.reset()
        self._freeze()
        out = SolveResult(converged=True)
        for n in (range(1, nsteps + 1) if steps is None else steps):
            lam = n / nsteps
            conv, hist, kry, res = self.newton(lam, tol=tol, maxiter=maxiter)
            out.iters.append(len(hist))
            out.residuals.append(hist)
            out.cg_iters.append(kry)
            out.solver_used.append([res.solver_used[i] for i in range(len(kry))])
            out.times.append({"assemble": res.t_assemble, "bc": res.t_bc, "setup": res.t_setup,
                              "solve": res.t_solve, "total": res.t_total})
            if verbose:
                print(f"load step {n}/{nsteps} (lambda={lam:.4f}) newton={len(hist)} krylov={kry} "
                      f"|R|={hist[-1]:.3e} t={res.t_total:.3f}s")
            if not conv:
                out.converged = False
                warnings.warn(f"load step {n} did not converge in {maxiter} iterations")
                break
            self.commit()
            if callback is not None:
                callback(n, self)
        return out

    def write_vtu(self, filename):
        """write_vtu (Visualization.jl:129-203): ASCII .vtu of the committed state."""
        from .vtu import write_vtu
        return write_vtu(filename, self)

    def attach_net(self, net, mode="precondition", stride=1):
        """Learned preconditioner (pfem_attach_net): 'log' samples, 'precondition', 'off'."""
        from .learned import attach
        attach(self, net, mode, stride)
        return self

    def stats(self):
        st = L.Stats()
        L.check(L.lib().pfem_get_stats(self._h, C.byref(st)), "get_stats")
        return st

    def bench_spmv(self, prec=0, reps=20):
        sec = C.c_double()
        L.check(L.lib().pfem_bench_spmv(self._h, int(prec), int(reps), C.byref(sec)), "bench_spmv")
        return sec.value


def element_batch(kind, material: J2Material, Xe, ue, epsp=None, abar=None, cpinv=None, commit=False,
                  beta=None):
    """Per-element Fe (n,24), Ke (n,24,24) and stress (n,8,6) computed by the production
    GPU kernels for independent elements (Elements.jl:405-494 finite strain, :323-360 small
    strain; Kirchhoff stress for finite strain).  History arrays (n,8,6)/(n,8) -- beta is
    the back stress -- are updated in place when commit is true."""
    Xe = np.ascontiguousarray(Xe, dtype=np.float64).reshape(-1, 24)
    ue = np.ascontiguousarray(ue, dtype=np.float64).reshape(-1, 24)
    n = Xe.shape[0]
    if epsp is None:
        epsp = np.zeros((n, 8, 6))
    if abar is None:
        abar = np.zeros((n, 8))
    if cpinv is None:
        cpinv = np.tile(np.array([1.0, 1, 1, 0, 0, 0]), (n, 8, 1))
    if beta is None:
        beta = np.zeros((n, 8, 6))
    for a in (epsp, beta, abar, cpinv):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise ValueError("history arrays must be C-contiguous float64")
    Fe = np.empty((n, 24))
    Ke = np.empty((n, 576))
    sig = np.empty((n, 8, 6))
    f = L.Fail()
    m = material._c()
    rc = L.lib().pfem_element_batch(_KIND[kind] if isinstance(kind, str) else int(kind), C.byref(m), n,
                                    L.ptr(Xe), L.ptr(ue), L.ptr(epsp), L.ptr(beta), L.ptr(abar), L.ptr(cpinv),
                                    int(commit),
                                    L.ptr(Fe), L.ptr(Ke), L.ptr(sig), C.byref(f))
    if rc == L.PFEM_ERR_INVERTED:
        raise ElementInversionError(f.e, f.g, f.J)
    L.check(rc, "element_batch")
    return Fe, Ke.reshape(n, 24, 24).transpose(0, 2, 1).copy(), sig
