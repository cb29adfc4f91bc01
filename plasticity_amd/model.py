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
import sys
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


_OPTION_ALIASES = {
    "fcg_directions": "gcr_directions",   # directions kept by the learned preconditioner's flexible GCR
    "line_search": "line_search",
    "wcycle": "amg_wcycle",
}


def default_options(**kw):
    """pfem_default_options with fields overridden by keyword (names of pfem_options), plus:
    rho_iters=(cold, warm); gcr_fallback=True/False (flexible GCR or BiCGStab after a CG
    breakdown); stream_prolongation=True/False; precision / smoother / amg_rebuild by name."""
    o = L.Options()
    L.lib().pfem_default_options(C.byref(o))
    for k, v in kw.items():
        k = _OPTION_ALIASES.get(k, k)
        if k == "rho_iters":
            o.rho_iters_cold, o.rho_iters_warm = int(v[0]), int(v[1])
            continue
        if k == "gcr_fallback":
            o.krylov_fallback = 0 if v else 1
            continue
        if k == "stream_prolongation":
            o.prolong_rowwise = 0 if v else 1
            continue
        if k == "precision" and isinstance(v, str):
            v = {"fp64": L.PREC_FP64, "fp32_ir": L.PREC_FP32_IR}[v]
        if k == "smoother" and isinstance(v, str):
            v = {"jacobi": L.SMOOTH_JACOBI, "chebyshev": L.SMOOTH_CHEBYSHEV, "chebyshev4": L.SMOOTH_CHEBYSHEV4}[v]
        if k == "amg_rebuild" and isinstance(v, str):
            v = {"per_load_step": L.AMG_PER_LOAD_STEP, "every_iteration": L.AMG_EVERY_ITERATION,
                 "keep_p": L.AMG_KEEP_P, "keep_p_until_fallback": L.AMG_KEEP_P_UNTIL_FALLBACK,
                 "reuse_coarse": L.AMG_REUSE_COARSE}[v]
        if isinstance(v, bool):
            v = int(v)
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k!r}")
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
        # committed load factor, increment, fast successes in a row, successes needed to grow
        self._lam_c, self._dl, self._streak, self._need = 0.0, None, 0, None

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
        self._lam_c, self._dl, self._streak, self._need = 0.0, None, 0, None
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
        the committed stress sig (Kirchhoff for finite strain, Cauchy for small)."""
        ep, bt, ab, cp, sg = (np.empty((self.ngp, 6)), np.empty((self.ngp, 6)), np.empty(self.ngp),
                              np.empty((self.ngp, 6)), np.empty((self.ngp, 6)))
        L.check(L.lib().pfem_get_state(self._h, L.ptr(ep), L.ptr(bt), L.ptr(ab), L.ptr(cp), L.ptr(sg)),
                "get_state")
        return {"epsp": ep, "beta": bt, "abar": ab, "cpinv": cp, "sig": sg}

    def set_state(self, epsp=None, abar=None, cpinv=None, beta=None):
        args = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (epsp, beta, abar, cpinv)]
        L.check(L.lib().pfem_set_state(self._h, *[None if a is None else L.ptr(a) for a in args]), "set_state")

    def trial_stress(self):
        """Stress of the last assembly (the trial state's sigma written by assemble!), ngp x 6."""
        out = np.empty((self.ngp, 6))
        L.check(L.lib().pfem_get_trial_stress(self._h, L.ptr(out)), "get_trial_stress")
        return out

    def nodal_displacements(self):
        """3 x nnodes (Model.jl:169-178)."""
        return self.U.reshape(-1, 3).T.copy()

    def gauss_kirchhoff(self):
        """Committed Kirchhoff stress tau, ngp x 6 (Model.jl:226)."""
        return self.get_state()["sig"]

    def equivalent_plastic_strain(self):
        """Committed abar (Model.jl:231)."""
        return self.get_state()["abar"]

    def gauss_stress(self):
        """Committed Cauchy stress, ngp x 6 (Model.jl:188-208), computed on the device: the
        stored stress for the small-strain element, tau / det F with the un-barred F at the
        current U for the finite-strain elements."""
        out = np.empty((self.ngp, 6))
        L.check(L.lib().pfem_gauss_stress(self._h, L.ptr(out)), "gauss_stress")
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

    def impose_dirichlet(self, lam, which=None, U=None, R=None):
        """impose_dirichlet! on the device K and R.  which=None: both frozen sets (the two
        calls of newton!, Solver.jl:354-355); "ramp" / "fixed": one call for that set.  A host
        R (ndof array) is sent in and overwritten with the result; U (ndof) replaces the
        device U first."""
        if which is None:
            L.check(L.lib().pfem_impose_dirichlet(self._h, float(lam)), "impose_dirichlet")
            return R
        Ua = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        if R is not None:
            assert R.dtype == np.float64 and R.flags.c_contiguous
        st = {"ramp": L.BC_RAMPED, "fixed": L.BC_FIXED}[which]
        L.check(L.lib().pfem_impose_dirichlet_set(self._h, st, float(lam), None if Ua is None else L.ptr(Ua),
                                                  None if R is None else L.ptr(R)), "impose_dirichlet_set")
        return R

    def subtract_external(self, lam):
        L.check(L.lib().pfem_subtract_external(self._h, float(lam)), "subtract_external")

    def residual_norm(self):
        out = C.c_double()
        L.check(L.lib().pfem_residual_norm(self._h, C.byref(out)), "residual_norm")
        return out.value

    def linear_solve(self, rtol, rebuild_pc=True, strict=False, b=None):
        """_linear_solve!(dU, K, b, ...) -> (dU, niter, solver_used, true_relres); b=None
        solves against -R on the device.  Statuses 2 (not converged) and 3 (breakdown) keep
        the correction the fallbacks produced, as the reference does, unless strict."""
        it, used, rel = C.c_int(), C.c_int(), C.c_double()
        dU = np.empty(self.ndof)
        ba = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        rc = L.lib().pfem_linear_solve(self._h, None if ba is None else L.ptr(ba), float(rtol), int(rebuild_pc),
                                       C.byref(it), C.byref(used), C.byref(rel), L.ptr(dU))
        if rc in (L.PFEM_ERR_NOT_CONVERGED, L.PFEM_ERR_NON_SPD) and not strict:
            warnings.warn(f"linear solve: {L.lib().pfem_status_string(rc).decode()}")
        else:
            L.check(rc, "linear_solve")
        return dU, it.value, used.value, rel.value

    def update_displacement(self):
        L.check(L.lib().pfem_update_displacement(self._h), "update_displacement")

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
            err = ElementInversionError(f.e, f.g, f.J)
            err.result = res  # the iterations done before the inversion
            raise err
        L.check(rc, "newton")
        n = res.iters
        hist = list(res.residuals[:n])
        kry = list(res.krylov_iters[:res.solves])
        return bool(res.converged), hist, kry, res

    def commit(self):
        f = L.Fail()
        rc = L.lib().pfem_commit(self._h, C.byref(f))
        if rc == L.PFEM_ERR_INVERTED:
            raise ElementInversionError(f.e, f.g, f.J)
        L.check(rc, "commit")

    def rollback(self):
        """U back to its value at the last commit (the committed history is never changed by a
        trial assembly)."""
        L.check(L.lib().pfem_rollback(self._h), "rollback")

    def advance(self, target, nsteps, tol=1e-8, maxiter=25, cutback=0, grow_after=2, grow_iters=6,
                verbose=False):
        """Advance the committed load factor to `target` (a point n/nsteps of the load path)
        with newton! + commit, bisecting the increment on failure (lagrangian/docs/SCALING.md
        :268-274; the reference's solve! only warns and stops, Solver.jl:463-467).

        An increment whose Newton loop does not converge, or inverts an element, is undone
        (pfem_rollback) and halved -- at most `cutback` halvings below the nominal increment
        1/nsteps.  After `grow_after` converged increments in a row that each needed at most
        `grow_iters` Newton iterations the increment doubles again, never above 1/nsteps; an
        increment that would pass `target` is cut to end on it.  Every failure doubles the
        number of such increments needed before the next doubling (at most 8 grow_after),
        and a converged increment of the full nominal size brings it back to grow_after, so
        an increment size that keeps failing is not retried at every load step.
        Returns the list of attempted increments, each a dict (lam, converged, inverted,
        newton iterations, residuals, krylov, times); raises RuntimeError when the increment
        would have to fall below its limit (or at the first failure when cutback = 0)."""
        nominal = 1.0 / nsteps
        if self._dl is None:
            self._dl = nominal
        if self._need is None:
            self._need = grow_after
        dl_min = nominal / 2.0 ** cutback
        out = []
        while True:
            lam_c = self._lam_c
            lam = target if lam_c + self._dl >= target - 1e-12 * nominal else lam_c + self._dl
            rec = {"lam": lam, "converged": False, "inverted": False}
            try:
                conv, hist, kry, res = self.newton(lam, tol=tol, maxiter=maxiter)
                rec.update(converged=conv, iters=len(hist), residuals=hist, krylov=kry,
                           solver_used=list(res.solver_used[:res.solves]), stopped=res.stopped,
                           times={"assemble": res.t_assemble, "bc": res.t_bc, "setup": res.t_setup,
                                  "solve": res.t_solve, "total": res.t_total})
            except ElementInversionError as e:
                r = getattr(e, "result", None)
                n = r.iters if r is not None else 0
                rec.update(inverted=True, error=str(e), iters=n,
                           residuals=list(r.residuals[:n]) if r is not None else [],
                           krylov=list(r.krylov_iters[:r.solves]) if r is not None else [], solver_used=[],
                           times={"assemble": r.t_assemble, "bc": r.t_bc, "setup": r.t_setup,
                                  "solve": r.t_solve, "total": r.t_total} if r is not None else {})
            out.append(rec)
            if verbose:
                print(f"  lambda {lam:.6f} (+{lam - lam_c:.3e}): {'ok' if rec['converged'] else 'failed'} "
                      f"newton={rec['iters']} krylov={rec['krylov']}" + (" inverted" if rec["inverted"] else ""),
                      file=sys.stderr, flush=True)
            if rec["converged"]:
                self.commit()
                self._lam_c = lam
                if lam - lam_c >= nominal * (1.0 - 1e-9):
                    self._need = grow_after
                self._streak = self._streak + 1 if rec["iters"] <= grow_iters else 0
                if self._streak >= self._need and self._dl < nominal:
                    self._dl = min(2.0 * self._dl, nominal)
                    self._streak = 0
                if lam == target:
                    return out
                continue
            half = 0.5 * (lam - lam_c)
            if cutback <= 0 or half < dl_min * (1.0 - 1e-9):
                if cutback > 0:  # with cutback = 0, U stays where newton! left it (Solver.jl:463-467)
                    self.rollback()
                err = RuntimeError(f"load step to lambda={target:.6f} failed at lambda={lam:.6f} "
                                   f"(increment {lam - lam_c:.3e}, cut-back limit {dl_min:.3e})")
                err.records = out
                raise err
            self.rollback()
            self._dl = half
            self._streak = 0
            self._need = min(2 * self._need, 8 * grow_after)

    def solve(self, nsteps=10, tol=1e-8, maxiter=25, verbose=False, callback=None, steps=None, cutback=0,
              grow_after=2, grow_iters=6):
        """solve! (Solver.jl:419-475): reset, freeze BCs, ramp lambda = n/nsteps, commit after
        each converged load step.  cutback > 0 bisects failed increments (advance); with the
        default 0 the first failed load step ends the solve, as in the reference."""
        self.reset()
        self._freeze()
        out = SolveResult(converged=True)
        out.increments = []
        out.cutbacks = 0
        for n in (range(1, nsteps + 1) if steps is None else steps):
            lam = n / nsteps
            try:
                recs = self.advance(lam, nsteps, tol=tol, maxiter=maxiter, cutback=cutback, grow_after=grow_after,
                                    grow_iters=grow_iters)
                ok = True
            except RuntimeError as e:
                recs, ok = getattr(e, "records", []), False
                if cutback <= 0:
                    warnings.warn(f"load step {n} did not converge in {maxiter} iterations")
            out.increments += recs
            out.cutbacks += sum(1 for r in recs if not r["converged"])
            last = [r for r in recs if r["converged"]]
            for r in last:
                out.iters.append(r["iters"])
                out.residuals.append(r["residuals"])
                out.cg_iters.append(r["krylov"])
                out.solver_used.append(r["solver_used"])
                out.times.append(r["times"])
            if verbose:
                print(f"load step {n}/{nsteps} (lambda={lam:.4f}) increments={len(recs)} "
                      f"newton={[r['iters'] for r in recs]}")
            if not ok:
                out.converged = False
                break
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

    def apply_K(self, x, prec=0):
        """y = K x with the current device K: prec 0 FP64, 1 the FP32 copy, 2 the FP32 +
        BF16 split copy (FP64 arithmetic)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.ndof,):
            raise ValueError(f"x must have {self.ndof} entries")
        y = np.empty(self.ndof)
        L.check(L.lib().pfem_apply_K(self._h, int(prec), x.ctypes.data_as(L._dp), y.ctypes.data_as(L._dp)),
                "apply_K")
        return y

    def bench_spmv(self, prec=0, reps=20):
        sec = C.c_double()
        L.check(L.lib().pfem_bench_spmv(self._h, int(prec), int(reps), C.byref(sec)), "bench_spmv")
        return sec.value

    def bench_vcycle(self, prec=0, reps=20):
        """Seconds and algorithmic bytes of one AMG cycle and of one fused fine-level
        Chebyshev step (prec 0 FP64, 1 FP32) on the current hierarchy."""
        v = [C.c_double() for _ in range(4)]
        L.check(L.lib().pfem_bench_vcycle(self._h, int(prec), int(reps), *[C.byref(x) for x in v]), "bench_vcycle")
        return {"cycle_s": v[0].value, "cycle_bytes": v[1].value, "step_s": v[2].value, "step_bytes": v[3].value}


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
