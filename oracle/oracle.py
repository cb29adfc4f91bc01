"""ctypes binding of the C oracle and a Python Newton driver (TEST INFRASTRUCTURE ONLY).

Reference functions restated (lagrangian/src):
  Materials.jl:82-250      -> Material, return_map
  FiniteStrain.jl:148-547  -> finite_kinematics, finite_stress_update, dPdF
  Elements.jl:405-531      -> element_finite
  Mesh.jl:30-59            -> box_mesh (+ seeded jitter)
  Assembly.jl:62-416       -> build_csc, greedy_colors, OracleModel.assemble
  BoundaryConditions.jl:62-109 -> impose_dirichlet
  Solver.jl:159-475        -> forcing_term, OracleModel.newton / solve
Indices are 0-based.  Layouts follow the reference (see pfem_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

__all__ = [
    "lib", "Material", "return_map", "eig_sym3", "finite_kinematics",
    "finite_stress_update", "dPdF", "element_finite", "ref_grads", "box_mesh",
    "polar_RU", "dR_polar", "dtau_dbeta", "element_small",
    "jitter_u01", "build_csc", "greedy_colors", "colour_lists", "on_face",
    "impose_dirichlet", "forcing_term", "ElementInversionError", "OracleModel",
    "rigid_body_modes", "OracleAMG",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")


def _load():
    if not os.path.exists(_SO):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return C.CDLL(_SO)


lib = _load()

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class _Mat(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("E", "nu", "sy0", "Hiso", "Hkin", "ssat", "delta", "G", "K", "lam")] + \
               [("C", C.c_double * 36)]


class _Kin(C.Structure):
    _fields_ = [("ee", C.c_double * 6), ("b", C.c_double * 3), ("n", C.c_double * 9),
                ("Finv", C.c_double * 9), ("J", C.c_double), ("ok", C.c_int)]


lib.or_material_init.argtypes = [C.POINTER(_Mat)] + [C.c_double] * 7
lib.or_material_init.restype = C.c_int
lib.or_return_map.argtypes = [C.POINTER(_Mat), _dp, _dp, _dp, C.c_double, _dp, _dp, _dp,
                              C.POINTER(C.c_double), _dp]
lib.or_eig_sym3.argtypes = [_dp, _dp, _dp]
lib.or_finite_kinematics.argtypes = [_dp, _dp, C.POINTER(_Kin)]
lib.or_finite_stress_update.argtypes = [C.POINTER(_Mat), C.POINTER(_Kin), _dp, _dp, _dp, C.c_double,
                                        _dp, _dp, _dp, C.POINTER(C.c_double), _dp, C.c_void_p,
                                        _dp, _dp, _dp]
lib.or_polar_RU.argtypes = [_dp, _dp, _dp]
lib.or_dR_polar.argtypes = [_dp, _dp, _dp, _dp]
lib.or_dtau_dbeta.argtypes = [C.POINTER(_Mat), _dp, _dp, C.c_double, _dp]
lib.or_dPdF.argtypes = [C.POINTER(_Kin), _dp, _dp, _dp, _dp, C.c_void_p, C.c_void_p, C.c_void_p,
                        C.c_void_p, _dp]
lib.or_ref_grads.argtypes = [_dp, _dp, _dp, C.POINTER(C.c_double)]
lib.or_element_finite.argtypes = [C.c_int, C.POINTER(_Mat), _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                  C.c_int, _dp, _dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
lib.or_element_small.argtypes = [C.POINTER(_Mat), _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]
lib.or_element_finite.restype = C.c_int
lib.or_box_mesh.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [C.c_double, C.c_uint64, _dp, _ip]
lib.or_jitter_u01.argtypes = [C.c_uint64, C.c_int64, C.c_int]
lib.or_jitter_u01.restype = C.c_double
lib.or_build_csc.argtypes = [C.c_int64, C.c_int64, _ip, _ip, C.c_void_p, C.POINTER(C.c_int64)]
lib.or_greedy_colors.argtypes = [C.c_int64, C.c_int64, _ip, _i32p]
lib.or_greedy_colors.restype = C.c_int
lib.or_assemble.argtypes = [C.c_int, C.POINTER(_Mat), C.c_void_p, C.c_int64, C.c_int64, _dp, _ip, _ip, _ip,
                            C.c_int, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp,
                            C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                            C.POINTER(C.c_double)]
lib.or_assemble.restype = C.c_int
lib.or_csc_spmv_omp.argtypes = [C.c_int64, _ip, _ip, _dp, _dp, _dp, C.c_int]
lib.or_amg_setup.argtypes = [C.c_int64, _ip, _ip, _dp, _dp, np.ctypeslib.ndpointer(dtype=np.uint8),
                             C.c_int64, C.c_int]
lib.or_amg_setup.restype = C.c_void_p
lib.or_amg_free.argtypes = [C.c_void_p]
lib.or_amg_levels.argtypes = [C.c_void_p, _ip]
lib.or_amg_levels.restype = C.c_int
lib.or_amg_pcg.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int]
lib.or_amg_pcg.restype = C.c_int
lib.or_amg_vcycle.argtypes = [C.c_void_p, _dp, _dp]
lib.or_amg_apply_A.argtypes = [C.c_void_p, _dp, _dp]
lib.or_impose_dirichlet.argtypes = [C.c_int64, _ip, _ip, _dp, _dp, C.c_int64, _ip, _dp, C.c_int,
                                    C.c_double, _dp]
lib.or_gauss_stress.argtypes = [C.c_int, C.c_int64, _dp, _ip, _dp, _dp, _dp]


class ElementInversionError(Exception):
    """Mirror of FiniteStrain.ElementInversionError(e, g, J) (FiniteStrain.jl:42-52)."""

    def __init__(self, e, g, J):
        super().__init__(f"element {e} GP {g} inverted (J = {J})")
        self.e, self.g, self.J = e, g, J


class Material:
    """J2Material (Materials.jl:82-108)."""

    def __init__(self, E, nu, sy0, Hiso=0.0, Hkin=0.0, ssat=None, delta=0.0):
        self._m = _Mat()
        ssat = sy0 if ssat is None else ssat
        rc = lib.or_material_init(C.byref(self._m), E, nu, sy0, Hiso, Hkin, ssat, delta)
        if rc:
            raise ValueError(f"invalid J2 material (guard {rc})")

    def __getattr__(self, k):
        if k == "Cmat":
            return np.array(self._m.C[:]).reshape(6, 6)
        return getattr(self._m, k)

    @property
    def ref(self):
        return C.byref(self._m)


def _v(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
    if n is not None:
        assert a.size == n
    return a


def return_map(mat, eps, epn, bn, an):
    sig, ep, bt, D = (np.zeros(6), np.zeros(6), np.zeros(6), np.zeros(36))
    a = C.c_double()
    lib.or_return_map(mat.ref, _v(eps, 6), _v(epn, 6), _v(bn, 6), an, sig, ep, bt, C.byref(a), D)
    return sig, ep, bt, a.value, D.reshape(6, 6)


def eig_sym3(A):
    A = np.asarray(A, dtype=np.float64)
    w, V = np.zeros(3), np.zeros(9)
    lib.or_eig_sym3(_v(A.T), w, V)  # column-major in
    return w, V.reshape(3, 3).T


def _F9(F):
    return _v(np.asarray(F, dtype=np.float64).T)  # column-major 3x3


def finite_kinematics(F, cpi=(1, 1, 1, 0, 0, 0)):
    k = _Kin()
    lib.or_finite_kinematics(_F9(F), _v(cpi, 6), C.byref(k))
    return k


def finite_stress_update(mat, kin, epn, an, want_cpi=True, F=None, bref=None, full=False):
    """finite_stress_update (FiniteStrain.jl:292-346).  F (3x3) is needed when mat.Hkin > 0
    (polar rotation of the back stress); bref is the reference-configuration back stress.
    Returns (tau, ep, abar, D, cpi_new), or with full=True also (bref_new, beta_sp, R, U)."""
    tau, ep, D, cn = np.zeros(6), np.zeros(6), np.zeros(36), np.zeros(6)
    bnew, bsp, R, U = np.zeros(6), np.zeros(6), np.zeros(9), np.zeros(9)
    if F is None:
        assert mat.Hkin == 0, "F is required with kinematic hardening"
        F = np.eye(3)
    a = C.c_double()
    lib.or_finite_stress_update(mat.ref, C.byref(kin), _F9(F), _v(epn, 6),
                                _v(np.zeros(6) if bref is None else bref, 6), an, tau, ep, bnew,
                                C.byref(a), D, cn.ctypes.data if want_cpi else None, bsp, R, U)
    out = (tau, ep, a.value, D.reshape(6, 6), cn)
    if full:
        out += (bnew, bsp, R.reshape(3, 3, order="F"), U.reshape(3, 3, order="F"))
    return out


def polar_RU(F):
    R, U = np.zeros(9), np.zeros(9)
    lib.or_polar_RU(_F9(F), R, U)
    return R.reshape(3, 3, order="F"), U.reshape(3, 3, order="F")


def dR_polar(R, U, dF):
    dR = np.zeros(9)
    lib.or_dR_polar(_F9(R), _F9(U), _F9(dF), dR)
    return dR.reshape(3, 3, order="F")


def dtau_dbeta(mat, ee, bsp, an):
    out = np.zeros(36)
    lib.or_dtau_dbeta(mat.ref, _v(ee, 6), _v(bsp, 6), an, out)
    return out.reshape(6, 6)


def dPdF(kin, cpi, D, tau, F, bref=None, R=None, U=None, dtdb=None):
    A = np.zeros(81)
    kin_args = [None] * 4
    if bref is not None:
        kin_args = [_v(bref, 6), _F9(R), _F9(U), _v(np.asarray(dtdb).reshape(6, 6))]
    lib.or_dPdF(C.byref(kin), _v(cpi, 6), _v(np.asarray(D).reshape(6, 6)), _v(tau, 6), _F9(F),
                *[None if x is None else x.ctypes.data for x in kin_args], A)
    return A.reshape(9, 9, order="F")


def ref_grads(Xe, xi):
    dNdX = np.zeros(24)
    d = C.c_double()
    lib.or_ref_grads(_v(Xe, 24), _v(xi, 3), dNdX, C.byref(d))
    return dNdX.reshape(8, 3), d.value


def element_finite(fbar, mat, Xe, ue, epsp=None, abar=None, cpinv=None, commit=False, beta=None):
    """Fe (24), Ke (24x24), sig (8x6).  State arrays (8x6 / 8) updated in place on commit;
    beta is the reference-configuration back stress (zero when omitted)."""
    if epsp is None:
        epsp = np.zeros((8, 6))
    if abar is None:
        abar = np.zeros(8)
    if cpinv is None:
        cpinv = np.tile([1.0, 1, 1, 0, 0, 0], (8, 1))
    if beta is None:
        beta = np.zeros((8, 6))
    for arr in (epsp, beta, abar, cpinv):
        assert arr.dtype == np.float64 and arr.flags.c_contiguous
    sig = np.zeros((8, 6))
    Fe, Ke = np.zeros(24), np.zeros(576)
    g, J = C.c_int(), C.c_double()
    rc = lib.or_element_finite(int(fbar), mat.ref, _v(Xe, 24), _v(ue, 24), epsp, beta, abar, cpinv, sig,
                               int(commit), Fe, Ke, C.byref(g), C.byref(J))
    if rc:
        raise ElementInversionError(0, g.value, J.value)
    return Fe, Ke.reshape(24, 24, order="F"), sig


def element_small(mat, Xe, ue, epsp=None, beta=None, abar=None, commit=False):
    """Small-strain Hex8 (Elements.jl:323-360): Fe (24), Ke (24x24), sig (8x6)."""
    epsp = np.zeros((8, 6)) if epsp is None else epsp
    beta = np.zeros((8, 6)) if beta is None else beta
    abar = np.zeros(8) if abar is None else abar
    for arr in (epsp, beta, abar):
        assert arr.dtype == np.float64 and arr.flags.c_contiguous
    sig = np.zeros((8, 6))
    Fe, Ke = np.zeros(24), np.zeros(576)
    lib.or_element_small(mat.ref, _v(Xe, 24), _v(ue, 24), epsp, beta, abar, sig, int(commit), Fe, Ke)
    return Fe, Ke.reshape(24, 24, order="F"), sig


# ---------------------------------------------------------------- mesh
def box_mesh(lx, ly, lz, nx, ny, nz, jitter=0.0, seed=0):
    """nodes (nnodes, 3), conn (nelem, 8) 0-based (Mesh.jl:30-59)."""
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    ne = nx * ny * nz
    nodes = np.zeros((nn, 3))
    conn = np.zeros((ne, 8), dtype=np.int64)
    lib.or_box_mesh(lx, ly, lz, nx, ny, nz, jitter, seed, nodes, conn)
    return nodes, conn


def jitter_u01(seed, node, comp):
    return lib.or_jitter_u01(seed, node, comp)


def build_csc(nnodes, conn):
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    colptr = np.zeros(3 * nnodes + 1, dtype=np.int64)
    nnz = C.c_int64()
    lib.or_build_csc(nnodes, conn.shape[0], conn, colptr, None, C.byref(nnz))
    rowval = np.zeros(nnz.value, dtype=np.int64)
    lib.or_build_csc(nnodes, conn.shape[0], conn, colptr, rowval.ctypes.data, C.byref(nnz))
    return colptr, rowval


def greedy_colors(nnodes, conn):
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    col = np.zeros(conn.shape[0], dtype=np.int32)
    nc = lib.or_greedy_colors(nnodes, conn.shape[0], conn, col)
    if nc < 0:
        raise RuntimeError("more than 64 colours")
    return col, nc


def colour_lists(col, nc):
    """colour_ptr, colour_order: element ids colour by colour, ascending within."""
    order = np.argsort(col, kind="stable").astype(np.int64)
    ptr = np.zeros(nc + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(np.bincount(col, minlength=nc))
    return ptr, order


def on_face(nodes, face):
    """on_face (Mesh.jl:82-104), 0-based node ids."""
    lo, hi = nodes.min(axis=0), nodes.max(axis=0)
    L = float((hi - lo).max())
    tol = 1e-8 * max(L, 1.0)
    ax = "xyz".index(face[0])
    ref = lo[ax] if face[1:] == "min" else hi[ax]
    return np.nonzero(np.abs(nodes[:, ax] - ref) <= tol)[0].astype(np.int64)


def impose_dirichlet(K, R, dofs, values, ramp, lam, U):
    """In place on scipy CSC K (int64 indices) and R."""
    dofs = np.ascontiguousarray(dofs, dtype=np.int64)
    lib.or_impose_dirichlet(K.shape[0], K.indptr, K.indices, K.data, R, len(dofs), dofs,
                            _v(values), int(ramp), lam, _v(U))


def forcing_term(it, rk, rkm1, eta_prev, eta_min=1e-8, eta_max=0.1, gamma=0.9, alpha=1.5):
    """_forcing_term, Eisenstat-Walker choice 2 (Solver.jl:159-171); it is 1-based."""
    if it == 1 or rkm1 <= 0:
        return eta_max
    eta = gamma * (rk / rkm1) ** alpha
    safe = gamma * eta_prev ** alpha
    if safe > 0.1:
        eta = max(eta, safe)
    return min(max(eta, eta_min), eta_max)


class OracleAMG:
    """PCG + smoothed-aggregation AMG (rigid-body near-null space, damped Jacobi smoother)
    on the oracle's CSC K -- the reference's CG+AMG linear solve (Solver.jl:235-290)."""

    def __init__(self, K, nodes, isbc, coarse_max=1000, nthreads=0):
        self.K = K
        self.n = K.shape[0]
        self._h = lib.or_amg_setup(self.n, K.indptr, K.indices, K.data,
                                   np.ascontiguousarray(nodes, dtype=np.float64),
                                   np.ascontiguousarray(isbc, dtype=np.uint8), coarse_max, nthreads)

    def levels(self):
        sz = np.zeros(16, dtype=np.int64)
        n = lib.or_amg_levels(self._h, sz)
        return list(sz[:n])

    def solve(self, b, rtol, itmax=200):
        """-> (x, iterations); iterations < 0 flags a breakdown or no convergence."""
        x = np.zeros(self.n)
        it = lib.or_amg_pcg(self._h, _v(b), x, rtol, itmax)
        return x, it

    def apply_A(self, x):
        y = np.zeros(self.n)
        lib.or_amg_apply_A(self._h, _v(x), y)
        return y

    def vcycle(self, b):
        x = np.zeros(self.n)
        lib.or_amg_vcycle(self._h, _v(b), x)
        return x

    def __del__(self):
        if getattr(self, "_h", None):
            lib.or_amg_free(self._h)
            self._h = None


def rigid_body_modes(nodes):
    """_rigid_body_modes (Solver.jl:117-129): ndof x 6."""
    nn = nodes.shape[0]
    B = np.zeros((3 * nn, 6))
    x, y, z = nodes[:, 0], nodes[:, 1], nodes[:, 2]
    B[0::3, 0] = 1.0
    B[1::3, 1] = 1.0
    B[2::3, 2] = 1.0
    B[1::3, 3] = -z
    B[2::3, 3] = y
    B[0::3, 4] = z
    B[2::3, 4] = -x
    B[0::3, 5] = -y
    B[1::3, 5] = x
    return B


class OracleModel:
    """Model + solve!/newton! restated on the oracle (Model.jl:63-162, Solver.jl:317-475).

    The linear solve is a direct sparse solve (scipy) -- the reference's F-bar
    default (Solver.jl:237-240).
    """

    def __init__(self, nodes, conn, mat, element="finite_fbar", nthreads=0, material_id=None):
        import scipy.sparse as sp
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.conn = np.ascontiguousarray(conn, dtype=np.int64)
        mats = list(mat) if isinstance(mat, (list, tuple)) else [mat]
        self.mat = mats[0]
        self._mats = (_Mat * len(mats))(*[m._m for m in mats])
        self.matid = None if material_id is None else np.ascontiguousarray(material_id, dtype=np.int32)
        if self.matid is not None:
            assert self.matid.shape == (self.conn.shape[0],) and 0 <= self.matid.min() and \
                self.matid.max() < len(mats)
        self.kind = {"finite": 0, "finite_fbar": 1, "small": 2}[element]
        self.nn, self.ne = self.nodes.shape[0], self.conn.shape[0]
        self.ndof = 3 * self.nn
        self.colptr, self.rowval = build_csc(self.nn, self.conn)
        self.K = sp.csc_matrix((np.zeros(self.rowval.size), self.rowval, self.colptr),
                               shape=(self.ndof, self.ndof))
        self.K.indices = self.rowval
        self.K.indptr = self.colptr
        col, nc = greedy_colors(self.nn, self.conn)
        self.colors, self.ncolors = col, nc
        self.cptr, self.corder = colour_lists(col, nc)
        self.nthreads = nthreads
        self.U = np.zeros(self.ndof)
        self.U_c = np.zeros(self.ndof)   # U at the last commit (restored after a failed increment)
        self.R = np.zeros(self.ndof)
        ngp = 8 * self.ne
        self.st_c = self._new_state(ngp)
        self.dir_dofs, self.dir_vals, self.dir_ramp = [], [], []
        self.neu_dofs, self.neu_vals = [], []

    @staticmethod
    def _new_state(ngp):
        cp = np.zeros((ngp, 6))
        cp[:, :3] = 1.0
        return {"epsp": np.zeros((ngp, 6)), "beta": np.zeros((ngp, 6)), "abar": np.zeros(ngp),
                "cpinv": cp, "sig": np.zeros((ngp, 6))}

    # BC builders (Model.jl:102-145)
    def fix(self, nodes, comp="all"):
        for n in nodes:
            for c in ((0, 1, 2) if comp == "all" else ("xyz".index(comp),)):
                self.dir_dofs.append(3 * int(n) + c)
                self.dir_vals.append(0.0)
                self.dir_ramp.append(False)

    def prescribe(self, nodes, comp, value, ramp=True):
        for n in nodes:
            for c in ((0, 1, 2) if comp == "all" else ("xyz".index(comp),)):
                self.dir_dofs.append(3 * int(n) + c)
                self.dir_vals.append(float(value))
                self.dir_ramp.append(bool(ramp))

    def load(self, nodes, comp, value, distribute=False):
        per = value / len(nodes) if distribute else value
        for n in nodes:
            for c in ((0, 1, 2) if comp == "all" else ("xyz".index(comp),)):
                self.neu_dofs.append(3 * int(n) + c)
                self.neu_vals.append(float(per))

    def constrain(self, dofs, values, ramp):
        self.dir_dofs += [int(d) for d in dofs]
        self.dir_vals += [float(v) for v in values]
        self.dir_ramp += [bool(r) for r in ramp]

    def freeze_bcs(self):
        """_freeze_bcs (Solver.jl:192-218): dedup, last wins, split ramp / fixed."""
        last = {}
        for d, v, r in zip(self.dir_dofs, self.dir_vals, self.dir_ramp):
            last[d] = (v, r)
        rd = [(d, v) for d, (v, r) in last.items() if r]
        fd = [(d, v) for d, (v, r) in last.items() if not r]
        mk = lambda L: (np.array([d for d, _ in L], dtype=np.int64),  # noqa: E731
                        np.array([v for _, v in L], dtype=np.float64))
        return mk(rd), mk(fd)

    def assemble(self, U, state, commit=False, nthreads=None):
        """assemble_threaded! (Assembly.jl:340-416): K.data and R overwritten."""
        fe, fg, fJ = C.c_int64(), C.c_int(), C.c_double()
        rc = lib.or_assemble(self.kind, self._mats, None if self.matid is None else self.matid.ctypes.data,
                             self.nn, self.ne, self.nodes, self.conn,
                             self.colptr, self.rowval, self.ncolors, self.cptr, self.corder,
                             _v(U), state["epsp"], state["beta"], state["abar"], state["cpinv"], state["sig"],
                             int(commit), self.K.data, self.R,
                             self.nthreads if nthreads is None else nthreads,
                             C.byref(fe), C.byref(fg), C.byref(fJ))
        if rc:
            raise ElementInversionError(fe.value, fg.value, fJ.value)
        return self.K, self.R

    def newton(self, lam, Fext, bcs, tol=1e-8, maxiter=25, solver=None, diverge=0.0):
        """newton! (Solver.jl:317-389).  `solver(K, b, eta)` -> (dU, niter).  diverge > 0 stops
        the step (not converged) when, from the third iteration on, ||R|| > diverge times the
        smallest earlier ||R|| (the device driver's optional early stop; the reference runs
        maxiter iterations)."""
        import scipy.sparse.linalg as spla
        if solver is None:
            solver = lambda K, b, eta: (spla.spsolve(K.tocsc(), b), 0)  # noqa: E731
        (rd, rv), (fd, fv) = bcs
        hist, cg = [], []
        ref = 1.0
        fext_norm = float(np.sqrt(np.sum(Fext ** 2)))
        rkm1, eta_prev, conv = 0.0, 0.1, False
        for it in range(1, maxiter + 1):
            trial = {k: v.copy() for k, v in self.st_c.items()}      # :344
            K, R = self.assemble(self.U, trial, commit=False)         # :347
            R -= Fext                                                 # :349
            impose_dirichlet(K, R, rd, rv, True, lam, self.U)         # :354
            impose_dirichlet(K, R, fd, fv, False, lam, self.U)        # :355
            rnorm = float(np.sqrt(np.sum(R ** 2)))
            hist.append(rnorm)
            if it == 1:
                ref = max(rnorm, fext_norm, 1.0)
            if rnorm <= tol * ref:
                conv = True
                break
            if diverge > 0 and it >= 3 and not rnorm <= diverge * min(hist[:-1]):
                break
            eta = forcing_term(it, rnorm, rkm1, eta_prev)
            dU, n = solver(K, -R, eta)                                # :378-379 (also at maxiter)
            cg.append(n)
            self.U += dU                                              # :383
            rkm1, eta_prev = rnorm, eta
        return conv, hist, cg

    def commit(self):
        """solve!'s commit (Solver.jl:470-471)."""
        trial = {k: v.copy() for k, v in self.st_c.items()}
        self.assemble(self.U, trial, commit=True)                     # :470
        self.st_c = trial                                              # :471
        self.U_c = self.U.copy()

    def gauss_stress(self):
        """gauss_stress (Model.jl:188-208): committed tau / det F (un-barred F at the current U)."""
        out = np.zeros((8 * self.ne, 6))
        lib.or_gauss_stress(self.kind, self.ne, self.nodes, self.conn, _v(self.U),
                            np.ascontiguousarray(self.st_c["sig"]), out)
        return out

    def solve(self, nsteps=10, tol=1e-8, maxiter=25, solver=None, callback=None, cutback=0, grow_after=2,
              grow_iters=6, diverge=0.0, steps=None, records=None):
        """solve! (Solver.jl:419-475) with the direct linear solve.

        cutback > 0: load-step bisection (lagrangian/docs/SCALING.md:268-274).  A failed
        increment (Newton not converged, or an inverted element) is undone (U back to the
        committed U; the committed history is untouched) and halved, at most `cutback`
        times below 1/nsteps; after `grow_after` converged increments in a row that each took
        at most `grow_iters` Newton iterations it doubles again, never above 1/nsteps, and an
        increment that would pass the load step's end n/nsteps is cut to end there.  Every
        failure doubles the successes needed before the next doubling (at most
        8 grow_after); a converged increment of the full nominal size resets them.  `records` (a list) receives one
        (lam, converged, newton iterations) tuple per attempt."""
        self.U[:] = 0.0
        self.U_c = self.U.copy()
        self.st_c = self._new_state(8 * self.ne)
        bcs = self.freeze_bcs()
        Fext0 = np.zeros(self.ndof)
        for d, v in zip(self.neu_dofs, self.neu_vals):
            Fext0[d] += v
        iters, res, cgs, allconv = [], [], [], True
        nominal = 1.0 / nsteps
        dl, streak, lam_c, need = nominal, 0, 0.0, grow_after
        dl_min = nominal / 2.0 ** cutback
        for n in (range(1, nsteps + 1) if steps is None else steps):
            target = n / nsteps
            ok = True
            while True:
                lam = target if lam_c + dl >= target - 1e-12 * nominal else lam_c + dl
                Fext = np.zeros(self.ndof)
                for d, v in zip(self.neu_dofs, self.neu_vals):
                    Fext[d] += lam * v                                 # assemble_neumann! :44-49
                try:
                    conv, hist, cg = self.newton(lam, Fext, bcs, tol=tol, maxiter=maxiter, solver=solver,
                                                 diverge=diverge)
                except ElementInversionError:
                    conv, hist, cg = False, [], []
                if records is not None:
                    records.append((lam, conv, len(hist)))
                if conv:
                    iters.append(len(hist))
                    res.append(hist)
                    cgs.append(cg)
                    self.commit()
                    if lam - lam_c >= nominal * (1.0 - 1e-9):
                        need = grow_after
                    lam_c = lam
                    streak = streak + 1 if len(hist) <= grow_iters else 0
                    if streak >= need and dl < nominal:
                        dl, streak = min(2.0 * dl, nominal), 0
                    if lam == target:
                        break
                    continue
                half = 0.5 * (lam - lam_c)
                if cutback <= 0 or half < dl_min * (1.0 - 1e-9):
                    if cutback > 0:
                        self.U = self.U_c.copy()
                    iters.append(len(hist))
                    res.append(hist)
                    cgs.append(cg)
                    ok = False
                    break
                self.U = self.U_c.copy()
                dl, streak, need = half, 0, min(2 * need, 8 * grow_after)
            if not ok:
                allconv = False
                break
            if callback is not None:
                callback(n, self)
        return allconv, iters, res, cgs
