"""Error behaviour of the host side, without a GPU: the J2Material validity guards of the
reference (Materials.jl:90-103: E > 0, -1 < nu < 0.5, sy0 > 0, Hiso >= 0, Hkin >= 0,
ssat >= sy0, delta >= 0, and no saturation together with kinematic hardening) in the
Python mirror (ValueError, as the reference throws ArgumentError) and in the library
(PFEM_ERR_BAD_ARG from pfem_element_batch, which checks the material before it looks
for a device), and the argument checks every entry point makes before any HIP call."""
import ctypes as C

import numpy as np
import pytest

import plasticity_amd as pa
import plasticity_amd._lib as L

VALID = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9, Hkin=0.0, ssat=250e6, delta=0.0)
INVALID = [
    dict(E=0.0), dict(E=-1.0), dict(nu=0.5), dict(nu=-1.0), dict(nu=0.7), dict(sy0=0.0), dict(sy0=-5.0),
    dict(Hiso=-1.0), dict(Hkin=-1.0), dict(ssat=200e6), dict(delta=-0.1),
    dict(ssat=300e6, delta=10.0, Hkin=1e9),  # saturation with kinematic hardening
]


def _batch(mat):
    """pfem_element_batch on one undeformed unit cube element with material `mat`."""
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                 dtype=np.float64).ravel()
    u = np.zeros(24)
    epsp, beta, cp, sig = (np.zeros(48) for _ in range(4))
    cp.reshape(8, 6)[:, :3] = 1.0
    abar, Fe, Ke = np.zeros(8), np.zeros(24), np.zeros(576)
    m = L.Material(**mat)
    f = L.Fail()
    return L.lib().pfem_element_batch(L.HEX8_FINITE_FBAR, C.byref(m), 1, L.ptr(X), L.ptr(u), L.ptr(epsp),
                                      L.ptr(beta), L.ptr(abar), L.ptr(cp), 0, L.ptr(Fe), L.ptr(Ke), L.ptr(sig),
                                      C.byref(f))


@pytest.mark.parametrize("bad", INVALID)
def test_material_guards_python_and_library(bad):
    kw = dict(VALID, **bad)
    with pytest.raises(ValueError):
        pa.J2Material(**kw)
    assert _batch(kw) == L.PFEM_ERR_BAD_ARG


def test_valid_material_passes_the_guards():
    pa.J2Material(**VALID)
    pa.J2Material(**dict(VALID, nu=-0.99, Hkin=5e8))            # auxetic, kinematic
    pa.J2Material(**dict(VALID, ssat=400e6, delta=16.0))        # Voce saturation
    # the library gets past the material check: no device here (or the call succeeds on one)
    assert _batch(VALID) in (L.PFEM_OK, L.PFEM_ERR_NO_DEVICE)


def test_entry_points_reject_bad_arguments_before_touching_a_device():
    lib = L.lib()
    BAD = L.PFEM_ERR_BAD_ARG
    out = C.c_void_p()
    # pfem_create: missing descriptor, bad element kind, no materials, empty mesh
    assert lib.pfem_create(None, None, C.byref(out)) == BAD
    nodes = np.zeros(24)
    conn = np.arange(8, dtype=np.int64)
    mat = L.Material(**VALID)
    for kw in (dict(element_kind=7), dict(nmaterials=0), dict(nnodes=0), dict(nelem=0)):
        d = dict(nnodes=8, nelem=1, nodes=L.ptr(nodes), conn=L.ptr(conn), material_id=None, nmaterials=1,
                 materials=C.cast(C.pointer(mat), C.c_void_p), element_kind=L.HEX8_FINITE_FBAR)
        d.update(kw)
        md = L.MeshDesc(**d)
        assert lib.pfem_create(C.byref(md), None, C.byref(out)) == BAD, kw
    # calls on a missing context
    res, f = L.NewtonResult(), L.Fail()
    assert lib.pfem_newton(None, 1.0, 1e-8, 10, C.byref(res), C.byref(f)) == BAD
    assert lib.pfem_assemble(None, None, 0, None, C.byref(f)) == BAD
    assert lib.pfem_reset(None) == BAD
    assert lib.pfem_commit(None, C.byref(f)) == BAD
    ni, nu_, rr = C.c_int(), C.c_int(), C.c_double()
    assert lib.pfem_linear_solve(None, None, 1e-8, 1, C.byref(ni), C.byref(nu_), C.byref(rr), None) == BAD
    assert lib.pfem_impose_dirichlet_set(None, 0, 1.0, None, None) == BAD
    assert lib.pfem_rollback(None) == BAD
    assert lib.pfem_get_trial_stress(None, None) == BAD
    assert lib.pfem_gauss_stress(None, None) == BAD
    assert lib.pfem_set_dirichlet(None, 0, None, None, None) == BAD
    assert lib.pfem_attach_net(None, None, 0, 1) == BAD
    d4 = [C.c_double() for _ in range(4)]
    assert lib.pfem_bench_vcycle(None, 0, 5, *[C.byref(v) for v in d4]) == BAD
    assert lib.pfem_bench_spmv(None, 0, 5, C.byref(d4[0])) == BAD
    xp = nodes.ctypes.data_as(L._dp)
    assert lib.pfem_apply_K(None, 0, xp, xp) == BAD
    assert lib.pfem_net_set_loss(None, 1) == BAD and lib.pfem_net_set_conv_kernel(None, 1) == BAD
    # stand-alone entry points
    assert lib.pfem_box_mesh(1.0, 1.0, 1.0, 0, 2, 2, 0.1, 0, L.ptr(nodes), L.ptr(conn)) == BAD
    assert lib.pfem_box_mesh(1.0, 1.0, 1.0, 2, 2, 2, -0.1, 0, L.ptr(nodes), L.ptr(conn)) == BAD
    assert lib.pfem_dense_inverse(0, L.ptr(nodes), L.ptr(nodes)) == BAD
    assert _batch_kind(9) == BAD
    net = C.c_void_p()
    assert lib.pfem_net_create(None, C.byref(net)) == BAD
    nd = L.NetDesc(nx=8, ny=8, nz=8, capacity=0, seed=0)
    nd.channels[:] = [16, 32, 48, 64]  # 48 is not a supported width
    assert lib.pfem_net_create(C.byref(nd), C.byref(net)) == BAD
    for code in range(8):
        assert lib.pfem_status_string(code)


def _batch_kind(kind):
    z = np.zeros(576)
    m = L.Material(**VALID)
    f = L.Fail()
    return L.lib().pfem_element_batch(kind, C.byref(m), 1, L.ptr(z), L.ptr(z), L.ptr(z), L.ptr(z), L.ptr(z),
                                      L.ptr(z), 0, L.ptr(z), L.ptr(z), L.ptr(z), C.byref(f))


def test_element_repeating_a_node_is_rejected():
    """pfem_create refuses an element whose 8 node ids are not distinct (the scatter gives
    every node pair of an element its own K slot and first-touch bit); the check runs
    before any device call."""
    lib = L.lib()
    nodes = np.zeros(3 * 9)
    conn = np.array([0, 1, 2, 3, 4, 5, 6, 6], dtype=np.int64)
    mat = L.Material(**VALID)
    md = L.MeshDesc(nnodes=9, nelem=1, nodes=L.ptr(nodes), conn=L.ptr(conn), material_id=None, nmaterials=1,
                    materials=C.cast(C.pointer(mat), C.c_void_p), element_kind=L.HEX8_FINITE_FBAR)
    out = C.c_void_p()
    assert lib.pfem_create(C.byref(md), None, C.byref(out)) == L.PFEM_ERR_BAD_ARG
    assert "repeats node 6" in lib.pfem_status_string(L.PFEM_ERR_BAD_ARG).decode()
    conn[7] = 7
    assert lib.pfem_create(C.byref(md), None, C.byref(out)) in (L.PFEM_ERR_NO_DEVICE, L.PFEM_OK)
    if out.value:
        lib.pfem_destroy(out)


def test_named_options_and_their_defaults():
    o = pa.default_options()
    assert o.amg_rebuild == L.AMG_KEEP_P_UNTIL_FALLBACK and o.amg_wcycle == 0 and o.krylov_fallback == 0
    assert o.line_search == 0 and o.newton_diverge == 0.0 and o.fp32_krylov == 0 and o.host_build == 0
    assert o.assembly_chunk == 0 and o.smoother_unfused == 0 and list(o.reserved) == [0] * 3 and o.coarse_degree == 0
    assert o.krylov_sync == 0 and o.split_operator == 1
    o = pa.default_options(amg_rebuild="per_load_step", wcycle=True, gcr_fallback=False, rho_iters=(30, 5),
                           stream_prolongation=False, line_search=3, newton_diverge=10.0)
    assert (o.amg_rebuild, o.amg_wcycle, o.krylov_fallback, o.rho_iters_cold, o.rho_iters_warm) == (0, 1, 1, 30, 5)
    assert o.prolong_rowwise == 1 and o.line_search == 3 and o.newton_diverge == 10.0
    with pytest.raises(TypeError):
        pa.default_options(no_such_option=1)
