"""Model set-up on the device (build.hip): the box mesh generator is bit-identical to the
host and oracle generators; the device-built pattern, scatter slots and colours equal the
host-built ones (pfem_options.host_build = 1), so K and R assembled through either are
bit-identical -- for box numbering (device parity colours) and for a permuted element
order (device pattern, host greedy colours)."""
import numpy as np
import pytest

import oracle as O
import plasticity_amd as pa
from plasticity_amd.mesh import Mesh

pytestmark = pytest.mark.gpu
MAT = dict(E=200e9, nu=0.3, sy0=250e6, Hiso=1e9)


@pytest.mark.parametrize("dims,jit,seed", [((13, 11, 9), 0.1, 3), ((40, 40, 40), 0.1, 0), ((5, 7, 3), 0.0, 0),
                                           ((1, 1, 1), 0.1, 2)])
def test_device_box_mesh_is_bit_identical(dims, jit, seed, need_gpu):
    print(f"[box mesh] {dims} jitter {jit} seed {seed}", flush=True)
    h = pa.box_mesh(1.0, 2.0, 0.5, *dims, jitter=jit, seed=seed)
    d = pa.box_mesh(1.0, 2.0, 0.5, *dims, jitter=jit, seed=seed, device=True)
    assert np.array_equal(h.nodes, d.nodes) and np.array_equal(h.elements, d.elements)
    on, oc = O.box_mesh(1.0, 2.0, 0.5, *dims, jitter=jit, seed=seed)
    assert np.array_equal(on, d.nodes) and np.array_equal(oc, d.elements)


def _pair(mesh, element="finite_fbar"):
    oh = pa.default_options()
    oh.host_build = 1
    md = pa.Model(mesh, pa.J2Material(**MAT), element=element)
    mh = pa.Model(mesh, pa.J2Material(**MAT), element=element, options=oh)
    return md, mh


def _compare(md, mh, mesh, seed):
    rp_d, ci_d, _ = md.get_K_bsr()
    rp_h, ci_h, _ = mh.get_K_bsr()
    assert np.array_equal(rp_d, rp_h) and np.array_equal(ci_d, ci_h)
    cd, nd = md.colors()
    ch, nh = mh.colors()
    assert nd == nh and np.array_equal(cd, ch)
    rng = np.random.default_rng(seed)
    F = np.eye(3) + 0.01 * rng.normal(size=(3, 3))
    U = (mesh.nodes @ F.T - mesh.nodes).ravel() + 1e-4 * rng.normal(size=3 * mesh.nnodes)
    Rd = md.assemble(U)
    Rh = mh.assemble(U)
    _, _, vd = md.get_K_bsr()
    _, _, vh = mh.get_K_bsr()
    assert np.array_equal(Rd, Rh) and np.array_equal(vd, vh)


def test_device_build_equals_host_build_on_a_box(need_gpu):
    mesh = pa.box_mesh(1, 1, 1, 9, 7, 6, jitter=0.1, seed=5)
    print(f"[device build] box {mesh.nelem} elements", flush=True)
    md, mh = _pair(mesh)
    sd, sh = md.stats(), mh.stats()
    print(f"  built on device: {sd.built_on_device} (host-built: {sh.built_on_device}); "
          f"pattern time {sd.t_create_pattern * 1e3:.1f} ms vs {sh.t_create_pattern * 1e3:.1f} ms", flush=True)
    assert sd.built_on_device == 3 and sh.built_on_device == 0
    assert sd.nnzb == sh.nnzb and sd.ncolors == sh.ncolors == 8
    _compare(md, mh, mesh, 1)


def test_device_build_with_permuted_elements_uses_host_colours(need_gpu):
    base = pa.box_mesh(1, 1, 1, 6, 5, 4, jitter=0.1, seed=6)
    perm = np.random.default_rng(7).permutation(base.nelem)
    mesh = Mesh(base.nodes, base.elements[perm])
    print(f"[device build] permuted {mesh.nelem} elements", flush=True)
    md, mh = _pair(mesh)
    assert md.stats().built_on_device == 1 and mh.stats().built_on_device == 0
    _compare(md, mh, mesh, 2)
    # the colouring is valid: no two elements of one colour share a node
    col, nc = md.colors()
    for c in range(nc):
        nodes = mesh.elements[col == c].ravel()
        assert np.unique(nodes).size == nodes.size


def test_device_build_small_strain_and_notched_bar(need_gpu):
    from plasticity_amd.mesh import notched_bar_mesh
    mesh = notched_bar_mesh(8, jitter=0.05, seed=7)
    print(f"[device build] notched bar {mesh.nelem} elements, small strain", flush=True)
    md, mh = _pair(mesh, element="small")
    assert md.stats().built_on_device == 3
    _compare(md, mh, mesh, 3)
