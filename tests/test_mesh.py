"""Host-side mesh generation and colouring: bit-identical to the oracle (integer/byte
work), jitter invariants, parity colours, notch and inclusion generators."""
import numpy as np
import pytest

import oracle as O
import plasticity_amd as pa


@pytest.mark.parametrize("dims,jit,seed", [((8, 8, 8), 0.0, 0), ((6, 5, 4), 0.1, 3), ((10, 10, 10), 0.1, 0)])
def test_box_mesh_bit_identical_to_oracle(dims, jit, seed):
    m = pa.box_mesh(1, 1, 1, *dims, jitter=jit, seed=seed)
    n2, c2 = O.box_mesh(1, 1, 1, *dims, jitter=jit, seed=seed)
    assert np.array_equal(m.nodes, n2) and np.array_equal(m.elements, c2)


def test_jitter_invariants():
    n = 10
    m0 = pa.box_mesh(1, 1, 1, n, n, n)
    m = pa.box_mesh(1, 1, 1, n, n, n, jitter=0.1, seed=0)
    d = m.nodes - m0.nodes
    assert np.abs(d).max() <= 0.1 / n + 1e-15
    bnd = np.any((m0.nodes == 0) | (m0.nodes == 1), axis=1)
    assert np.all(d[bnd] == 0) and np.all(np.abs(d[~bnd]).max(axis=1) > 0)
    # reference numbering (Mesh.jl:30-59): x fastest
    assert np.array_equal(m0.nodes[1], [0.1, 0, 0]) and np.array_equal(m0.elements[0], [0, 1, 12, 11, 121, 122, 133, 132])
    u = pa.jitter_u01(0, np.arange(1000), 0)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.05


def test_product_colours_equal_oracle_and_parity():
    m = pa.box_mesh(1, 1, 1, 5, 7, 3, jitter=0.1, seed=2)
    col, nc = pa.greedy_colors(m)
    col2, nc2 = O.greedy_colors(m.nnodes, m.elements)
    assert nc == nc2 == 8 and np.array_equal(col, col2)


def test_notched_bar_section_reduction():
    m = pa.notched_bar_mesh(20, length=1.0, notch_depth=0.2)
    X = m.nodes
    mid = np.isclose(X[:, 0], 0.5)
    w = X[mid, 1].max() - X[mid, 1].min()
    assert w ** 2 == pytest.approx(0.8, rel=1e-12)  # 20% area reduction at mid-length
    ends = np.isclose(X[:, 0], 0.0)
    assert X[ends, 1].max() - X[ends, 1].min() == pytest.approx(1.0)


def test_inclusions_volume_fraction():
    m = pa.box_mesh(1, 1, 1, 12, 12, 12)
    mid = pa.ellipsoid_inclusions(m, 0.2, seed=1)
    assert 0.2 <= mid.mean() < 0.35
    assert np.array_equal(mid, pa.ellipsoid_inclusions(m, 0.2, seed=1))


def test_ellipsoid_inclusions_equal_brute_force():
    """The index-box search returns exactly the element set of testing every centroid."""
    from plasticity_amd.mesh import ellipsoid_inclusions, jitter_u01

    def brute(mesh, vf, seed, semi_axes=(0.08, 0.05, 0.05)):
        cent = mesh.nodes[mesh.elements].mean(axis=1)
        lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
        mid = np.zeros(mesh.nelem, dtype=np.int32)
        a = np.asarray(semi_axes) * (hi - lo)
        k = 0
        while mid.sum() < vf * mesh.nelem:
            u = jitter_u01(seed, np.arange(6 * k, 6 * k + 6, dtype=np.uint64) + np.uint64(1 << 40), 0)
            c = lo + u[:3] * (hi - lo)
            th, ph, ps = u[3] * 2 * np.pi, u[4] * np.pi, u[5] * 2 * np.pi
            Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
            Ry = np.array([[np.cos(ph), 0, np.sin(ph)], [0, 1, 0], [-np.sin(ph), 0, np.cos(ph)]])
            Rx = np.array([[1, 0, 0], [0, np.cos(ps), -np.sin(ps)], [0, np.sin(ps), np.cos(ps)]])
            d = (cent - c) @ (Rz @ Ry @ Rx)
            mid[((d / a) ** 2).sum(axis=1) <= 1.0] = 1
            k += 1
        return mid

    for n, seed, vf in [(24, 1, 0.2), (17, 5, 0.3), (30, 2, 0.1)]:
        m = pa.box_mesh(1, 1, 1, n, n, n, jitter=0.1, seed=seed)
        assert np.array_equal(ellipsoid_inclusions(m, vf, seed), brute(m, vf, seed))
