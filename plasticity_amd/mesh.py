"""Mesh generation and node selection (host side, numpy).

box_mesh        Mesh.jl:30-59 (lexicographic nodes, VTK hex connectivity), plus the
                seeded interior jitter used by the GPU-scale workloads: every interior
                node moves by (2u-1)*frac*h per axis, u = splitmix64-counter uniform in
                [0,1) keyed by (seed, 3*node + component).  Boundary nodes never move,
                and frac = 0 reproduces box_mesh exactly.
on_face         Mesh.jl:82-104 (tolerance 1e-8*max(L,1)).
select_nodes    Mesh.jl:66-74.
dof             Mesh.jl:111 (0-based: 3*node + comp).
greedy_colors   Assembly.jl:172-188 as first-fit colouring in element order.
"""
from __future__ import annotations

import numpy as np

__all__ = ["Mesh", "box_mesh", "jitter_u01", "on_face", "select_nodes", "dof", "greedy_colors",
           "notched_bar_mesh", "ellipsoid_inclusions"]

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def jitter_u01(seed: int, node, comp) -> np.ndarray:
    """Counter-based uniform [0,1): splitmix64(splitmix64(seed) ^ (3*node + comp)) >> 11 * 2^-53."""
    s = _splitmix64(np.array([seed], dtype=np.uint64))[0]
    key = (np.asarray(node, dtype=np.uint64) * np.uint64(3) + np.asarray(comp, dtype=np.uint64))
    h = _splitmix64(s ^ key)
    return (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


class Mesh:
    """nodes (nnodes, 3) float64, elements (nelem, 8) int64, 0-based (Mesh.jl:17-22)."""

    def __init__(self, nodes, elements, dims=None):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.elements = np.ascontiguousarray(elements, dtype=np.int64)
        self.nnodes = self.nodes.shape[0]
        self.nelem = self.elements.shape[0]
        self.dims = dims

    def __repr__(self):
        return f"Mesh(nnodes={self.nnodes}, nelem={self.nelem})"


def box_mesh(lx, ly, lz, nx, ny, nz, jitter=0.0, seed=0, device=False) -> Mesh:
    """device=True generates the same mesh (bit for bit) with pfem_box_mesh on the GPU."""
    if device:
        from . import _lib as L
        nodes = np.empty(((nx + 1) * (ny + 1) * (nz + 1), 3))
        el = np.empty((nx * ny * nz, 8), dtype=np.int64)
        L.check(L.lib().pfem_box_mesh(float(lx), float(ly), float(lz), int(nx), int(ny), int(nz), float(jitter),
                                      int(seed), L.ptr(nodes), L.ptr(el)), "pfem_box_mesh")
        return Mesh(nodes, el, dims=(nx, ny, nz, lx, ly, lz))
    npx, npy, npz = nx + 1, ny + 1, nz + 1
    k, j, i = np.meshgrid(np.arange(npz), np.arange(npy), np.arange(npx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    nodes = np.empty((npx * npy * npz, 3))
    nodes[:, 0] = lx * i / nx
    nodes[:, 1] = ly * j / ny
    nodes[:, 2] = lz * k / nz
    if jitter != 0.0:
        interior = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
        ids = np.nonzero(interior)[0].astype(np.uint64)
        h = (lx / nx, ly / ny, lz / nz)
        for c in range(3):
            u = jitter_u01(seed, ids, c)
            nodes[ids, c] += (2 * u - 1) * jitter * h[c]
    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()

    def nid(a, b, c):
        return (c * npy + b) * npx + a

    el = np.stack([nid(ei, ej, ek), nid(ei + 1, ej, ek), nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek),
                   nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1), nid(ei + 1, ej + 1, ek + 1),
                   nid(ei, ej + 1, ek + 1)], axis=1).astype(np.int64)
    return Mesh(nodes, el, dims=(nx, ny, nz, lx, ly, lz))


def select_nodes(mesh: Mesh, pred) -> np.ndarray:
    x, y, z = mesh.nodes[:, 0], mesh.nodes[:, 1], mesh.nodes[:, 2]
    return np.nonzero(pred(x, y, z))[0].astype(np.int64)


def on_face(mesh: Mesh, face: str) -> np.ndarray:
    lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
    L = float((hi - lo).max())
    tol = 1e-8 * max(L, 1.0)
    if face not in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        raise ValueError(f"unknown face {face}")
    ax = "xyz".index(face[0])
    ref = lo[ax] if face.endswith("min") else hi[ax]
    return np.nonzero(np.abs(mesh.nodes[:, ax] - ref) <= tol)[0].astype(np.int64)


def dof(node, comp):
    """0-based global dof of (node, comp), comp in 0..2 (Mesh.jl:111)."""
    return 3 * np.asarray(node) + np.asarray(comp)


def greedy_colors(mesh: Mesh):
    """First-fit colouring in element order: colour[e], ncolours."""
    mask = np.zeros(mesh.nnodes, dtype=np.uint64)
    col = np.zeros(mesh.nelem, dtype=np.int32)
    nc = 0
    for e in range(mesh.nelem):
        nodes = mesh.elements[e]
        forb = np.bitwise_or.reduce(mask[nodes])
        c = 0
        while (int(forb) >> c) & 1:
            c += 1
        col[e] = c
        nc = max(nc, c + 1)
        mask[nodes] |= np.uint64(1 << c)
    return col, nc


def notched_bar_mesh(n, length=1.0, notch_depth=0.2, jitter=0.0, seed=0) -> Mesh:
    """Square bar [0,L]x[0,1]x[0,1] on an n^3 structured grid whose mid-length
    cross-section is reduced by `notch_depth` (area fraction) with a circular notch:
    the lateral coordinates (y, z) about the bar axis are scaled by
    1 - d*(1 - ((x - L/2)/r)^2) for |x - L/2| < r, r = L/10, with d chosen so the
    mid-section area is (1 - notch_depth) of the full section."""
    m = box_mesh(length, 1.0, 1.0, n, n, n, jitter=jitter, seed=seed)
    X = m.nodes
    r = length / 10.0
    d = 1.0 - np.sqrt(1.0 - notch_depth)
    xi = (X[:, 0] - length / 2) / r
    f = np.where(np.abs(xi) < 1.0, 1.0 - d * (1.0 - xi ** 2), 1.0)
    X[:, 1] = 0.5 + (X[:, 1] - 0.5) * f
    X[:, 2] = 0.5 + (X[:, 2] - 0.5) * f
    return m


def ellipsoid_inclusions(mesh: Mesh, volume_fraction, seed, semi_axes=(0.08, 0.05, 0.05)) -> np.ndarray:
    """Per-element material id (0 matrix, 1 inclusion): ellipsoids with random centres
    and orientations (splitmix64 stream keyed by `seed`) are added until the element
    centroids inside any ellipsoid reach `volume_fraction` of all elements."""
    cent = mesh.nodes[mesh.elements].mean(axis=1)
    lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
    mid = np.zeros(mesh.nelem, dtype=np.int32)
    a = np.asarray(semi_axes, dtype=np.float64) * (hi - lo)
    k = 0
    target = volume_fraction * mesh.nelem
    count = 0
    # Box numbering (dims known): only centroids in the index box around an ellipsoid
    # can lie inside it (|centroid - c| <= max semi-axis; jitter moves a centroid by less
    # than one cell), so each ellipsoid tests that block instead of every element.
    box = mesh.dims is not None and mesh.nelem == mesh.dims[0] * mesh.dims[1] * mesh.dims[2]
    if box:
        nd = np.array(mesh.dims[:3])
        h = (hi - lo) / nd
        cg = cent.reshape(nd[2], nd[1], nd[0], 3)
        ids = np.arange(mesh.nelem).reshape(nd[2], nd[1], nd[0])
    while count < target and k < 100000:
        u = jitter_u01(seed, np.arange(6 * k, 6 * k + 6, dtype=np.uint64) + np.uint64(1 << 40), 0)
        c = lo + u[:3] * (hi - lo)
        th, ph, ps = u[3] * 2 * np.pi, u[4] * np.pi, u[5] * 2 * np.pi
        Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        Ry = np.array([[np.cos(ph), 0, np.sin(ph)], [0, 1, 0], [-np.sin(ph), 0, np.cos(ph)]])
        Rx = np.array([[1, 0, 0], [0, np.cos(ps), -np.sin(ps)], [0, np.sin(ps), np.cos(ps)]])
        R = Rz @ Ry @ Rx
        if box:
            r = a.max()
            i0 = np.maximum(np.floor((c - r - lo) / h).astype(int) - 1, 0)
            i1 = np.minimum(np.ceil((c + r - lo) / h).astype(int) + 1, nd)
            sub = cg[i0[2]:i1[2], i0[1]:i1[1], i0[0]:i1[0]].reshape(-1, 3)
            sid = ids[i0[2]:i1[2], i0[1]:i1[1], i0[0]:i1[0]].ravel()
        else:
            sub, sid = cent, np.arange(mesh.nelem)
        d = (sub - c) @ R
        inside = sid[((d / a) ** 2).sum(axis=1) <= 1.0]
        count += int(np.count_nonzero(mid[inside] == 0))
        mid[inside] = 1
        k += 1
    return mid
