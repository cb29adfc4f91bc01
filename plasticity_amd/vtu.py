"""Result output: ASCII VTK unstructured-grid files (.vtu) for ParaView / VisIt.

Same fields as the reference's `write_vtu` (lagrangian/src/Visualization.jl:129-203):
point data `Displacement`; cell data (Gauss-point values averaged per element)
`Stress` (Cauchy = Kirchhoff / det F of the un-barred F for the finite-strain
elements, Model.jl:188-208), `Strain` (total small strain B u, engineering shear,
Visualization.jl:41-58), `VonMises`, `MeanStress` and `EqPlasticStrain`.  Hex8 node
order is VTK's (cell type 12).
"""
from __future__ import annotations

import numpy as np

VOIGT_NAMES = ("xx", "yy", "zz", "xy", "yz", "zx")
# Gauss points (Elements.jl:34-42 order) and the Hex8 natural node coordinates (:26-30)
_NAT = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                 [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
_GQ = 1.0 / np.sqrt(3.0)
_GAUSS = _NAT * _GQ


def _dshape(xi):
    """dN_a/dxi_j at natural point xi (8 x 3)."""
    s = _NAT
    d = np.empty((8, 3))
    d[:, 0] = s[:, 0] * (1 + s[:, 1] * xi[1]) * (1 + s[:, 2] * xi[2]) / 8
    d[:, 1] = s[:, 1] * (1 + s[:, 0] * xi[0]) * (1 + s[:, 2] * xi[2]) / 8
    d[:, 2] = s[:, 2] * (1 + s[:, 0] * xi[0]) * (1 + s[:, 1] * xi[1]) / 8
    return d


def gauss_strain(nodes, elements, U):
    """Total small strain at every Gauss point (ngp x 6, engineering shear, gp = 8 e + g)
    from the reference configuration: eps = sym(grad u) with grad u = ue^T dN/dX."""
    X = np.asarray(nodes, dtype=np.float64)[elements]           # (ne, 8, 3)
    ue = np.asarray(U, dtype=np.float64).reshape(-1, 3)[elements]  # (ne, 8, 3)
    out = np.empty((elements.shape[0], 8, 6))
    for g in range(8):
        dN = _dshape(_GAUSS[g])                                   # (8, 3)
        Jm = np.einsum("eai,aj->eij", X, dN)                      # dX_i/dxi_j
        dNdX = np.einsum("aj,eji->eai", dN, np.linalg.inv(Jm))   # dN_a/dX_i
        H = np.einsum("eak,eai->eki", ue, dNdX)                  # du_k/dX_i
        out[:, g, 0] = H[:, 0, 0]
        out[:, g, 1] = H[:, 1, 1]
        out[:, g, 2] = H[:, 2, 2]
        out[:, g, 3] = H[:, 0, 1] + H[:, 1, 0]
        out[:, g, 4] = H[:, 1, 2] + H[:, 2, 1]
        out[:, g, 5] = H[:, 2, 0] + H[:, 0, 2]
    return out.reshape(-1, 6)


def von_mises(s):
    s = np.asarray(s)
    return np.sqrt(0.5 * ((s[..., 0] - s[..., 1]) ** 2 + (s[..., 1] - s[..., 2]) ** 2 + (s[..., 2] - s[..., 0]) ** 2)
                   + 3.0 * (s[..., 3] ** 2 + s[..., 4] ** 2 + s[..., 5] ** 2))


def _cell_avg(field, nelem):
    f = np.asarray(field).reshape(nelem, 8, -1)
    return f.mean(axis=1)


def _array(name, data, ncomp, dtype="Float64", comps=None):
    extra = " ".join(f'ComponentName{i}="{c}"' for i, c in enumerate(comps)) + " " if comps else ""
    head = f'        <DataArray type="{dtype}" Name="{name}" NumberOfComponents="{ncomp}" {extra}format="ascii">\n'
    fmt = "%d" if dtype.startswith(("Int", "UInt")) else "%.17g"
    body = "\n".join("          " + " ".join(fmt % v for v in row) for row in np.atleast_2d(data).reshape(-1, ncomp))
    return head + body + "\n        </DataArray>\n"


def write_vtu(filename, model):
    """Write `model`'s committed state as filename(.vtu); returns the path written.
    Call after a converged solve: Stress and EqPlasticStrain are the committed values,
    Strain is recomputed from the current displacement."""
    fn = filename if str(filename).endswith(".vtu") else f"{filename}.vtu"
    mesh = model.mesh
    nn, ne = mesh.nnodes, mesh.nelem
    U = model.U
    sig = _cell_avg(model.gauss_stress(), ne)
    eps = _cell_avg(gauss_strain(mesh.nodes, mesh.elements, U), ne)
    abar = _cell_avg(model.equivalent_plastic_strain(), ne)[:, 0]
    with open(fn, "w") as f:
        f.write('<?xml version="1.0"?>\n')
        f.write('<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian">\n')
        f.write("  <UnstructuredGrid>\n")
        f.write(f'    <Piece NumberOfPoints="{nn}" NumberOfCells="{ne}">\n')
        f.write("      <Points>\n")
        f.write(_array("Points", mesh.nodes, 3).replace(' Name="Points"', ""))
        f.write("      </Points>\n      <Cells>\n")
        f.write(_array("connectivity", mesh.elements, 8, "Int64").replace(' NumberOfComponents="8"', ""))
        f.write(_array("offsets", 8 * np.arange(1, ne + 1), 1, "Int64").replace(' NumberOfComponents="1"', ""))
        f.write(_array("types", np.full(ne, 12), 1, "UInt8").replace(' NumberOfComponents="1"', ""))
        f.write("      </Cells>\n")
        f.write('      <PointData Vectors="Displacement">\n')
        f.write(_array("Displacement", U.reshape(-1, 3), 3))
        f.write("      </PointData>\n")
        f.write('      <CellData Scalars="VonMises" Tensors="Stress">\n')
        f.write(_array("Stress", sig, 6, comps=VOIGT_NAMES))
        f.write(_array("Strain", eps, 6, comps=VOIGT_NAMES))
        f.write(_array("VonMises", von_mises(sig), 1))
        f.write(_array("MeanStress", sig[:, :3].mean(axis=1), 1))
        f.write(_array("EqPlasticStrain", abar, 1))
        f.write("      </CellData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n")
    return fn
