"""VTU output (plasticity_amd/vtu.py, after Visualization.jl:129-203): the small-strain
recovery is exact for affine displacement fields on distorted Hex8 meshes, von Mises has
its closed forms, and the written file parses with the expected arrays.  CPU only (the
model is a stand-in carrying the same accessors as plasticity_amd.Model)."""
import xml.etree.ElementTree as ET

import numpy as np

from plasticity_amd.mesh import box_mesh
from plasticity_amd.vtu import gauss_strain, von_mises, write_vtu


def test_strain_exact_for_affine_fields_on_jittered_mesh():
    mesh = box_mesh(1.0, 2.0, 1.5, 3, 4, 2, jitter=0.2, seed=9)
    A = np.array([[1e-3, 2e-3, -1e-3], [4e-4, -2e-3, 5e-4], [3e-3, 1e-4, 7e-4]])
    U = (mesh.nodes @ A.T).ravel()
    eps = gauss_strain(mesh.nodes, mesh.elements, U)
    S = 0.5 * (A + A.T)
    expect = np.array([S[0, 0], S[1, 1], S[2, 2], 2 * S[0, 1], 2 * S[1, 2], 2 * S[2, 0]])
    assert eps.shape == (8 * mesh.nelem, 6)
    np.testing.assert_allclose(eps, np.tile(expect, (eps.shape[0], 1)), rtol=0, atol=1e-15)


def test_von_mises_closed_forms():
    assert np.isclose(von_mises(np.array([3.0, 0, 0, 0, 0, 0])), 3.0)
    assert np.isclose(von_mises(np.array([0, 0, 0, 2.0, 0, 0])), 2.0 * np.sqrt(3.0))
    assert np.isclose(von_mises(np.array([5.0, 5.0, 5.0, 0, 0, 0])), 0.0)


class _StubModel:
    def __init__(self, mesh, rng):
        self.mesh = mesh
        self.U = rng.normal(size=3 * mesh.nnodes) * 1e-3
        self._sig = rng.normal(size=(8 * mesh.nelem, 6))
        self._abar = rng.random(8 * mesh.nelem)

    def gauss_stress(self):
        return self._sig

    def equivalent_plastic_strain(self):
        return self._abar


def _arrays(path):
    root = ET.parse(path).getroot()
    out = {}
    for da in root.iter("DataArray"):
        vals = np.array(da.text.split(), dtype=float)
        out[da.get("Name", "Points")] = (vals, da)
    return root, out


def test_written_file_parses_with_expected_fields(tmp_path):
    rng = np.random.default_rng(4)
    mesh = box_mesh(1.0, 1.0, 1.0, 2, 3, 2, jitter=0.1, seed=2)
    m = _StubModel(mesh, rng)
    fn = write_vtu(str(tmp_path / "block"), m)
    assert fn.endswith(".vtu")
    root, arr = _arrays(fn)
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == mesh.nnodes and int(piece.get("NumberOfCells")) == mesh.nelem
    np.testing.assert_allclose(arr["Points"][0].reshape(-1, 3), mesh.nodes, rtol=1e-15)
    np.testing.assert_array_equal(arr["connectivity"][0].reshape(-1, 8), mesh.elements)
    np.testing.assert_array_equal(arr["offsets"][0], 8 * np.arange(1, mesh.nelem + 1))
    assert set(arr["types"][0]) == {12}
    np.testing.assert_allclose(arr["Displacement"][0], m.U, rtol=1e-15)
    sig = m._sig.reshape(mesh.nelem, 8, 6).mean(axis=1)
    np.testing.assert_allclose(arr["Stress"][0].reshape(-1, 6), sig, rtol=1e-14)
    np.testing.assert_allclose(arr["VonMises"][0], von_mises(sig), rtol=1e-14)
    np.testing.assert_allclose(arr["MeanStress"][0], sig[:, :3].mean(axis=1), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(arr["EqPlasticStrain"][0], m._abar.reshape(-1, 8).mean(axis=1), rtol=1e-14)
    assert arr["Stress"][1].get("ComponentName3") == "xy"
