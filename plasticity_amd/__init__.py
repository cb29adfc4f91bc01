"""plasticity_amd -- MI355X-native Newton-step hot path of the lagrangian/ finite-strain
J2 plasticity solver (vinaysrao1/plasticity): Hex8 F-bar element integration, colour-
ordered 3x3 BSR assembly, Dirichlet imposition and PCG + smoothed-aggregation AMG, as
hand-written gfx950 HIP kernels behind the C-ABI of include/pfem.h (libpfem.so).
"""
from .mesh import Mesh, box_mesh, dof, ellipsoid_inclusions, greedy_colors, jitter_u01, notched_bar_mesh, \
    on_face, select_nodes  # noqa: F401
from .model import ElementInversionError, J2Material, Model, SolveResult, default_options, \
    element_batch  # noqa: F401
from .vtu import gauss_strain, von_mises, write_vtu  # noqa: F401

__all__ = ["Mesh", "box_mesh", "dof", "on_face", "select_nodes", "jitter_u01", "greedy_colors",
           "notched_bar_mesh", "ellipsoid_inclusions", "J2Material", "Model", "SolveResult",
           "ElementInversionError", "element_batch", "default_options", "write_vtu", "gauss_strain",
           "von_mises"]
