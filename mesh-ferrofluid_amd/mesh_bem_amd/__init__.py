"""MI355X-native mesh-BEM hot path (RWG MoM and Laplace BEM).

Python host code mirrors the reference's interfaces (src/mesh_BEM,
src/utils) and calls hand-written HIP kernels through the C ABI of
include/mesh_bem_amd.h.
"""

from .types import (
    AssemblyMode,
    BEMType,
    BoundaryType,
    LayerType,
    OrderType,
    PanelRelationType,
    Precision,
)
from .quadrature import get_gaussion_integration_points_and_weights, radon7
from .geometry import (
    compute_face_areas,
    compute_face_normals,
    compute_panel_relation,
    compute_vert_areas,
    compute_vert_normals,
    get_edge_unique,
    icosphere,
    geodesic_sphere,
)
from .kernels import MeshMOMInterpolateJsKernel, MeshMOMRhsKernel, MeshMOMZmatKernel
