"""Integer enums of the mesh-BEM interface.

The values are part of the drop-in interface: callers pass them as plain ints
(``BEM_type=int(BEMType.HELMHOLTZ_MOM)``), exactly as with the reference
(src/utils/types.py:81-154).
"""

from enum import Enum


class _IntEnum(Enum):
    def __int__(self):
        return self.value


class BoundaryType(_IntEnum):
    """Per-vertex boundary flag (src/utils/types.py:81-86)."""

    DIRICHLET = 0
    NEUMANN = 1


class LayerType(_IntEnum):
    """Boundary-integral operator kind (src/utils/types.py:89-96)."""

    RHS = 0
    SINGLE_LAYER = 1
    DOUBLE_LAYER = 2
    HYPERSINGULAR_LAYER = 3


class BEMType(_IntEnum):
    """PDE / formulation selector (src/utils/types.py:108-136)."""

    LAPLACE = 0
    HELMHOLTZ_BEM = 1
    HELMHOLTZ_MOM = 2

    @staticmethod
    def use_complex(BEM_type: int) -> bool:
        return BEM_type in (int(BEMType.HELMHOLTZ_BEM), int(BEMType.HELMHOLTZ_MOM))

    @staticmethod
    def belong_Helmholtz(BEM_type: int) -> bool:
        return BEM_type in (int(BEMType.HELMHOLTZ_BEM), int(BEMType.HELMHOLTZ_MOM))


class OrderType(_IntEnum):
    """Shape-function order on a triangle (src/utils/types.py:139-144)."""

    PLANAR = 0
    LINEAR = 1


class PanelRelationType(_IntEnum):
    """Number of vertex indices two triangles share (src/utils/types.py:147-154)."""

    SEPARATE = 0
    SAME_VERTEX = 1
    SAME_EDGE = 2
    SAME_FACE = 3


class AssemblyMode(_IntEnum):
    """Which integration rule the MoM / Laplace assembly follows.

    REF reproduces the reference kernel: collapsed Gauss-Legendre rules of
    order Q (test) and Q-1 (source), CFIE terms, the fixed incidence of
    src/mesh_BEM/cuda/mesh_MOM_3d.cu:24-70.
    NS is the EFIE with the 7-point degree-5 triangle rule, analytic FP64
    1/R terms on near triangle pairs and parameterised plane waves.
    """

    REF = 0
    NS = 1


class Precision(_IntEnum):
    F32 = 0
    F64 = 1
