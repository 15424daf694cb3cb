"""capytaine_amd: the Capytaine boundary-element hot path on AMD MI355X (gfx950).

Delhommeau Green-function evaluation, assembly of the influence matrices S
and K, and their dense LU solve run as hand-written HIP kernels in
libcapytaine_amd.so, reached through a C-ABI (include/capytaine_amd.h) with
ctypes.  The Python API mirrors capytaine's: ``Delhommeau``,
``DefaultMatrixEngine`` (alias ``BasicMatrixEngine``), ``BEMSolver``,
``RadiationProblem``, ``DiffractionProblem``, ``FloatingBody`` and mesh
generators.
"""
from .meshes import (Mesh, mesh_sphere, mesh_parallelepiped, mesh_rectangle, mesh_ellipsoid,
                     mesh_vertical_cylinder, ReflectionSymmetricMesh, RotationSymmetricMesh)
from .green_functions import Delhommeau, XieDelhommeau, FinGreen3D, HAMS_GF, GreenFunctionEvaluationError
from .engines import DefaultMatrixEngine, BasicMatrixEngine, DeviceMatrix, LUDecomposedDeviceMatrix
from .problems import (FloatingBody, RadiationProblem, DiffractionProblem, rigid_body_dofs,
                       TranslationDof, RotationDof, froude_krylov_force, airy_waves_velocity,
                       airy_waves_potential, airy_waves_pressure, airy_waves_free_surface_elevation)
from .solver import BEMSolver, problems_for_body
from ._lib import CptError

__version__ = "0.1.0"
