"""salvus_amd: MI355X-native spectral-element time stepping (tensor quads and
hexes; Scalar, Elastic2D, Elastic3D) behind the reference's
Element/Physics/Problem interface, through the C ABI of include/salvus_sem.h.
"""
from . import meshgen  # noqa: F401
from .sem import (ACOUSTIC, ELASTIC, DEFAULT_QUIRKS, PATH_FUSED, PATH_PLAIN,  # noqa: F401
                  QUIRK_E3D_SOURCE, QUIRK_NEWMARK3D, QUIRK_QUAD_JACOBIAN, Batch, Mesh, Order2Newmark, SemError,
                  device_count, reference_step_count, topology)
