"""PETITE shower transport on AMD MI355X (gfx950).

Drop-in classes with the reference's interface (kjkellyphys/PETITE):
``Shower``, ``DarkShower``, ``Particle`` and the module registries, backed by hand-written HIP
kernels in libpetite_hip.so (include/petite_hip.h).
"""
from .constants import (dimensionalities, mass_dict, process_code, process_PIDS,  # noqa: F401
                        target_information)
from .particle import Particle, default_ids  # noqa: F401
from .shower import Shower  # noqa: F401
from .batch import ParticleBatch, PRIMARY_DTYPE  # noqa: F401
from .dark_shower import DarkBatch, DarkShower  # noqa: F401
from .detector import detector_cut, transverse_position  # noqa: F401

__all__ = ["Shower", "DarkShower", "DarkBatch", "Particle", "ParticleBatch", "PRIMARY_DTYPE", "process_code", "process_PIDS",
           "dimensionalities", "mass_dict", "target_information", "default_ids", "detector_cut",
           "transverse_position"]
