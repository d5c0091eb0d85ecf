"""ion_gym_amd -- MI355X-native ion-flight engine for ion_gym.

The batched replacement of ion_gym's numba flight kernels
(ion_gym/physics/tracer_rz.py, tracer3d.py) as hand-written gfx950 HIP
kernels behind a C ABI (include/ion_gym_amd.h), with the reference's
Python flight API (build_run-style fly functions and the ensemble driver)
kept on top.  Importing the package does not load the library; the first
engine call does, and fails loudly when it is not built.
"""
__version__ = "0.1.0"
