"""Host-side mirror of the parts of ion_gym.physics the flight path reads:
constants, gas scalars, SDS transport numbers, the r-z field build, the
per-ion envelope and the ensemble driver, plus the engine-backed fly
functions that replace the numba kernels."""
