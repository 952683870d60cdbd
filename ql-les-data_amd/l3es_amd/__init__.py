"""l3es_amd: MI355X (gfx950) HIP implementation of the ql-les-data spectral step.

Host-side mirror of the reference package ``l3es`` for the hot path only:
spectral_solver (RHS, RK4, forcing, setup, driver), init_fields, integrator
(particle interpolation), relax (SPH relaxation), turbulence (particle -> grid spectra), combined (DNS + particles),
dataset (packing of the frames and their statistics) and the spectral utilities (filter, spectrum, dissipation,
SPH density).  All arithmetic runs in
``ql-les-data_amd/lib/libqlles.so`` through the C ABI in ``include/qlles.h``.
"""

from . import _lib  # noqa: F401
from .spectral_solver import (comp_dt, custom_jit_integrate, rhs_wrapper, rk4_wrapper,  # noqa: F401
                              set_up_solver, simulate)
from .integrator import shift_fn, spectral_interpolator_wrapper  # noqa: F401
from .batch import Batch  # noqa: F401
from .relax import relax_advect, relax_wrapper  # noqa: F401
from .utils import rho_computer  # noqa: F401
from .turbulence import u_and_spectrum_from_ur, ur_to_u_dft_wrapper, ur_to_u_mls_wrapper  # noqa: F401
from .dataset import compute_statistics, single_frames_to_dataset  # noqa: F401
