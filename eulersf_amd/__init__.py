"""eulersf_amd -- MI355X (gfx950) drop-in for the stress, coupled-solve and
advection stages of EULERSF's coupled solid-fluid substep.

The compute path is libesf.so (hand-written HIP kernels behind the C ABI of
include/esf.h); this package is the Python view of that ABI plus the cfg
format.  Importing it does not load the library; the first call does, and it
fails loudly when the library has not been built.
"""
from ._abi import (ESF_MIXED, ESF_PLAIN_FP64, EsfError, exported_symbols, load)  # noqa: F401
from .cfg import Cfg, params_from_cfg, unit_cube_ball_cfg  # noqa: F401
from .integrator import CoupledIntegrator3D, corotation_batch, step_batch  # noqa: F401
