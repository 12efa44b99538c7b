"""rkhs_ba_amd -- MI355X (gfx950) CVO / RKHS-BA registration.

The compute path is ``librkhs_ba_amd.so`` (hand-written HIP kernels behind the C
ABI in include/cvo_c.h); this package is the Python host mirror of the
reference ``cvo::CvoGPU`` interface plus seeded synthetic scenes.
"""
from ._abi import CvoParams, PARAM_NAMES  # noqa: F401
from .runtime import (Association, CvoError, CvoFrame, CvoGPU, CvoResultInfo, DeviceArray, DeviceCloud,  # noqa: F401
                      default_params, icosahedral_rotations, lib, read_params_yaml, version)
from . import synth  # noqa: F401

__all__ = ["CvoGPU", "CvoFrame", "CvoParams", "CvoError", "CvoResultInfo", "Association", "DeviceArray", "DeviceCloud",
           "default_params", "read_params_yaml", "synth", "lib", "version", "icosahedral_rotations"]
