"""ffip_amd: MI355X (gfx950) FDTD time-stepping engine behind ffip's Python API.

The public names mirror the reference package's ``ffip/__init__.py`` (geometry
and media, sources, Simulation and its helpers, and the native ``transpose``),
so scripts written for the reference can switch with ``import ffip_amd as ffip``.
The hot path runs in libffip_hip.so (hand-written HIP kernels); there is no CPU
fallback and the library must be built (``python -m ffip_amd.build``).
"""
from ._lib import FfipError, LibSim, TO_convolve, TO_convolve_transpose, transpose  # noqa: F401
from .polygon import check_inside  # noqa: F401
from .geom import (Box, DeybeSusceptibility, DrudeSusceptibility, General_Medium_Box,  # noqa: F401
                   GeometricObject, LorentzianSusceptibility, Medium, Sphere, Susceptibility, Two_Medium_Box,
                   Vector3, cmp_shape, getgrid, metadata_to_pts, planar_polygon, sph2cart)
from .simulation import (PML, Adjoint_Flux, Adjoint_Source, Adjoint_Volume, Fields_DFT, Flux_Box,  # noqa: F401
                         Flux_Region, Geom_Info, Simulation, Symmetry, direction_dict, field_match_objective,
                         run_until_dft, run_until_fields_decay, run_until_time, side_dict)
from .source import Gaussian1, Inhom_Source, Source, Source_Function  # noqa: F401
