"""photonprop — host side of the MI355X photon-propagation library.

Python mirror of PhotonProp.jl's modules (src/photon_prop.jl:1-22): Utils,
Medium, Spectral (spectrum), Emission, Detection and the propagation entry
points.  Compute runs in libphotonprop.so (hand-written gfx950 HIP kernels);
this package only converts inputs and outputs.
"""
from . import _abi, detection, emission, lightyield, medium, modelling, spectrum, utils  # noqa: F401
from ._abi import HIT_DTYPE, PhotonPropError, load_library  # noqa: F401
from .detection import DetectionSphere  # noqa: F401
from .emission import AngularEmissionProfile, PhotonSource  # noqa: F401
from .medium import (MediumTables, SeawaterConfigMedium, WaterProperties,  # noqa: F401
                     get_absorption_length, get_refractive_index, get_scattering_length,
                     make_cascadia_medium_properties, medium_tables)
from .propagation import (Context, REQUEST_DTYPE, make_requests, propagate_distance,  # noqa: F401
                          request_from, result_table)
from .serving import Server, run_open_loop, run_open_loop_native, run_per_request_native  # noqa: F401
from .spectrum import (CherenkovSpectrum, InverseSquareSpectrum, Monochromatic,  # noqa: F401
                       frank_tamm, frank_tamm_inverted_cdf, frank_tamm_norm)
