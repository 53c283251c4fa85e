"""MI355X-native (gfx950) filterbank dedispersion hot path with sigpyproc's API.

Hot path: sub-byte unpack -> incoherent dedispersion / subbanding -> DM-trial sweep ->
boxcar matched filter, as hand-written HIP kernels in ``libspp_hip.so`` (C ABI, see
``include/spp_hip.h``), behind drop-ins for sigpyproc's ``FilReader``, ``Filterbank``,
``TimeSeries``, ``MatchedFilter`` and the ``core.kernels`` operators.
"""

from sigpyproc3_amd.header import Header
from sigpyproc3_amd.readers import FilReader
from sigpyproc3_amd.timeseries import TimeSeries

__version__ = "0.1.0"

__all__ = ["FilReader", "Header", "TimeSeries", "__version__"]
