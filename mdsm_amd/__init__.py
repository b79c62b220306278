"""mdsm_amd -- MI355X-native brute-force dedispersion and single-pulse
detection (the MDSM dedisperser/detector stage) behind a C ABI.

The product is ``libmdsm_amd.so`` (HIP kernels for gfx950 + C++ host code,
``include/mdsm_abi.h``); ``mdsm_amd.abi`` is its ctypes binding.
"""
from . import abi  # noqa: F401
