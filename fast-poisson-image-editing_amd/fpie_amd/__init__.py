"""MI355X-native Poisson image editing: the ``hip`` solver backend.

``fpie_amd.process`` mirrors the reference's processor API
(EquProcessor/GridProcessor, fpie/process.py); ``fpie_amd.core_hip`` holds
the solver cores with the reference core interface, backed by the gfx950
HIP library ``libfpie_hip.so`` through the C ABI in include/fpie_hip.h.
"""

__version__ = "0.1.0"
