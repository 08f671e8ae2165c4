"""MI355X-native LV-ELVA time step of FastIsostasy (HIP / CDNA4).

The product is the C ABI in include/fastiso.h, implemented by libfastiso.so
(fastisostasy_amd/csrc, built by fastisostasy_amd.build).  This package holds
the Python mirror of the reference's isos_class / bsl_class API over that ABI.
"""
from .isostasy import (BslClass, IsosClass, Batch, bsl_init, bsl_update, isos_end, isos_init,
                       isos_init_ref, isos_init_state, isos_update, make_params, padded_size)

__all__ = ["BslClass", "IsosClass", "Batch", "bsl_init", "bsl_update", "isos_end", "isos_init",
           "isos_init_ref", "isos_init_state", "isos_update", "make_params", "padded_size"]
