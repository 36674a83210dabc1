"""MI355X-native RBRP column interpolative decomposition.

The computation lives in the HIP library ``librbrp.so`` (C ABI in
``include/rbrp.h``); this package is its host-side mirror.  Importing the
package does not load the library; the first call does, and raises
ImportError if it has not been built.
"""

from .rbrp import (CPQR, RBRP, Batch, ColID, Handle, RandAdapLUPP, RandCPQR,  # noqa: F401
                   RandCPQR_OS, RandLUPP, RandLUPP_OS, batched_launch_shape, batched_tune, colmajor, gaussian_sketch, interp_os,
                   rbrp_colid, rbrp_colid_batched, uniforms)

__all__ = ["CPQR", "RBRP", "Batch", "ColID", "Handle", "RandAdapLUPP", "RandCPQR", "RandCPQR_OS", "RandLUPP",
           "RandLUPP_OS", "batched_launch_shape", "batched_tune", "colmajor", "gaussian_sketch", "interp_os", "rbrp_colid",
           "rbrp_colid_batched", "uniforms"]
