"""MI355X-native implementation of the afft transform hot path.

The product is the C-ABI shared library ``libafft_mi355x.so`` (include/fft.h,
include/afft_gpu.h) built from ``fft_amd/csrc`` by ``fft_amd.build``.  This
package is its Python front end: ``fft_amd.afft`` mirrors afft's C++ wrapper
classes (fft.hpp) and adds the batched ``Plan``.
"""
from ._lib import (afft_auto, afft_gpu, afft_interleaved, afft_split, lib,  # noqa: F401
                   last_error, LIB_PATH)
from . import afft  # noqa: F401
